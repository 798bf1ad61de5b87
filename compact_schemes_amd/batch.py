"""Batched subdomains of the HIP library (cs_batch_*): derivative along x, y, z
and the compact filter along x, y, z of many independent subdomains in one
launch per phase; see include/compact_schemes.h."""
import ctypes

from . import Dims, _p, _stream, check, lib, tune

OP_DERIV, OP_FILTER = 0, 1


class Batch:
    """dims: list of (lxi, let, lze); nbc: one [[3 codes], [3 codes]] for all
    subdomains or a list of them."""

    def __init__(self, dims, nbc):
        self.dims = [tuple(int(v) for v in d) for d in dims]
        n = len(self.dims)
        if nbc and isinstance(nbc[0][0], int):
            nbc = [nbc] * n
        flat = [int(v) for b in nbc for row in b for v in row]
        arr = (Dims * n)(*[Dims(*d) for d in self.dims])
        h = ctypes.c_void_p()
        check(lib().cs_batch_create(n, arr, (ctypes.c_int * len(flat))(*flat), ctypes.byref(h)))
        self.h = h
        self.count = n
        self._keep = None

    def close(self):
        if getattr(self, "h", None):
            lib().cs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bind(self, f, d):
        """f: list of count device tensors (filtered in place); d: list of
        3*count device tensors, d[3*s + axis] = derivative of f[s] along axis."""
        assert len(f) == self.count and len(d) == 3 * self.count
        fa = (ctypes.c_void_p * self.count)(*[_p(t) for t in f])
        da = (ctypes.c_void_p * (3 * self.count))(*[_p(t) for t in d])
        check(lib().cs_batch_bind(self.h, fa, da))
        self._keep = (f, d)

    def deriv(self, stream=None):
        check(lib().cs_batch_deriv(self.h, _stream(stream)))

    def filter(self, stream=None):
        check(lib().cs_batch_filter(self.h, _stream(stream)))

    def step(self, stream=None):
        check(lib().cs_batch_step(self.h, _stream(stream)))

    def set_persistent(self, mode):
        """Grid mode: 0 one workgroup per item, 1 persistent grid, 2 one launch
        per subdomain and operator (the single-subdomain kernels)."""
        check(lib().cs_batch_set_persistent(self.h, int(mode)))

    def get_persistent(self):
        m = ctypes.c_int()
        check(lib().cs_batch_get_persistent(self.h, ctypes.byref(m)))
        return m.value

    def capture(self, stream):
        check(lib().cs_batch_capture(self.h, _stream(stream)))

    def replay(self, stream=None):
        check(lib().cs_batch_replay(self.h, _stream(stream)))

    def info(self):
        """(work items, kernel launches) of the phases deriv, filter x, y, z."""
        items = (ctypes.c_int * 4)()
        launches = (ctypes.c_int * 4)()
        check(lib().cs_batch_info(self.h, items, launches))
        return list(items), list(launches)

    def autotune(self, reps=3, stream=None):
        check(lib().cs_batch_autotune(self.h, int(reps), _stream(stream)))


def tune_set(op, n, K, start=0, end=0, precision=0):
    """chunk length K of the line kernels (see compact_schemes_amd.tune)"""
    tune.set_line(op, n, K, start, end, precision)


def tune_get(op, n, start=0, end=0, precision=0):
    return tune.get_line(op, n, start, end, precision)


tune_clear = tune.clear
tune_save = tune.save
tune_load = tune.load
