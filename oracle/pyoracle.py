"""ctypes binding of the CPU oracle (oracle/cs_oracle.cpp).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to compare against.  The product path never imports
this module.  All arrays are float64 numpy arrays in the reference layout
(SoA, variable-major, x fastest).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build(force=False):
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        V, P = None, ctypes.c_void_p
        sig = {
            "orc_interface_table": (V, [ctypes.c_int, _D, _D]),
            "orc_build_deriv_rows": (V, [ctypes.c_int, _D, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, _D, _D]),
            "orc_solve_tridiag": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, _D]),
            "orc_filter_coeffs": (V, [ctypes.c_int, ctypes.c_double, _D]),
            "orc_cyclic_solve": (V, [ctypes.c_int, ctypes.c_double, _D, _D]),
            "orc_gcbc_matrices": (V, [_D, ctypes.c_double, _D, ctypes.c_double, _D, ctypes.c_int, _D, _D, _D]),
            "orc_create": (P, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _I, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
            "orc_destroy": (V, [P]),
            "orc_set_interface_table": (V, [P, _D, _D]),
            "orc_drva": (_D, [P, ctypes.c_int]),
            "orc_send": (_D, [P, ctypes.c_int]),
            "orc_recv": (_D, [P, ctypes.c_int]),
            "orc_stress": (_D, [P]),
            "orc_heat": (_D, [P]),
            "orc_buffer": (_D, [P]),
            "orc_buffer_deriv": (_D, [P]),
            "orc_fill": (V, [P, ctypes.c_int, _D, ctypes.c_int]),
            "orc_mpigo": (V, [P, ctypes.c_int]),
            "orc_deriv": (V, [P, ctypes.c_int, _D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "orc_deriv_field": (V, [P, ctypes.c_int, _D, _D, ctypes.c_int]),
            "orc_filter": (V, [P, ctypes.c_int, _D, _D]),
            "orc_primitives": (V, [ctypes.c_long, _D, _D, _D]),
            "orc_fluxes_pre": (V, [ctypes.c_int, ctypes.c_long, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D]),
            "orc_time_step": (ctypes.c_double, [ctypes.c_long, _D, _D, _D, _D, _D, _D, _D, ctypes.c_double, _D]),
            "orc_calc_viscous": (V, [P, _D, _D, _D, _D, _D, _D]),
            "orc_calc_fluxes": (V, [P, ctypes.c_int, _D, _D, _D, _D, _D, _D, _D]),
            "orc_gcbc": (ctypes.c_int, [P, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, ctypes.c_double]),
            "orc_wall_condition_update": (V, [P, _D, _I, _D]),
            "orc_rk_stage": (V, [ctypes.c_long, _D, _D, _D, _D, ctypes.c_double]),
            "orc_rhs": (ctypes.c_int, [P, ctypes.c_int, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, ctypes.c_double]),
            "orc_rk4_step": (ctypes.c_double, [P, ctypes.c_int, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D,
                                               _D, _D, _D, _D, _D, _D, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
            "orc_num_threads": (ctypes.c_int, []),
            "orc_set_num_threads": (V, [ctypes.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    """double* of a C-contiguous float64 array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, "need C-contiguous float64"
    return a.ctypes.data_as(_D)


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I)


def interface_table(corrected=False):
    pbco = np.zeros(16)
    pbci = np.zeros(16)
    lib().orc_interface_table(int(corrected), ptr(pbco), ptr(pbci))
    return pbco, pbci


def filter_coeffs(M, af):
    an = np.zeros(M + 1)
    lib().orc_filter_coeffs(M, af, ptr(an))
    return an


def solve_tridiag(a, b, c, r, solver):
    a, b, c, r = (np.array(v, dtype=np.float64) for v in (a, b, c, r))
    x = np.zeros_like(r)
    rc = lib().orc_solve_tridiag(len(r), solver, ptr(a), ptr(b), ptr(c), ptr(r), ptr(x))
    if rc != 0:
        raise ValueError("PCR needs a power-of-two length")
    return x


def build_deriv_rows(f, nstart, nend, recv4, pbci):
    f = np.ascontiguousarray(f, dtype=np.float64)
    n = len(f)
    a, b, c, r = (np.zeros(n) for _ in range(4))
    recv4 = np.ascontiguousarray(recv4, dtype=np.float64)
    lib().orc_build_deriv_rows(n, ptr(f), nstart, nend, ptr(recv4), ptr(np.ascontiguousarray(pbci)),
                               ptr(a), ptr(b), ptr(c), ptr(r))
    return a, b, c, r


def gcbc_matrices(cm, rho, mom, p, umf, fixed=True):
    R = np.zeros((5, 5))
    L = np.zeros((5, 5))
    info = np.zeros(3)
    lib().orc_gcbc_matrices(ptr(np.ascontiguousarray(cm, dtype=np.float64)), float(rho),
                            ptr(np.ascontiguousarray(mom, dtype=np.float64)), float(p),
                            ptr(np.ascontiguousarray(umf, dtype=np.float64)), int(fixed),
                            ptr(R), ptr(L), ptr(info))
    return R, L, info


def face_size(dims, axis):
    lxi, let, lze = dims
    return (let * lze, lxi * lze, lxi * let)[axis]


class Context:
    """numerics_base + physics_base state of one subdomain, on the CPU.

    dims = (lxi, let, lze); nbc = [[face0 x,y,z], [face1 x,y,z]] BC codes.
    table: 0 = as-shipped interface table, 1 = D1-corrected.
    solver: 0 = reference PCR where valid else Thomas, 1 = Thomas, 2 = dense.
    """

    def __init__(self, dims, nbc, table=0, solver=0, af=0.49):
        self.dims = tuple(int(d) for d in dims)
        self.lmx = self.dims[0] * self.dims[1] * self.dims[2]
        self.nbc = np.ascontiguousarray(np.asarray(nbc, dtype=np.int32).reshape(2, 3))
        self._h = lib().orc_create(self.dims[0], self.dims[1], self.dims[2], iptr(self.nbc),
                                   int(table), int(solver), float(af))

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().orc_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def _view(self, p, n):
        return np.ctypeslib.as_array(p, shape=(n,))

    def drva(self, axis):
        fs = face_size(self.dims, axis)
        return self._view(lib().orc_drva(self._h, axis), 2 * 5 * fs).reshape(2, 5, fs)

    def send(self, axis):
        fs = face_size(self.dims, axis)
        return self._view(lib().orc_send(self._h, axis), 5 * 4 * fs).reshape(5, 2, 2, fs)

    def recv(self, axis):
        fs = face_size(self.dims, axis)
        return self._view(lib().orc_recv(self._h, axis), 5 * 4 * fs).reshape(5, 2, 2, fs)

    def stress(self):
        return self._view(lib().orc_stress(self._h), 6 * self.lmx).reshape(6, self.lmx).copy()

    def heat(self):
        return self._view(lib().orc_heat(self._h), 3 * self.lmx).reshape(3, self.lmx).copy()

    def set_interface_table(self, pbco, pbci):
        lib().orc_set_interface_table(self._h, ptr(np.ascontiguousarray(pbco, dtype=np.float64)),
                                      ptr(np.ascontiguousarray(pbci, dtype=np.float64)))

    def fill(self, axis, field, var):
        lib().orc_fill(self._h, axis, ptr(field), var)

    def mpigo(self, var):
        lib().orc_mpigo(self._h, var)

    def deriv(self, axis, fin, fout, nstart, nend, var):
        lib().orc_deriv(self._h, axis, ptr(fin), ptr(fout), nstart, nend, var)

    def deriv_field(self, axis, fin, var=0):
        out = np.zeros(self.lmx)
        lib().orc_deriv_field(self._h, axis, ptr(np.ascontiguousarray(fin.reshape(-1))), ptr(out), var)
        return out

    def filter(self, axis, fin):
        out = np.zeros(self.lmx)
        lib().orc_filter(self._h, axis, ptr(np.ascontiguousarray(fin.reshape(-1))), ptr(out))
        return out

    def calc_viscous(self, de, ssk, xim, etm, zem, yaco):
        lib().orc_calc_viscous(self._h, ptr(de), ptr(ssk), ptr(xim), ptr(etm), ptr(zem), ptr(yaco))

    def calc_fluxes(self, visc, qa, p, de, xim, etm, zem, umf):
        lib().orc_calc_fluxes(self._h, int(visc), ptr(qa), ptr(p), ptr(de), ptr(xim), ptr(etm), ptr(zem),
                              ptr(np.asarray(umf, dtype=np.float64)))

    def gcbc(self, cm, qa, de, p, yaco, sbcc, umf, dudtmf, dtwi, drvb=(None, None, None)):
        rc = lib().orc_gcbc(self._h, ptr(cm[0]), ptr(cm[1]), ptr(cm[2]),
                            ptr(drvb[0]), ptr(drvb[1]), ptr(drvb[2]),
                            ptr(qa), ptr(de), ptr(p), ptr(yaco), ptr(sbcc),
                            ptr(np.asarray(umf, dtype=np.float64)),
                            ptr(np.asarray(dudtmf, dtype=np.float64)), float(dtwi))
        if rc != 0:
            raise ValueError("inter-curv face needs drvb")

    def wall_condition_update(self, qa, npex, umf):
        lib().orc_wall_condition_update(self._h, ptr(qa), iptr(npex), ptr(np.asarray(umf, dtype=np.float64)))

    def rhs(self, visc, qa, de, p, ssk, xim, etm, zem, yaco, cm, sbcc, umf, dudtmf=(0.0, 0.0, 0.0), dtwi=0.0):
        rc = lib().orc_rhs(self._h, int(visc), ptr(qa), ptr(de), ptr(p), ptr(ssk), ptr(xim), ptr(etm),
                           ptr(zem), ptr(yaco), ptr(cm[0]), ptr(cm[1]), ptr(cm[2]), ptr(sbcc),
                           ptr(np.asarray(umf, dtype=np.float64)), ptr(np.asarray(dudtmf, dtype=np.float64)),
                           float(dtwi))
        if rc != 0:
            raise ValueError("inter-curv face needs drvb")

    def rk4_step(self, visc, q, ssk, xim, etm, zem, yaco, cm, sbcc, umf, cfl=0.95, dt=0.0, filt=True,
                 dudtmf=(0.0, 0.0, 0.0)):
        lmx = self.lmx
        q0 = np.zeros(5 * lmx)
        de = np.zeros(5 * lmx)
        p = np.zeros(lmx)
        work = np.zeros(lmx)
        return lib().orc_rk4_step(self._h, int(visc), ptr(q), ptr(q0), ptr(de), ptr(p), ptr(work), ptr(ssk),
                                  ptr(xim), ptr(etm), ptr(zem), ptr(yaco), ptr(cm[0]), ptr(cm[1]), ptr(cm[2]),
                                  ptr(sbcc), ptr(np.asarray(umf, dtype=np.float64)),
                                  ptr(np.asarray(dudtmf, dtype=np.float64)), float(cfl), float(dt), int(filt))


def primitives(qa, lmx):
    de = np.zeros(5 * lmx)
    p = np.zeros(lmx)
    lib().orc_primitives(lmx, ptr(qa), ptr(de), ptr(p))
    return de, p


def fluxes_pre(visc, lmx, qa, p, de, xim, etm, zem, stress, heat, umf):
    buf = np.zeros(15 * lmx)
    lib().orc_fluxes_pre(int(visc), lmx, ptr(buf), ptr(qa), ptr(p), ptr(de), ptr(xim), ptr(etm), ptr(zem),
                         ptr(stress), ptr(heat), ptr(np.asarray(umf, dtype=np.float64)))
    return buf


def time_step(lmx, xim, etm, zem, de, yaco, ssk, umf, cfl):
    out = np.zeros(2)
    dte = lib().orc_time_step(lmx, ptr(xim), ptr(etm), ptr(zem), ptr(de), ptr(yaco), ptr(ssk),
                              ptr(np.asarray(umf, dtype=np.float64)), float(cfl), ptr(out))
    return dte, out


def rk_stage(lmx, q, q0, de, yaco, dtk):
    lib().orc_rk_stage(lmx, ptr(q), ptr(q0), ptr(de), ptr(yaco), float(dtk))


def num_threads():
    return lib().orc_num_threads()


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))
