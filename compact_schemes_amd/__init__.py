"""compact_schemes_amd -- MI355X (gfx950) compact-scheme hot path.

The product is the C-ABI shared library ``libcompact_schemes_amd.so`` built
from ``csrc/`` (declared in ``include/compact_schemes.h``).  This module is a
thin ctypes binding over it for Python callers and tests: device memory and
streams come from PyTorch, every computation runs in the HIP library.  There
is no CPU fallback: importing works without a GPU, but every call needs the
library and a device, and raises otherwise.
"""
import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcompact_schemes_amd.so")
CSRC = os.path.join(_HERE, "csrc")

BC_NON_REFLECTIVE = 10
BC_WALL_INVISCID = 20
BC_WALL_VISCOUS = 25
BC_INTER_CURV = 30
BC_INTER_STRAIGHT = 35
BC_INTER_SUBDOMAINS = 40
BC_PERIODIC = 45

ERRORS = {1: "CS_ERR_ARG", 2: "CS_ERR_SHAPE", 3: "CS_ERR_HIP", 4: "CS_ERR_UNSUPPORTED", 5: "CS_ERR_STATE"}

# every symbol include/compact_schemes.h declares (checked by the CPU tests)
EXPORTS = [
    "cs_last_error", "cs_version",
    "cs_numerics_create", "cs_numerics_destroy", "cs_deriv_setup", "cs_set_interface_table",
    "cs_get_interface_table", "cs_numerics_buffers", "cs_numerics_ndf", "cs_fill_buffers", "cs_mpigo", "cs_exchange", "cs_metrics",
    "cs_deriv1d", "cs_deriv2d", "cs_deriv", "cs_filter", "cs_set_filter_alpha",
    "cs_physics_create", "cs_physics_destroy", "cs_physics_set_frame", "cs_physics_buffers",
    "cs_primitives", "cs_calc_fluxes", "cs_calc_viscous_shear_stress", "cs_calc_time_step",
    "cs_calc_time_step_noalloc", "cs_gcbc_go", "cs_wall_condition_update", "cs_rk_stage",
    "cs_set_precision", "cs_deriv_adjoint", "cs_filter_adjoint", "cs_solver_create", "cs_solver_destroy", "cs_solver_rhs", "cs_solver_sweep", "cs_solver_tune", "cs_solver_tune_filter",
    "cs_solver_get_tune", "cs_solver_autotune",
    "cs_solver_set_order", "cs_solver_set_precision", "cs_solver_set_metrics", "cs_solver_rhs_vjp", "cs_solver_rk4_vjp", "cs_solver_tape_size",
    "cs_solver_rk4_step_tape", "cs_solver_rk4_vjp_tape", "cs_solver_adjoint_mode",
    "cs_batch_create", "cs_batch_destroy", "cs_batch_bind", "cs_batch_deriv", "cs_batch_filter", "cs_batch_step",
    "cs_batch_capture", "cs_batch_replay", "cs_batch_get_persistent", "cs_batch_set_persistent", "cs_batch_info", "cs_batch_autotune", "cs_tune_set", "cs_tune_get",
    "cs_tune_set_sweep", "cs_tune_get_sweep", "cs_tune_set_filter_sweep", "cs_tune_get_filter_sweep", "cs_tune_clear", "cs_tune_save", "cs_tune_load",
    "cs_solver_time_step", "cs_solver_rk4_step",
    "cs_init_pulse", "cs_boundary_weights", "cs_config_defaults", "cs_config_load", "cs_config_dims", "cs_config_grid", "cs_vtk_write",
]


class CsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Dims(ctypes.Structure):
    _fields_ = [("lxi", ctypes.c_int), ("let", ctypes.c_int), ("lze", ctypes.c_int)]


class Point(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double), ("z", ctypes.c_double)]


class SolverConfig(ctypes.Structure):
    _fields_ = [("dims", Dims), ("nbc", ctypes.c_int * 6), ("viscous", ctypes.c_int), ("umf", Point),
                ("mu", ctypes.c_double), ("metric", ctypes.c_double * 3), ("yaco", ctypes.c_double),
                ("cfl", ctypes.c_double), ("alpha_f", ctypes.c_double), ("filter", ctypes.c_int)]


class RunConfig(ctypes.Structure):
    """cs_run_config: the run configuration (config.yaml schema + run section)."""
    _fields_ = [("nblocks", ctypes.c_int), ("nbpc", ctypes.c_int * 3), ("lximb", ctypes.c_int),
                ("letmb", ctypes.c_int), ("lzemb", ctypes.c_int), ("nbbc", ctypes.c_int * 6),
                ("mbcd", ctypes.c_int * 6), ("doml0", ctypes.c_double), ("doml1", ctypes.c_double),
                ("domh", ctypes.c_double), ("span", ctypes.c_double), ("amach", Point),
                ("cfl", ctypes.c_double), ("tmax", ctypes.c_double), ("ndata", ctypes.c_int),
                ("nsteps", ctypes.c_int), ("viscous", ctypes.c_int), ("re", ctypes.c_double),
                ("alpha_f", ctypes.c_double), ("filter", ctypes.c_int), ("pulse_amp", ctypes.c_double),
                ("pulse_width", ctypes.c_double)]


_lib = None


def build(force=True):
    """Compile the HIP library for gfx950 in-tree (hipcc cross-compiles without a GPU)."""
    args = ["make", "-s", "-C", CSRC, "-j8"]
    if force:
        args.insert(2, "-B")
    subprocess.run(args, check=True)


def lib():
    """The loaded C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                               "(the HIP library is required; there is no fallback)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.c_double
        PP = ctypes.POINTER(ctypes.c_void_p)
        DP = ctypes.POINTER(ctypes.c_double)
        sig = {
            "cs_last_error": (ctypes.c_char_p, []),
            "cs_version": (ctypes.c_char_p, []),
            "cs_numerics_create": (I, [ctypes.POINTER(Dims), ctypes.POINTER(ctypes.c_int), PP]),
            "cs_numerics_destroy": (I, [P]),
            "cs_deriv_setup": (I, [P]),
            "cs_set_interface_table": (I, [P, DP, DP]),
            "cs_get_interface_table": (I, [P, DP, DP]),
            "cs_numerics_buffers": (I, [P, I, PP, PP, PP]),
            "cs_numerics_ndf": (I, [P, ctypes.POINTER(ctypes.c_int)]),
            "cs_fill_buffers": (I, [P, P, I, I, I, P]),
            "cs_mpigo": (I, [P, I, P]),
            "cs_exchange": (I, [P, I, P, I, I, I, P]),
            "cs_metrics": (I, [P, P, P, P, P, P, P, P]),
            "cs_deriv1d": (I, [P, I, P, P, I, I, I, P]),
            "cs_deriv2d": (I, [P, I, P, P, I, I, I, I, P]),
            "cs_deriv": (I, [P, I, P, P, I, P]),
            "cs_filter": (I, [P, I, P, P, P]),
            "cs_set_filter_alpha": (I, [P, D]),
            "cs_set_precision": (I, [P, I]),
            "cs_deriv_adjoint": (I, [P, I, P, P, P]),
            "cs_filter_adjoint": (I, [P, I, P, P, P]),
            "cs_physics_create": (I, [ctypes.POINTER(Dims), I, Point, PP]),
            "cs_physics_destroy": (I, [P]),
            "cs_physics_set_frame": (I, [P, Point, Point]),
            "cs_physics_buffers": (I, [P, PP, PP, PP, PP]),
            "cs_primitives": (I, [ctypes.POINTER(Dims), P, P, P, P]),
            "cs_calc_fluxes": (I, [P, P, P, P, P, P, P, P, P]),
            "cs_calc_viscous_shear_stress": (I, [P, P, P, P, P, P, P, P, P]),
            "cs_calc_time_step": (I, [P, P, P, P, P, P, P, D, DP, DP, P]),
            "cs_calc_time_step_noalloc": (I, [ctypes.POINTER(Dims), Point, P, P, P, P, P, P, D, P, P]),
            "cs_gcbc_go": (I, [P, PP, PP, P, P, P, P, P, Point, Point, D, P]),
            "cs_wall_condition_update": (I, [P, P, P, Point, P]),
            "cs_rk_stage": (I, [ctypes.POINTER(Dims), P, P, P, P, D, P]),
            "cs_solver_create": (I, [ctypes.POINTER(SolverConfig), PP]),
            "cs_solver_destroy": (I, [P]),
            "cs_solver_rhs": (I, [P, P, P, P, P]),
            "cs_solver_sweep": (I, [P, I, P, P, P, P]),
            "cs_solver_tune": (I, [P, I, I, I, I]),
            "cs_solver_tune_filter": (I, [P, I, I, I, I]),
            "cs_solver_get_tune": (I, [P, I, P]),
            "cs_solver_set_metrics": (I, [P, P, P, P, P, P, P, P]),
            "cs_solver_autotune": (I, [P, P, P, I, P]),
            "cs_solver_set_order": (I, [P, I]),
            "cs_solver_set_precision": (I, [P, I]),
            "cs_solver_rhs_vjp": (I, [P, P, P, P, P, P]),
            "cs_solver_rk4_vjp": (I, [P, P, P, D, P, P]),
            "cs_solver_tape_size": (I, [P, ctypes.POINTER(ctypes.c_size_t)]),
            "cs_solver_rk4_step_tape": (I, [P, P, P, D, P, P, P]),
            "cs_solver_rk4_vjp_tape": (I, [P, P, P, D, P, P]),
            "cs_solver_adjoint_mode": (I, [P, I]),
            "cs_batch_create": (I, [I, P, P, P]),
            "cs_batch_destroy": (I, [P]),
            "cs_batch_bind": (I, [P, P, P]),
            "cs_batch_deriv": (I, [P, P]),
            "cs_batch_filter": (I, [P, P]),
            "cs_batch_step": (I, [P, P]),
            "cs_batch_capture": (I, [P, P]),
            "cs_batch_set_persistent": (I, [P, I]),
            "cs_batch_get_persistent": (I, [P, ctypes.POINTER(ctypes.c_int)]),
            "cs_batch_replay": (I, [P, P]),
            "cs_batch_info": (I, [P, P, P]),
            "cs_batch_autotune": (I, [P, I, P]),
            "cs_tune_set": (I, [I, I, I, I, I, I]),
            "cs_tune_get": (I, [I, I, I, I, I, P]),
            "cs_tune_set_sweep": (I, [I, I, I, I, I, I, I, I]),
            "cs_tune_get_sweep": (I, [I, I, I, I, I, P, P]),
            "cs_tune_set_filter_sweep": (I, [I, I, I, I, I, I]),
            "cs_tune_get_filter_sweep": (I, [I, I, I, P, P]),
            "cs_tune_clear": (I, []),
            "cs_tune_save": (I, [ctypes.c_char_p]),
            "cs_tune_load": (I, [ctypes.c_char_p]),
            "cs_init_pulse": (I, [ctypes.POINTER(Dims), Point, D, D, P, P]),
            "cs_boundary_weights": (I, [ctypes.POINTER(Dims), P, P]),
            "cs_config_defaults": (I, [ctypes.POINTER(RunConfig)]),
            "cs_config_load": (I, [ctypes.c_char_p, ctypes.POINTER(RunConfig)]),
            "cs_config_dims": (I, [ctypes.POINTER(RunConfig), ctypes.POINTER(Dims), ctypes.POINTER(ctypes.c_int)]),
            "cs_config_grid": (I, [ctypes.POINTER(RunConfig), DP]),
            "cs_vtk_write": (I, [ctypes.c_char_p, ctypes.POINTER(Dims), DP, I, ctypes.POINTER(ctypes.c_char_p),
                                 ctypes.POINTER(DP), I]),
            "cs_solver_time_step": (I, [P, P, DP, P]),
            "cs_solver_rk4_step": (I, [P, P, P, D, DP, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise CsError(rc, lib().cs_last_error().decode())
    return rc


def _p(t):
    """Device pointer of a contiguous CUDA/HIP tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("field must live on the GPU")
    if not t.is_contiguous():
        raise ValueError("field must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class Numerics:
    """numerics_pc<double> of one subdomain (src/numerics_base.hpp, src/numerics_pc.hpp)."""

    def __init__(self, dims, nbc, setup=True):
        self.dims = tuple(int(d) for d in dims)
        self.lmx = self.dims[0] * self.dims[1] * self.dims[2]
        flat = [int(v) for row in nbc for v in row]
        arr = (ctypes.c_int * 6)(*flat)
        h = ctypes.c_void_p()
        check(lib().cs_numerics_create(ctypes.byref(Dims(*self.dims)), arr, ctypes.byref(h)))
        self.h = h
        self.nbc = [flat[0:3], flat[3:6]]
        if setup:
            self.deriv_setup()

    def close(self):
        if getattr(self, "h", None):
            lib().cs_numerics_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, mode):
        """0: FP64 line solves; k = 1..4: FP32 solve + k FP64-residual refinement sweeps."""
        check(lib().cs_set_precision(self.h, int(mode)))

    def deriv_adjoint(self, axis, fin, fout, stream=None):
        """fout = D^T fin, the transpose of deriv along axis."""
        check(lib().cs_deriv_adjoint(self.h, int(axis), _p(fin), _p(fout), _stream(stream)))

    def filter_adjoint(self, axis, fin, fout, stream=None):
        """fout = F^T fin, the transpose of filter along axis."""
        check(lib().cs_filter_adjoint(self.h, int(axis), _p(fin), _p(fout), _stream(stream)))

    def deriv_setup(self):
        check(lib().cs_deriv_setup(self.h))

    def set_interface_table(self, pbco, pbci):
        a = (ctypes.c_double * 16)(*[float(v) for v in pbco])
        b = (ctypes.c_double * 16)(*[float(v) for v in pbci])
        check(lib().cs_set_interface_table(self.h, a, b))

    def interface_table(self):
        a = (ctypes.c_double * 16)()
        b = (ctypes.c_double * 16)()
        check(lib().cs_get_interface_table(self.h, a, b))
        return list(a), list(b)

    def ndf(self):
        a = (ctypes.c_int * 6)()
        check(lib().cs_numerics_ndf(self.h, a))
        return [list(a[0:3]), list(a[3:6])]

    def face_size(self, axis):
        lxi, let, lze = self.dims
        return (let * lze, lxi * lze, lxi * let)[axis]

    def buffer_ptrs(self, axis):
        """Device addresses (drva, send, recv) of axis `axis`."""
        d, s, r = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().cs_numerics_buffers(self.h, axis, ctypes.byref(d), ctypes.byref(s), ctypes.byref(r)))
        return d.value, s.value, r.value

    def buffers(self, axis):
        """Copies (drva, send, recv) of the context's device buffers as torch tensors."""
        d, s, r = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().cs_numerics_buffers(self.h, axis, ctypes.byref(d), ctypes.byref(s), ctypes.byref(r)))
        fs = self.face_size(axis)
        return (_wrap(d.value, 2 * 5 * fs).view(2, 5, fs), _wrap(s.value, 5 * 4 * fs).view(5, 2, 2, fs),
                _wrap(r.value, 5 * 4 * fs).view(5, 2, 2, fs))

    def fill_buffers(self, infield, lo, hi, nrt, stream=None):
        check(lib().cs_fill_buffers(self.h, _p(infield), lo, hi, nrt, _stream(stream)))

    def mpigo(self, var, stream=None):
        check(lib().cs_mpigo(self.h, var, _stream(stream)))

    def exchange(self, face, other, other_face, axis, var, stream=None):
        """Seam exchange with another context on this GPU (cs_exchange)."""
        check(lib().cs_exchange(self.h, int(face), other.h, int(other_face), int(axis), int(var), _stream(stream)))

    def metrics(self, xyz, period=None, stream=None):
        """(xim, etm, zem, yaco) from [3][lmx] coordinates (cs_metrics);
        period: 3x3 coordinate jumps over one period of each seam axis."""
        import torch
        lmx = self.lmx
        xim, etm, zem = (torch.empty(3 * lmx, dtype=torch.float64, device="cuda") for _ in range(3))
        yaco = torch.empty(lmx, dtype=torch.float64, device="cuda")
        per = None
        if period is not None:
            per = (ctypes.c_double * 9)(*[float(v) for row in period for v in row])
            per = ctypes.cast(per, ctypes.c_void_p)
        check(lib().cs_metrics(self.h, _p(xyz), per, _p(xim), _p(etm), _p(zem), _p(yaco), _stream(stream)))
        return xim, etm, zem, yaco

    def deriv1d(self, axis, fin, fout, nstart, nend, var=0, stream=None):
        check(lib().cs_deriv1d(self.h, axis, _p(fin), _p(fout), nstart, nend, var, _stream(stream)))

    def deriv2d(self, axis, fin, fout, nstart, nend, var, comp, stream=None):
        check(lib().cs_deriv2d(self.h, axis, _p(fin), _p(fout), nstart, nend, var, comp, _stream(stream)))

    def deriv(self, axis, fin, fout, var=0, stream=None):
        check(lib().cs_deriv(self.h, axis, _p(fin), _p(fout), var, _stream(stream)))

    def filter(self, axis, fin, fout, stream=None):
        check(lib().cs_filter(self.h, axis, _p(fin), _p(fout), _stream(stream)))

    def set_filter_alpha(self, af):
        check(lib().cs_set_filter_alpha(self.h, float(af)))


class Physics:
    """physics<Visc, double> of one subdomain (src/physics_base.hpp, src/physics_pc.hpp)."""

    def __init__(self, dims, viscous=True, umf=(0.0, 0.0, 0.0)):
        self.dims = tuple(int(d) for d in dims)
        self.lmx = self.dims[0] * self.dims[1] * self.dims[2]
        self.viscous = bool(viscous)
        h = ctypes.c_void_p()
        check(lib().cs_physics_create(ctypes.byref(Dims(*self.dims)), int(viscous), Point(*umf), ctypes.byref(h)))
        self.h = h
        self.umf = tuple(umf)
        self.dudtmf = (0.0, 0.0, 0.0)

    def close(self):
        if getattr(self, "h", None):
            lib().cs_physics_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_frame(self, umf, dudtmf=(0.0, 0.0, 0.0)):
        check(lib().cs_physics_set_frame(self.h, Point(*umf), Point(*dudtmf)))
        self.umf, self.dudtmf = tuple(umf), tuple(dudtmf)

    def buffers(self):
        """Copies (stress [6][lmx], heat [3][lmx]) of the device storage."""
        st, ht, b, bd = (ctypes.c_void_p() for _ in range(4))
        check(lib().cs_physics_buffers(self.h, ctypes.byref(st), ctypes.byref(ht), ctypes.byref(b), ctypes.byref(bd)))
        out = []
        for ptr_, n in ((st, 6), (ht, 3)):
            out.append(_wrap(ptr_.value, n * self.lmx).view(n, self.lmx) if ptr_.value else None)
        return tuple(out)

    def calc_fluxes(self, num, qa, p, de, xim, etm, zem, stream=None):
        check(lib().cs_calc_fluxes(self.h, num.h, _p(qa), _p(p), _p(de), _p(xim), _p(etm), _p(zem),
                                   _stream(stream)))

    def calc_viscous_shear_stress(self, num, de, ssk, xim, etm, zem, yaco, stream=None):
        check(lib().cs_calc_viscous_shear_stress(self.h, num.h, _p(de), _p(ssk), _p(xim), _p(etm), _p(zem),
                                                 _p(yaco), _stream(stream)))

    def calc_time_step(self, xim, etm, zem, de, yaco, ssk, cfl, stream=None):
        dte = ctypes.c_double()
        m2 = (ctypes.c_double * 2)()
        check(lib().cs_calc_time_step(self.h, _p(xim), _p(etm), _p(zem), _p(de), _p(yaco), _p(ssk), float(cfl),
                                      ctypes.byref(dte), m2, _stream(stream)))
        return dte.value, (m2[0], m2[1])


def calc_time_step_noalloc(dims, umf, xim, etm, zem, de, yaco, ssk, cfl, out3, stream=None):
    """cs_calc_time_step_noalloc: out3 (device, 3 doubles) <- dte, max0, max1;
    no context, no allocation, no synchronisation."""
    check(lib().cs_calc_time_step_noalloc(ctypes.byref(Dims(*dims)), Point(*umf), _p(xim), _p(etm), _p(zem), _p(de),
                                          _p(yaco), _p(ssk), float(cfl), _p(out3), _stream(stream)))


def init_pulse(dims, center, amp, width, qa, stream=None):
    """cs_init_pulse: uniform flow (stored velocity 0) with a Gaussian pressure pulse."""
    check(lib().cs_init_pulse(ctypes.byref(Dims(*dims)), Point(*center), float(amp), float(width), _p(qa),
                              _stream(stream)))


def boundary_weights(dims, sbcc, stream=None):
    """cs_boundary_weights: sbcc = 1 on boundary points, 0 inside."""
    check(lib().cs_boundary_weights(ctypes.byref(Dims(*dims)), _p(sbcc), _stream(stream)))


def primitives(dims, qa, de, p, stream=None):
    check(lib().cs_primitives(ctypes.byref(Dims(*dims)), _p(qa), _p(de), _p(p), _stream(stream)))


def gcbc_go(num, cm, qa, de, p, yaco, sbcc, umf, dudtmf=(0.0, 0.0, 0.0), dtwi=0.0, drvb=None, stream=None):
    cma = (ctypes.c_void_p * 3)(*[t.data_ptr() if t is not None else None for t in cm])
    dvb = None
    if drvb is not None:
        dvb = (ctypes.c_void_p * 3)(*[t.data_ptr() if t is not None else None for t in drvb])
    check(lib().cs_gcbc_go(num.h, cma, dvb, _p(qa), _p(de), _p(p), _p(yaco), _p(sbcc), Point(*umf),
                           Point(*dudtmf), float(dtwi), _stream(stream)))


def wall_condition_update(num, qa, npex, umf, stream=None):
    check(lib().cs_wall_condition_update(num.h, _p(qa), _p(npex), Point(*umf), _stream(stream)))


def rk_stage(dims, q, q0, de, yaco, dtk, stream=None):
    check(lib().cs_rk_stage(ctypes.byref(Dims(*dims)), _p(q), _p(q0), _p(de), _p(yaco), float(dtk),
                            _stream(stream)))


_hip = None


def _hiprt():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.restype = ctypes.c_int
    return _hip


def _wrap(addr, n):
    """Copy n doubles of library-owned device memory into a new torch tensor."""
    import torch
    if not addr:
        return None
    t = torch.empty(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = _hiprt().hipMemcpy(ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(addr), 8 * n, 3)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed ({rc})")
    return t


def copy_to(addr, t):
    """Copy a contiguous float64 torch tensor into library-owned device memory."""
    import torch
    torch.cuda.synchronize()
    rc = _hiprt().hipMemcpy(ctypes.c_void_p(addr), ctypes.c_void_p(t.data_ptr()), 8 * t.numel(), 3)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed ({rc})")


# ---- run configuration, grid, VTK output (host code) ------------------------
def load_config(path):
    """cs_config_load: the run configuration of a file in config.yaml's schema."""
    cfg = RunConfig()
    check(lib().cs_config_load(os.fsencode(path), ctypes.byref(cfg)))
    return cfg


def config_defaults():
    cfg = RunConfig()
    check(lib().cs_config_defaults(ctypes.byref(cfg)))
    return cfg


def config_dims(cfg):
    """(dims, nbc) of the configuration's block: dims = (lxi, let, lze), nbc = [2][3]."""
    d = Dims()
    nbc = (ctypes.c_int * 6)()
    check(lib().cs_config_dims(ctypes.byref(cfg), ctypes.byref(d), nbc))
    return (d.lxi, d.let, d.lze), [list(nbc[0:3]), list(nbc[3:6])]


def config_grid(cfg):
    """cs_config_grid: host array [3][lmx] of the block's point coordinates."""
    import numpy as np
    dims, _ = config_dims(cfg)
    xyz = np.empty(3 * dims[0] * dims[1] * dims[2], dtype=np.float64)
    check(lib().cs_config_grid(ctypes.byref(cfg), xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return xyz


def write_vtk(path, dims, xyz, fields, precision=32):
    """cs_vtk_write: legacy VTK structured grid of host arrays; fields = {name: [lmx] array}."""
    import numpy as np
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in fields.values()]
    names = (ctypes.c_char_p * max(1, len(arrs)))(*[n.encode() for n in fields.keys()])
    ptrs = (ctypes.POINTER(ctypes.c_double) * max(1, len(arrs)))(
        *[a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in arrs])
    check(lib().cs_vtk_write(os.fsencode(path), ctypes.byref(Dims(*dims)),
                             xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(arrs), names, ptrs,
                             int(precision)))
