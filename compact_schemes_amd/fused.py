"""The fused right-hand side and RK4 step of the HIP library (cs_solver).

Six line sweeps per RHS on a uniform Cartesian grid with the characteristic
BCs and the RK stage update inside the sweeps; see include/compact_schemes.h.
"""
import ctypes

import numpy as np

from . import Dims, Point, SolverConfig, _p, _stream, check, lib

# algorithmic HBM bytes per grid point of each sweep (FP64, viscous RHS):
# fields read + fields written once each
# by launch position: three gradient sweeps (the third closes the stress
# tensor), three flux sweeps (the third writes the RHS); sweep 6 = the last
# flux sweep with the RK stage update, which also reads q0
S1_AXES = {0: (0, 1, 2), 1: (1, 2, 0)}
S2_AXES = {0: (0, 1, 2), 1: (2, 1, 0)}


def sweep_names(order):
    a1, a2 = S1_AXES[order], S2_AXES[order]
    x = "xyz"
    return [f"S1{x[a1[0]]} grad", f"S1{x[a1[1]]} grad", f"S1{x[a1[2]]} grad+stress", f"S2{x[a2[0]]} flux",
            f"S2{x[a2[1]]} flux", f"S2{x[a2[2]]} flux", f"S2{x[a2[2]]} flux+RK stage", "filter x+y+z (3 launches)",
            "filter x", "filter y", "filter z"]


SWEEP_NAMES = sweep_names(1)
SWEEP_BYTES_VISC = [8 * (5 + 4), 8 * (5 + 4), 8 * (5 + 8 + 9), 8 * (5 + 4 + 5), 8 * (5 + 4 + 5 + 5),
                    8 * (5 + 4 + 5 + 5), 8 * (5 + 4 + 5 + 5 + 5), 3 * 8 * (5 + 5)] + [8 * (5 + 5)] * 3
SWEEP_BYTES_EULER = [0, 0, 0, 8 * (5 + 5), 8 * (5 + 5 + 5), 8 * (5 + 5 + 5), 8 * (5 + 5 + 5 + 5),
                     3 * 8 * (5 + 5)] + [8 * (5 + 5)] * 3
# viscous intermediates (4 gradients per gradient sweep, stress and heat flux) in FP32
SWEEP_BYTES_VISC_F32 = [8 * 5 + 4 * 4, 8 * 5 + 4 * 4, 8 * 5 + 4 * 8 + 4 * 9, 8 * (5 + 5) + 4 * 4,
                        8 * (5 + 5 + 5) + 4 * 4, 8 * (5 + 5 + 5) + 4 * 4, 8 * (5 + 5 + 5 + 5) + 4 * 4,
                        3 * 8 * (5 + 5)] + [8 * (5 + 5)] * 3
# the filter sweeps (one per axis) that close an RK4 step
FILTER_SWEEPS = [8, 9, 10]
# the sweeps of one RK stage, in launch order
RK_STAGE_SWEEPS = [0, 1, 2, 3, 4, 6]


def _uniform(a, name):
    v = float(a.reshape(-1)[0])
    if not np.all(a == v):
        raise ValueError(f"fused path needs a uniform {name}")
    return v


class FusedSolver:
    """cs_solver for one subdomain.  state: dict of numpy arrays as produced by
    compact_schemes_amd.workloads.flow_state (ssk must be uniform; metrics and
    yaco uniform and diagonal unless curvilinear=True, which hands the
    per-point metrics, yaco and face normals cm to cs_solver_set_metrics)."""

    def __init__(self, dims, nbc, state, viscous=True, cfl=0.95, alpha_f=0.49, filt=True, curvilinear=False):
        import torch
        self.dims = tuple(int(d) for d in dims)
        lmx = self.dims[0] * self.dims[1] * self.dims[2]
        self.lmx = lmx
        self.curvilinear = bool(curvilinear)
        xim, etm, zem = (np.asarray(state[k]) for k in ("xim", "etm", "zem"))
        if self.curvilinear:
            metric = (1.0, 1.0, 1.0)
        else:
            metric = (_uniform(xim[:lmx], "xi_x"), _uniform(etm[lmx:2 * lmx], "eta_y"),
                      _uniform(zem[2 * lmx:], "zeta_z"))
            for arr, zero in ((xim, (1, 2)), (etm, (0, 2)), (zem, (0, 1))):
                for z in zero:
                    if np.any(arr[z * lmx:(z + 1) * lmx] != 0):
                        raise ValueError("fused path needs diagonal metrics unless curvilinear=True")
        cfg = SolverConfig()
        cfg.dims = Dims(*self.dims)
        flat = [int(v) for row in nbc for v in row]
        cfg.nbc = (ctypes.c_int * 6)(*flat)
        cfg.viscous = int(viscous)
        cfg.umf = Point(*state["umf"])
        cfg.mu = _uniform(np.asarray(state["ssk"]), "ssk")
        cfg.metric = (ctypes.c_double * 3)(*metric)
        cfg.yaco = 1.0 if self.curvilinear else _uniform(np.asarray(state["yaco"]), "yaco")
        cfg.cfl = float(cfl)
        cfg.alpha_f = float(alpha_f)
        cfg.filter = int(filt)
        self.cfg = cfg
        self.viscous = bool(viscous)
        h = ctypes.c_void_p()
        check(lib().cs_solver_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.sbcc = torch.as_tensor(np.asarray(state["sbcc"], dtype=np.float64)).cuda().contiguous()
        if self.curvilinear:
            # device copies the solver reads per point; kept alive with it
            dev = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
            self.metric_fields = [dev(xim), dev(etm), dev(zem), dev(state["yaco"])] + [dev(c) for c in state["cm"]]
            check(lib().cs_solver_set_metrics(self.h, *[_p(t) for t in self.metric_fields]))
        self.de = torch.zeros(5 * lmx, dtype=torch.float64, device="cuda")
        self.num = None
        self.nbc_flat = flat
        self.order = 1
        self.precision = 0
        # the filter of cs_solver_rk4_step runs one fused sweep per axis where the
        # axis splits into equal chunks (x: 8-point chunks up to 512 points; y/z:
        # 8 or 16 points up to 256, 16 up to 512), else a line-solver launch
        self.filter_axes = [bool(filt) and ((n % 8 == 0 and n <= 256) or (a == 0 and n % 8 == 0 and n <= 512) or
                                            (a > 0 and n % 16 == 0 and n <= 512))
                            for a, n in enumerate(self.dims)]
        self.filter_fused = all(self.filter_axes)

    def close(self):
        if getattr(self, "h", None):
            lib().cs_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rhs(self, q, de=None, stream=None):
        de = self.de if de is None else de
        check(lib().cs_solver_rhs(self.h, _p(q), _p(de), _p(self.sbcc), _stream(stream)))
        return de

    def sweep(self, which, q, de=None, stream=None):
        de = self.de if de is None else de
        check(lib().cs_solver_sweep(self.h, int(which), _p(q), _p(de), _p(self.sbcc), _stream(stream)))

    def tune(self, axis, variant, lines=16, xcd=0):
        check(lib().cs_solver_tune(self.h, int(axis), int(variant), int(lines), int(xcd)))

    def tune_filter(self, axis, fields, lines=16, chunk=8):
        check(lib().cs_solver_tune_filter(self.h, int(axis), int(fields), int(lines), int(chunk)))

    def get_tune(self, axis):
        """((variant, lines, xcd) of the RHS sweeps, (fields, lines, chunk) of
        the fused filter sweep or (0, 0, 0)) along axis."""
        out = (ctypes.c_int * 6)()
        check(lib().cs_solver_get_tune(self.h, int(axis), out))
        return tuple(out[:3]), tuple(out[3:])

    def autotune(self, q, reps=3, stream=None):
        """Times the admissible sweep and filter shapes on q, applies the
        fastest and records them in the process-wide cache
        (compact_schemes_amd.tune); overwrites the solver's stage buffers."""
        check(lib().cs_solver_autotune(self.h, _p(q), _p(self.sbcc), int(reps), _stream(stream)))

    def set_precision(self, mode):
        """1: viscous intermediates stored in FP32; 0: FP64 (cs_solver_set_precision)."""
        check(lib().cs_solver_set_precision(self.h, int(mode)))
        self.precision = int(mode)

    def set_order(self, order):
        check(lib().cs_solver_set_order(self.h, int(order)))
        self.order = int(order)

    def kernel_name(self, which):
        """Name of the HIP kernel instantiation a sweep launches (as profilers
        print it, without the argument list)."""
        if which == 7:
            return " + ".join(self.kernel_name(w) for w in FILTER_SWEEPS)
        if which >= 8:
            ax = which - 8
            fields, lines, chunk = self.get_tune(ax)[1]
            sf = f"cs::fused::SF{'1' if fields == 1 else ''}<{ax}>"
            if ax == 0:
                return f"cs::fused::xt_kernel<8, {sf}, false>"
            return f"cs::fused::direct_kernel<{chunk}, {ax}, {lines}, {sf}, false>"
        ax = S1_AXES[self.order][which] if which < 3 else S2_AXES[self.order][min(which, 5) - 3]
        n = self.dims[ax]
        variant, lines, _ = self.get_tune(ax)[0]
        K = 8 if (n <= 256 or (ax == 0 and n % 8 == 0)) and variant != 2 else 16
        P = (n + K - 1) // K
        exact = n == P * K
        nbc = self.nbc_flat
        gcbc = nbc[ax] == 10 and nbc[3 + ax] == 10
        f32 = ("C" if self.curvilinear else "") + ("F" if (self.viscous and self.precision == 1) else "")
        if which < 3:
            sw, gc = f"cs::fused::S1{f32}<{ax}, {str(which == 2).lower()}>", "false"
        else:
            mode = {3: 0, 4: 1, 5: 2, 6: 3}[which]
            sw, gc = f"cs::fused::S2{f32}<{ax}, {mode}, {str(self.viscous).lower()}>", str(gcbc).lower()
        if variant == 1 and K == 8:
            return f"cs::fused::tile_kernel<{K}, {ax}, {lines}, {str(exact).lower()}, {sw}, {gc}>"
        if exact and ax == 0 and K == 8:
            return f"cs::fused::xt_kernel<{K}, {sw}, {gc}>"
        if exact:
            tl = 16 if ax == 0 or lines != 8 else 8
            return f"cs::fused::direct_kernel<{K}, {ax}, {tl}, {sw}, {gc}>"
        return f"cs::fused::ragged_kernel<{K}, {ax}, {sw}, {gc}>"

    def time_step(self, q, stream=None):
        dt = ctypes.c_double()
        check(lib().cs_solver_time_step(self.h, _p(q), ctypes.byref(dt), _stream(stream)))
        return dt.value

    def rk4_step(self, q, dt=0.0, report_dt=False, stream=None):
        out = ctypes.c_double()
        check(lib().cs_solver_rk4_step(self.h, _p(q), _p(self.sbcc), float(dt),
                                       ctypes.byref(out) if report_dt else None, _stream(stream)))
        return out.value if report_dt else None

    def rhs_vjp(self, q, lam, g=None, stream=None):
        """g = (dR/dq)^T lam at q (inviscid solvers)."""
        import torch
        g = torch.empty_like(lam) if g is None else g
        check(lib().cs_solver_rhs_vjp(self.h, _p(q), _p(lam), _p(g), _p(self.sbcc), _stream(stream)))
        return g

    def rk4_vjp(self, q0, lam, dt, stream=None):
        """lam <- (d step(q0) / d q0)^T lam for one RK4 step with fixed dt (in place)."""
        check(lib().cs_solver_rk4_vjp(self.h, _p(q0), _p(lam), float(dt), _p(self.sbcc), _stream(stream)))
        return lam

    def tape_size(self):
        """Doubles of the tape of one RK4 step (cs_solver_tape_size)."""
        n = ctypes.c_size_t()
        check(lib().cs_solver_tape_size(self.h, ctypes.byref(n)))
        return n.value

    def new_tape(self):
        import torch
        return torch.empty(self.tape_size(), dtype=torch.float64, device="cuda")

    def rk4_step_tape(self, q, tape, dt=0.0, report_dt=False, stream=None):
        """rk4_step that records the step's tape (inviscid solvers)."""
        out = ctypes.c_double()
        check(lib().cs_solver_rk4_step_tape(self.h, _p(q), _p(self.sbcc), float(dt),
                                            ctypes.byref(out) if report_dt else None, _p(tape), _stream(stream)))
        return out.value if report_dt else None

    def rk4_vjp_tape(self, tape, lam, dt, stream=None):
        """lam <- (d step / d q0)^T lam for the step that recorded `tape` (in place)."""
        check(lib().cs_solver_rk4_vjp_tape(self.h, _p(tape), _p(lam), float(dt), _p(self.sbcc), _stream(stream)))
        return lam

    def set_adjoint_mode(self, mode):
        """1: fused transposed sweeps (default, lines of 8-point chunks); 0: per-operator products."""
        check(lib().cs_solver_adjoint_mode(self.h, int(mode)))

    def time_sweeps(self, q, stream, reps=10, sweeps=None):
        """Average duration (ms) of each sweep, HIP events on the launch stream
        (default: the sweeps of one RK stage)."""
        import torch
        self.time_step(q, stream=stream)
        self.rhs(q, stream=stream)
        res = []
        nbytes = (SWEEP_BYTES_VISC_F32 if self.precision == 1 else SWEEP_BYTES_VISC) if self.viscous \
            else SWEEP_BYTES_EULER
        for w in (RK_STAGE_SWEEPS if sweeps is None else sweeps):
            if nbytes[w] == 0:
                continue
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            src = self.de if w >= 7 else q           # the filter runs in place on the scratch RHS
            if w >= 7:
                self.de.copy_(q)
            self.sweep(w, src, stream=stream)
            e0.record(stream)
            for _ in range(reps):
                self.sweep(w, src, stream=stream)
            e1.record(stream)
            torch.cuda.synchronize()
            res.append({"which": w, "name": sweep_names(self.order)[w], "kernel": self.kernel_name(w),
                        "ms": e0.elapsed_time(e1) / reps, "bytes": float(nbytes[w]) * self.lmx})
        return res

    def time_dominant_kernel(self, q, stream, reps=10):
        """The sweep with the longest launch among those of one RK stage."""
        sw = self.time_sweeps(q, stream, reps)
        dom = max(sw, key=lambda r: r["ms"])
        out = dict(dom)
        out["all"] = [{"sweep": r["name"], "kernel": r["kernel"], "ms": round(r["ms"], 5),
                       "GBs": round(r["bytes"] / (r["ms"] * 1e-3) / 1e9, 1)} for r in sw]
        return out
