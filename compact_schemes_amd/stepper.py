"""Reference-ordered right-hand side and RK4 step over the C-ABI operators.

The sequence is the driver loop of the reference (driver/canard.cu:132-205)
with the stage update and gcbc enabled: primitives -> calc_time_step (first
stage) -> calc_viscous_shear_stress -> calc_fluxes -> gcbc_go -> RK stage
update; after the fourth stage the compact filter is applied to the five
conserved variables along x, y and z.  Every operation is one call into the
HIP library; this module only sequences them.
"""
import torch

from . import Numerics, Physics, gcbc_go, primitives, rk_stage


class Stepper:
    def __init__(self, dims, nbc, viscous, umf, state):
        self.dims = tuple(dims)
        self.lmx = dims[0] * dims[1] * dims[2]
        self.num = Numerics(dims, nbc)
        self.phys = Physics(dims, viscous, umf)
        self.viscous = viscous
        self.umf = tuple(umf)
        t = lambda a: torch.as_tensor(a, dtype=torch.float64).cuda().contiguous()
        self.xim, self.etm, self.zem = t(state["xim"]), t(state["etm"]), t(state["zem"])
        self.yaco, self.ssk, self.sbcc = t(state["yaco"]), t(state["ssk"]), t(state["sbcc"])
        self.cm = [t(c) for c in state["cm"]]
        self.de = torch.zeros(5 * self.lmx, dtype=torch.float64, device="cuda")
        self.p = torch.zeros(self.lmx, dtype=torch.float64, device="cuda")
        self.q0 = torch.zeros(5 * self.lmx, dtype=torch.float64, device="cuda")
        self.work = torch.zeros(self.lmx, dtype=torch.float64, device="cuda")
        self.char = any(c in (10, 20, 25, 30) for row in nbc for c in row)

    def time_step(self, q, cfl):
        primitives(self.dims, q, self.de, self.p)
        dt, _ = self.phys.calc_time_step(self.xim, self.etm, self.zem, self.de, self.yaco, self.ssk, cfl)
        return dt

    def rhs(self, q, dtwi=0.0):
        primitives(self.dims, q, self.de, self.p)
        if self.viscous:
            self.phys.calc_viscous_shear_stress(self.num, self.de, self.ssk, self.xim, self.etm, self.zem, self.yaco)
        self.phys.calc_fluxes(self.num, q, self.p, self.de, self.xim, self.etm, self.zem)
        if self.char:
            gcbc_go(self.num, self.cm, q, self.de, self.p, self.yaco, self.sbcc, self.umf, (0.0, 0.0, 0.0), dtwi)
        return self.de

    def rk4_step(self, q, dt=None, cfl=0.95, filt=True):
        self.q0.copy_(q)
        if dt is None:
            dt = self.time_step(q, cfl)
        for nk in range(1, 5):
            self.rhs(q, 1.0 / dt)
            rk_stage(self.dims, q, self.q0, self.de, self.yaco, dt / (4 - nk + 1))
        if filt:
            for v in range(5):
                f = q[v * self.lmx:(v + 1) * self.lmx]
                self.num.filter(0, f, self.work)
                self.num.filter(1, self.work, f)
                self.num.filter(2, f, self.work)
                f.copy_(self.work)
        return dt
