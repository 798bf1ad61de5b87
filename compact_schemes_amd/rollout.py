"""Gradients of long RK4 rollouts through the fused solver, with checkpoints.

The forward rollout keeps the state every `every` steps.  The adjoint runs
the segments last first: each segment is re-run from its checkpoint while
cs_solver_rk4_step_tape records one tape per step, then stepped backwards
with cs_solver_rk4_vjp_tape.  Memory: (steps / every) states plus `every`
tapes, instead of every stage of every step.

AcousticRollout is the acoustic workload on top of it: Euler flow at Mach 0.5
in x with an isentropic density perturbation rho' (p' = c0^2 rho', u' = 0),
loss = acoustic energy sum(p'^2 / (2 rho0 c0^2) + rho0 |u'|^2 / 2) on one
z-plane after a fixed number of steps, gradient with respect to rho'.
"""
import numpy as np


def checkpointed_vjp(solver, q0, nsteps, every, dt, loss_grad, tapes=None):
    """Runs nsteps RK4 steps of `solver` (fixed dt) from q0 (not modified),
    evaluates (L, lam) = loss_grad(q_final) and returns (L, lam0) with
    lam0 = (d q_final / d q0)^T lam."""
    import torch
    q = q0.clone()
    ck = {}
    for k in range(nsteps):
        if k % every == 0:
            ck[k] = q.clone()
        solver.rk4_step(q, dt=dt)
    L, lam = loss_grad(q)
    if tapes is None:
        tapes = [solver.new_tape() for _ in range(min(every, nsteps))]
    for k0 in sorted(ck, reverse=True):
        x = ck[k0].clone()
        m = min(k0 + every, nsteps) - k0
        for i in range(m):
            solver.rk4_step_tape(x, tapes[i], dt=dt)
        for i in reversed(range(m)):
            solver.rk4_vjp_tape(tapes[i], lam, dt)
    torch.cuda.synchronize()
    return L, lam


class AcousticRollout:
    """The acoustic-energy loss on the plane z = `plane` after `nsteps` RK4
    steps of inviscid flow (faces `nbc`, dt fixed from the CFL limit of the
    initial state), as a function of the initial density perturbation.
    curvilinear=True runs it on the sine-skewed grid of
    workloads.curvilinear_state (per-point metrics in every sweep)."""

    def __init__(self, dims, nbc, nsteps=40, every=8, plane=96, amp=1e-4, seed=5, cfl=0.5, curvilinear=False,
                 skew=1.5):
        import torch
        from . import workloads as wl
        from .fused import FusedSolver
        self.dims = tuple(dims)
        self.lmx = dims[0] * dims[1] * dims[2]
        self.nsteps, self.every, self.plane = nsteps, every, plane
        st = (wl.curvilinear_state(dims, seed=seed, skew=skew, amp=amp) if curvilinear
              else wl.flow_state(dims, seed=seed, amp=amp))
        st0 = wl.flow_state(dims, seed=seed, amp=0.0)
        self.base = torch.from_numpy(st0["qa"]).cuda()
        lmx = self.lmx
        self.rho1 = torch.from_numpy(st["qa"][:lmx] - st0["qa"][:lmx]).cuda()     # rho' of the pulse
        self.gam = wl.GAM
        self.p0, self.rho0, self.c0 = 1.0 / self.gam, 1.0, 1.0
        self.solver = FusedSolver(dims, nbc, st, viscous=False, cfl=cfl, curvilinear=curvilinear)
        self.dt = self.solver.time_step(self.initial(self.rho1))
        self.tapes = None

    def initial(self, r):
        q = self.base.clone()
        lmx = self.lmx
        q[:lmx] += r
        q[4 * lmx:] += (self.c0 * self.c0 / (self.gam - 1.0)) * r        # p' = c0^2 rho', u' = 0
        return q

    def loss_and_grad(self, q):
        import torch
        lxi, let, lze = self.dims
        qp = q.view(5, lze, let, lxi)[:, self.plane].detach().clone().requires_grad_(True)
        r, m, E = qp[0], qp[1:4], qp[4]
        pr = (self.gam - 1.0) * (E - 0.5 * (m * m).sum(0) / r)
        u2 = ((m / r) ** 2).sum(0)
        L = ((pr - self.p0) ** 2 / (2 * self.rho0 * self.c0 * self.c0) + 0.5 * self.rho0 * u2).sum()
        L.backward()
        lam = torch.zeros(5 * self.lmx, dtype=torch.float64, device="cuda")
        lam.view(5, lze, let, lxi)[:, self.plane] = qp.grad.detach()
        return float(L.detach()), lam

    def loss(self, r):
        q = self.initial(r)
        for _ in range(self.nsteps):
            self.solver.rk4_step(q, dt=self.dt)
        return self.loss_and_grad(q)[0]

    def gradient(self, r=None):
        """(L, dL/d rho') at r (default: the pulse)."""
        r = self.rho1 if r is None else r
        if self.tapes is None:
            self.tapes = [self.solver.new_tape() for _ in range(min(self.every, self.nsteps))]
        L, lam = checkpointed_vjp(self.solver, self.initial(r), self.nsteps, self.every, self.dt,
                                  self.loss_and_grad, self.tapes)
        lmx = self.lmx
        return L, lam[:lmx] + (self.c0 * self.c0 / (self.gam - 1.0)) * lam[4 * lmx:]

    def random_direction(self, seed):
        """N(0, 1) white noise over every grid point, numpy generator `seed`."""
        import torch
        return torch.from_numpy(np.random.default_rng(seed).standard_normal(self.lmx)).cuda()

    def fd_check(self, directions, eps=1e-7, grad=None, order=2):
        """Central differences against <grad, v> for each direction v; returns
        the per-direction arrays.  order 2: (L(r + e v) - L(r - e v)) / (2 e);
        order 4: (-L(r + 2e v) + 8 L(r + e v) - 8 L(r - e v) + L(r - 2e v)) / (12 e),
        whose truncation error is O(e^4) instead of O(e^2)."""
        import torch
        if grad is None:
            _, grad = self.gradient()
        if order not in (2, 4):
            raise ValueError("order must be 2 or 4")
        fds, ads = [], []
        for v in directions:
            lp = self.loss(self.rho1 + eps * v)
            lm = self.loss(self.rho1 - eps * v)
            if order == 2:
                fds.append((lp - lm) / (2 * eps))
            else:
                lpp = self.loss(self.rho1 + 2 * eps * v)
                lmm = self.loss(self.rho1 - 2 * eps * v)
                fds.append((-lpp + 8 * lp - 8 * lm + lmm) / (12 * eps))
            ads.append(float(torch.dot(grad, v)))
        return np.array(fds), np.array(ads)
