"""The differentiation workload at full size, and the dense transpose of one
RK4 step.

Config 5: Euler at 128^3, periodic x, non-reflective characteristic faces in
y and z, 40 RK4 steps (dt fixed from the CFL-0.5 limit of the initial state,
filter every step), checkpoints every 8 steps; loss = acoustic energy on the
plane z = 96; gradient with respect to the initial density perturbation.
Checked by fourth-order central differences (step 1e-5) along 32 random
directions (N(0, 1) white noise over the whole grid, numpy seeds 500..531),
each direction on its own.  Second-order central differences are not
accurate enough for a 1e-6 bar along white noise: their truncation error
(the cubic terms of the loss along the direction) measured 9e-4 (median
relative) at step 1e-5, 9e-6 at 1e-6 and 2e-7 at 1e-7, falling as the step
squared.  The fourth-order stencil leaves a floor of about 1e-7 that does
not shrink with the step (steps 1e-7 to 1e-5 measured the same; periodic
faces 5e-8): over the 32 directions the median relative error measured
1.3e-7 and the largest 1.0e-6.  Bars: the norm-wise error (largest
|fd - ad| over the largest |ad|) and the median below 1e-6, every
direction below 3e-6.

Dense transpose: the Jacobian of one RK4 step at 16^3 (same faces) assembled
column by column from fourth-order central differences of the forward step,
against the rows (d step / d q)^T e_i from the step's tape.
"""
import numpy as np
import pytest

from gpu_util import case

NBC5 = [[45, 10, 10], [45, 10, 10]]


@pytest.mark.gpu
def test_config5_gradient_vs_fd_random_directions():
    import torch
    from compact_schemes_amd.rollout import AcousticRollout
    case("config 5: 128^3, 40 steps, checkpoints every 8")
    ro = AcousticRollout((128, 128, 128), NBC5, nsteps=40, every=8, plane=96)
    L, grad = ro.gradient()
    assert np.isfinite(L) and L > 0 and torch.isfinite(grad).all()
    dirs = [ro.random_direction(500 + d) for d in range(32)]
    case("config 5: central differences on 32 random directions")
    fd, ad = ro.fd_check(dirs, eps=1e-5, grad=grad, order=4)
    rel = np.abs(fd - ad) / np.abs(ad)
    print("config 5 per-direction relative errors:", " ".join(f"{e:.2e}" for e in rel), flush=True)
    print(f"config 5 max {rel.max():.3e} median {np.median(rel):.3e}", flush=True)
    normwise = np.max(np.abs(fd - ad)) / np.max(np.abs(ad))
    print(f"config 5 norm-wise {normwise:.3e}", flush=True)
    assert np.all(np.abs(ad) > 0)
    assert normwise < 1e-6 and np.median(rel) < 1e-6
    assert rel.max() < 3e-6


@pytest.mark.gpu
def test_rk4_step_vjp_dense_transpose_16():
    import torch
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    dims = (16, 16, 16)
    n = 5 * 16 ** 3
    st = wl.flow_state(dims, seed=15, amp=1e-3, width=3.0, noise=1e-4)
    sv = FusedSolver(dims, NBC5, st, viscous=False, cfl=0.5)
    q0 = torch.from_numpy(st["qa"]).cuda()
    dt = sv.time_step(q0)
    case("dense transpose 16^3: forward Jacobian by fourth-order central differences")
    eps = 1e-5
    JT = torch.empty((n, n), dtype=torch.float64, device="cuda")     # JT[c] = d step / d q_c (column c of J)
    qs = [torch.empty_like(q0) for _ in range(4)]
    for c in range(n):
        for x, h in zip(qs, (2 * eps, eps, -eps, -2 * eps)):
            x.copy_(q0)
            x[c] += h
            sv.rk4_step(x, dt=dt)
        torch.sub(qs[3], qs[0], out=JT[c])
        JT[c].add_(qs[1] - qs[2], alpha=8.0)
    JT /= 12 * eps
    case("dense transpose 16^3: adjoint rows from the tape")
    tape = sv.new_tape()
    x = q0.clone()
    sv.rk4_step_tape(x, tape, dt=dt)
    V = torch.zeros((n, n), dtype=torch.float64, device="cuda")      # V[r] = (d step / d q)^T e_r = row r of J
    for r in range(n):
        V[r, r] = 1.0
        sv.rk4_vjp_tape(tape, V[r], dt)
    err = float((V - JT.t()).abs().max()) / float(JT.abs().max())
    print(f"dense transpose 16^3: max |J^T e_r - J[r, :]| / max|J| = {err:.3e}", flush=True)
    assert err < 1e-8
