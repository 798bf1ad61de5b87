"""Adjoint of the inviscid fused right-hand side and RK4 step
(cs_solver_rhs_vjp, cs_solver_rk4_vjp) against central finite differences,
and a torch.autograd.Function over the RK4 step (test-only) through
torch.autograd.gradcheck."""
import numpy as np
import pytest
import torch

from gpu_util import case, dev

CONFIG5 = [[45, 10, 10], [45, 10, 10]]     # periodic x, characteristic y and z
CL = [[10, 10, 10], [10, 10, 10]]
PER = [[45, 45, 45], [45, 45, 45]]


def _setup(dims, nbc, seed):
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    st = wl.flow_state(dims, seed=seed, noise=1e-2, width=3.0)
    sv = FusedSolver(dims, nbc, st, viscous=False)
    return st, sv


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(24, 20, 28), (32, 16, 20)])
@pytest.mark.parametrize("nbc", [CONFIG5, CL, PER], ids=["config5", "closure", "periodic"])
def test_rhs_vjp_matches_finite_differences(dims, nbc):
    st, sv = _setup(dims, nbc, 21)
    rng = np.random.default_rng(5)
    n = st["qa"].size
    q = dev(st["qa"])
    lam = dev(rng.standard_normal(n))
    v = dev(rng.standard_normal(n))
    case(f"rhs vjp dims={dims} nbc={nbc}")
    g = sv.rhs_vjp(q, lam).clone()
    eps = 1e-6
    rp = sv.rhs(q + eps * v).clone()
    rm = sv.rhs(q - eps * v).clone()
    fd = float(torch.dot(lam, (rp - rm) / (2 * eps)))
    ad = float(torch.dot(g, v))
    assert abs(fd - ad) <= 1e-7 * max(abs(ad), float(g.norm() * v.norm()) * 1e-3), (fd, ad)


@pytest.mark.gpu
@pytest.mark.parametrize("nbc", [CONFIG5, CL], ids=["config5", "closure"])
def test_rk4_vjp_matches_finite_differences(nbc):
    dims = (20, 24, 18)
    st, sv = _setup(dims, nbc, 22)
    rng = np.random.default_rng(6)
    n = st["qa"].size
    q0 = dev(st["qa"])
    dt = sv.time_step(q0)
    lam = dev(rng.standard_normal(n))
    v = dev(rng.standard_normal(n))

    def step(q):
        x = q.clone()
        sv.rk4_step(x, dt=dt)
        return x
    case(f"rk4 vjp nbc={nbc}")
    g = sv.rk4_vjp(q0, lam.clone(), dt)
    eps = 1e-6
    fd = float(torch.dot(lam, (step(q0 + eps * v) - step(q0 - eps * v)) / (2 * eps)))
    ad = float(torch.dot(g, v))
    assert abs(fd - ad) <= 1e-7 * abs(ad), (fd, ad)


class RK4Step(torch.autograd.Function):
    """One RK4 step of a FusedSolver with fixed dt as an autograd node (test-only)."""

    @staticmethod
    def forward(ctx, q, sv, dt):
        ctx.sv, ctx.dt = sv, dt
        ctx.save_for_backward(q)
        out = q.detach().clone()
        sv.rk4_step(out, dt=dt)
        return out

    @staticmethod
    def backward(ctx, grad):
        (q,) = ctx.saved_tensors
        lam = grad.detach().clone().contiguous()
        ctx.sv.rk4_vjp(q.detach().contiguous(), lam, ctx.dt)
        return lam, None, None


@pytest.mark.gpu
def test_gradcheck_rk4_step():
    dims = (16, 16, 16)
    st, sv = _setup(dims, CONFIG5, 23)
    q0 = dev(st["qa"])
    dt = sv.time_step(q0)
    q = q0.clone().requires_grad_(True)
    case("gradcheck rk4 step 16^3")
    assert torch.autograd.gradcheck(lambda x: RK4Step.apply(x, sv, dt), (q,), eps=1e-6, atol=1e-8, rtol=1e-5,
                                    fast_mode=True, nondet_tol=0.0)
