"""Initial-state and boundary-weight kernels (cs_init_pulse,
cs_boundary_weights), and the fused solver on Cartesian grids whose
coordinate decreases along an axis (the z axis of the reference's block grid,
src/grid.hpp:103), against the oracle."""
import numpy as np
import pytest

from gpu_util import case, dev, host, nrel

CL = [[10, 10, 10], [10, 10, 10]]
MIX = [[45, 10, 10], [45, 10, 10]]


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(32, 24, 20), (17, 40, 9)])
def test_init_pulse_matches_workload_state(dims):
    """The same state as workloads.flow_state (amp 1e-3, width 8, its seeded centre)."""
    import torch
    import compact_schemes_amd as cs
    from compact_schemes_amd import workloads as wl
    st = wl.flow_state(dims, seed=2)
    lmx = dims[0] * dims[1] * dims[2]
    q = torch.full((5 * lmx,), float("nan"), dtype=torch.float64, device="cuda")
    case(f"init pulse {dims}")
    cs.init_pulse(dims, st["center"], 1e-3, 8.0, q)
    qh = host(q)
    for v in range(5):
        sl = slice(v * lmx, (v + 1) * lmx)
        ref = st["qa"][sl]
        if v in (1, 2, 3):
            assert np.all(qh[sl] == 0.0)
        else:
            # exp on the device and in numpy may differ in the last bit
            assert np.max(np.abs(qh[sl] - ref)) <= 4e-16 * np.max(np.abs(ref)), v
    sb = torch.full((lmx,), float("nan"), dtype=torch.float64, device="cuda")
    cs.boundary_weights(dims, sb)
    assert np.array_equal(host(sb), wl.boundary_weights(dims))


@pytest.mark.gpu
def test_init_errors():
    import torch
    import compact_schemes_amd as cs
    q = torch.zeros(5 * 8, dtype=torch.float64, device="cuda")
    with pytest.raises(cs.CsError, match="CS_ERR_ARG"):
        cs.init_pulse((2, 2, 2), (0, 0, 0), 1e-3, 0.0, q)
    with pytest.raises(cs.CsError, match="CS_ERR_SHAPE"):
        cs.boundary_weights((0, 2, 2), q)


SPACINGS = [(1.2, 0.8, -0.5), (-0.7, 1.1, 0.9)]


@pytest.mark.gpu
@pytest.mark.parametrize("spacing", SPACINGS, ids=["z-decreasing", "x-decreasing"])
@pytest.mark.parametrize("nbc", [CL, MIX], ids=["closure", "mixed"])
@pytest.mark.parametrize("visc", [True, False], ids=["viscous", "euler"])
def test_fused_rhs_on_spaced_grid_matches_oracle(oracle, spacing, nbc, visc):
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    dims = (32, 28, 24)
    lmx = dims[0] * dims[1] * dims[2]
    st = wl.with_spacing(wl.flow_state(dims, seed=7, noise=1e-2, width=3.0), dims, spacing)
    sv = FusedSolver(dims, nbc, st, viscous=visc)
    q = dev(st["qa"])
    case(f"fused rhs spacing={spacing} nbc={nbc[0]} visc={visc}")
    de = host(sv.rhs(q))
    ctx = oracle.Context(dims, nbc)
    ref, p = np.zeros(5 * lmx), np.zeros(lmx)
    ctx.rhs(int(visc), st["qa"], ref, p, st["ssk"], st["xim"], st["etm"], st["zem"], st["yaco"], st["cm"],
            st["sbcc"], st["umf"])
    for m in range(5):
        sl = slice(m * lmx, (m + 1) * lmx)
        assert nrel(de[sl], ref[sl]) < 1e-12, f"variable {m}"


@pytest.mark.gpu
@pytest.mark.parametrize("spacing", SPACINGS, ids=["z-decreasing", "x-decreasing"])
def test_fused_rk4_step_on_spaced_grid_matches_oracle(oracle, spacing):
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    dims = (32, 28, 24)
    lmx = dims[0] * dims[1] * dims[2]
    st = wl.with_spacing(wl.flow_state(dims, seed=8, noise=1e-2, width=3.0), dims, spacing)
    sv = FusedSolver(dims, CL, st)
    q = dev(st["qa"])
    case(f"fused rk4 spacing={spacing}")
    dt = sv.rk4_step(q, report_dt=True)
    qh = host(q)
    ctx = oracle.Context(dims, CL)
    qo = st["qa"].copy()
    dto = ctx.rk4_step(1, qo, st["ssk"], st["xim"], st["etm"], st["zem"], st["yaco"], st["cm"], st["sbcc"],
                       st["umf"])
    assert abs(dt - dto) / dto < 1e-14
    for v in range(5):
        sl = slice(v * lmx, (v + 1) * lmx)
        assert nrel(qh[sl] - st["qa"][sl], qo[sl] - st["qa"][sl]) < 1e-10, f"variable {v}"


@pytest.mark.gpu
@pytest.mark.parametrize("spacing", SPACINGS, ids=["z-decreasing", "x-decreasing"])
def test_rhs_vjp_on_spaced_grid_matches_finite_differences(spacing):
    import torch
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    dims = (24, 20, 28)
    st = wl.with_spacing(wl.flow_state(dims, seed=21, noise=1e-2, width=3.0), dims, spacing)
    sv = FusedSolver(dims, [[45, 10, 10], [45, 10, 10]], st, viscous=False)
    rng = np.random.default_rng(5)
    n = st["qa"].size
    q = dev(st["qa"])
    lam = dev(rng.standard_normal(n))
    v = dev(rng.standard_normal(n))
    case(f"rhs vjp spacing={spacing}")
    g = sv.rhs_vjp(q, lam).clone()
    eps = 1e-6
    rp = sv.rhs(q + eps * v).clone()
    rm = sv.rhs(q - eps * v).clone()
    fd = float(torch.dot(lam, (rp - rm) / (2 * eps)))
    ad = float(torch.dot(g, v))
    assert abs(fd - ad) <= 1e-7 * max(abs(ad), float(g.norm() * v.norm()) * 1e-3), (fd, ad)
