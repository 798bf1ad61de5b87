"""GPU parity of the compressible right-hand side (HIP library vs CPU oracle).

Bar: FP64, norm-wise relative difference <= 1e-12 for every output
(pointwise kernels are the same arithmetic; the line solves differ in
algorithm, so differences are at rounding level).
"""
import numpy as np
import pytest

from gpu_util import case, dev, host, nrel

TOL = 1e-12
NBC_CL = [[10, 10, 10], [10, 10, 10]]


def _state(dims, seed=4, noise=1e-2):
    from compact_schemes_amd import workloads as wl
    return wl.flow_state(dims, seed=seed, noise=noise, width=3.0)


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(32, 28, 24), (40, 36, 20)])
@pytest.mark.parametrize("viscous", [True, False])
def test_rhs_stages_match_oracle(oracle, dims, viscous):
    """primitives -> calc_viscous_shear_stress -> calc_fluxes -> gcbc_go, each
    stage's outputs compared (physics_pc.hpp:83-227, gcbc.hpp:86-186)."""
    import compact_schemes_amd as cs
    lmx = dims[0] * dims[1] * dims[2]
    st = _state(dims)
    umf = st["umf"]
    ctx = oracle.Context(dims, NBC_CL)
    num = cs.Numerics(dims, NBC_CL)
    ph = cs.Physics(dims, viscous, umf)
    q = dev(st["qa"])
    de, p = dev(np.zeros(5 * lmx)), dev(np.zeros(lmx))
    g = {k: dev(st[k]) for k in ("xim", "etm", "zem", "yaco", "ssk", "sbcc")}
    cm = [dev(c) for c in st["cm"]]
    de_o, p_o = oracle.primitives(st["qa"], lmx)
    case(f"primitives dims={dims}")
    cs.primitives(dims, q, de, p)
    assert nrel(host(de), de_o) < 1e-14 and nrel(host(p), p_o) < 1e-14
    if viscous:
        case(f"viscous dims={dims}")
        ph.calc_viscous_shear_stress(num, de, g["ssk"], g["xim"], g["etm"], g["zem"], g["yaco"])
        ctx.calc_viscous(de_o, st["ssk"], st["xim"], st["etm"], st["zem"], st["yaco"])
        S, H = ph.buffers()
        assert nrel(host(S), ctx.stress()) < TOL
        assert nrel(host(H), ctx.heat()) < TOL
        assert nrel(host(de), de_o) < TOL
    case(f"fluxes dims={dims} viscous={viscous}")
    ph.calc_fluxes(num, q, p, de, g["xim"], g["etm"], g["zem"])
    ctx.calc_fluxes(int(viscous), st["qa"], p_o, de_o, st["xim"], st["etm"], st["zem"], umf)
    for m in range(5):
        assert nrel(host(de)[m * lmx:(m + 1) * lmx], de_o[m * lmx:(m + 1) * lmx]) < TOL
    for a in range(3):
        assert nrel(host(num.buffers(a)[0]), ctx.drva(a)) < TOL
    case(f"gcbc dims={dims}")
    cs.gcbc_go(num, cm, q, de, p, g["yaco"], g["sbcc"], umf, (0, 0, 0), 0.0)
    ctx.gcbc(st["cm"], st["qa"], de_o, p_o, st["yaco"], st["sbcc"], umf, (0, 0, 0), 0.0)
    for m in range(5):
        assert nrel(host(de)[m * lmx:(m + 1) * lmx], de_o[m * lmx:(m + 1) * lmx]) < TOL


@pytest.mark.gpu
def test_time_step_matches_oracle(oracle):
    """calc_time_step on the reference test's shape and distributions
    (tests/test_physics.cu:148-181)."""
    import compact_schemes_amd as cs
    dims = (256, 128, 64)
    n = 256 * 128 * 64
    rng = np.random.default_rng(13)
    xim, etm, zem = (rng.uniform(0, 1, 3 * n) for _ in range(3))
    de = rng.uniform(0, 1, 5 * n)
    yaco, ssk = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    umf = (0.3, 0.1, 0.1)
    ph = cs.Physics(dims, False, umf)
    case("time step 256x128x64")
    dte, m2 = ph.calc_time_step(dev(xim), dev(etm), dev(zem), dev(de), dev(yaco), dev(ssk), 0.95)
    ref, mo = oracle.time_step(n, xim, etm, zem, de, yaco, ssk, umf, 0.95)
    assert abs(dte - ref) / ref < 1e-15
    assert m2[0] == mo[0] and m2[1] == mo[1]          # a max is exact


@pytest.mark.gpu
def test_rk_stage_matches_oracle(oracle):
    import compact_schemes_amd as cs
    dims = (17, 13, 11)
    n = 17 * 13 * 11
    rng = np.random.default_rng(9)
    q0, de, q = rng.standard_normal(5 * n), rng.standard_normal(5 * n), np.zeros(5 * n)
    yaco = rng.uniform(0.5, 2, n)
    qd = dev(q)
    case("rk stage")
    cs.rk_stage(dims, qd, dev(q0), dev(de), dev(yaco), 0.0371)
    oracle.rk_stage(n, q, q0, de, yaco, 0.0371)
    assert nrel(host(qd), q) < 1e-15


@pytest.mark.gpu
@pytest.mark.parametrize("dims,nbc", [((32, 28, 24), NBC_CL), ((32, 24, 20), [[45, 10, 10], [45, 10, 10]])])
def test_rk4_step_matches_oracle(oracle, dims, nbc):
    """One RK4 step with time step, four stages and the filter, against the
    oracle's orc_rk4_step (the reference driver order, canard.cu:132-205)."""
    from compact_schemes_amd.stepper import Stepper
    st = _state(dims, seed=6)
    ctx = oracle.Context(dims, nbc)
    stp = Stepper(dims, nbc, True, st["umf"], st)
    q = dev(st["qa"])
    case(f"rk4 step dims={dims} nbc={nbc}")
    dt = stp.rk4_step(q)
    qo = st["qa"].copy()
    dto = ctx.rk4_step(1, qo, st["ssk"], st["xim"], st["etm"], st["zem"], st["yaco"], st["cm"], st["sbcc"],
                       st["umf"])
    assert abs(dt - dto) / dto < 1e-14
    lmx = dims[0] * dims[1] * dims[2]
    qh = host(q)
    for v in range(5):
        # compare the change over the step, the part the RHS produces
        d_gpu = qh[v * lmx:(v + 1) * lmx] - st["qa"][v * lmx:(v + 1) * lmx]
        d_ref = qo[v * lmx:(v + 1) * lmx] - st["qa"][v * lmx:(v + 1) * lmx]
        assert nrel(d_gpu, d_ref) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("nbc", [[[45, 45, 45], [45, 45, 45]], [[10, 10, 10], [10, 10, 10]]], ids=["periodic", "closure"])
def test_rk4_step_mixed_precision_matches_oracle(oracle, nbc):
    """The RK4 step with FP32 line solves + FP64-residual refinement against
    the FP64 oracle: the change of the state over the step within 1e-10."""
    from compact_schemes_amd.stepper import Stepper
    dims = (40, 36, 32)
    st = _state(dims, seed=9)
    ctx = oracle.Context(dims, nbc)
    stp = Stepper(dims, nbc, True, st["umf"], st)
    stp.num.set_precision(1)
    q = dev(st["qa"])
    case(f"mixed rk4 step nbc={nbc}")
    dt = stp.rk4_step(q)
    qo = st["qa"].copy()
    ctx.rk4_step(1, qo, st["ssk"], st["xim"], st["etm"], st["zem"], st["yaco"], st["cm"], st["sbcc"], st["umf"])
    lmx = dims[0] * dims[1] * dims[2]
    qh = host(q)
    assert dt > 0
    for v in range(5):
        d_gpu = qh[v * lmx:(v + 1) * lmx] - st["qa"][v * lmx:(v + 1) * lmx]
        d_ref = qo[v * lmx:(v + 1) * lmx] - st["qa"][v * lmx:(v + 1) * lmx]
        assert nrel(d_gpu, d_ref) < 1e-10, v


@pytest.mark.gpu
def test_walls_match_oracle(oracle):
    """Wall faces: gcbc setup from de (flag 0) and the wall update, then
    wall_condition_update (gcbc.hpp:188-228)."""
    import compact_schemes_amd as cs
    dims = (24, 20, 16)
    lmx = 24 * 20 * 16
    nbc = [[20, 25, 10], [25, 20, 10]]
    st = _state(dims, seed=8)
    umf = st["umf"]
    ctx = oracle.Context(dims, nbc)
    num = cs.Numerics(dims, nbc)
    ph = cs.Physics(dims, False, umf)
    q = dev(st["qa"])
    de, p = dev(np.zeros(5 * lmx)), dev(np.zeros(lmx))
    g = {k: dev(st[k]) for k in ("xim", "etm", "zem", "yaco", "sbcc")}
    cs.primitives(dims, q, de, p)
    de_o, p_o = oracle.primitives(st["qa"], lmx)
    case("walls: fluxes")
    ph.calc_fluxes(num, q, p, de, g["xim"], g["etm"], g["zem"])
    ctx.calc_fluxes(0, st["qa"], p_o, de_o, st["xim"], st["etm"], st["zem"], umf)
    case("walls: gcbc")
    cs.gcbc_go(num, [dev(c) for c in st["cm"]], q, de, p, g["yaco"], g["sbcc"], umf, (0.01, -0.02, 0.0), 7.0)
    ctx.gcbc(st["cm"], st["qa"], de_o, p_o, st["yaco"], st["sbcc"], umf, (0.01, -0.02, 0.0), 7.0)
    assert nrel(host(de), de_o) < TOL
    rng = np.random.default_rng(1)
    npex = rng.integers(0, 2, lmx).astype(np.int32)
    import torch
    npex_d = torch.from_numpy(npex).cuda()
    qo = st["qa"].copy()
    case("walls: wall_condition_update")
    cs.wall_condition_update(num, q, npex_d, umf)
    ctx.wall_condition_update(qo, npex, umf)
    assert nrel(host(q), qo) < 1e-14
