"""GPU parity of the fused RHS / RK4 path (cs_solver) against the CPU oracle.

Bar: FP64, norm-wise relative difference <= 1e-12 on the RHS; <= 1e-10 on the
change of the state over one RK4 step (the change is ~dt * RHS, so this is
the RHS bar carried through four stages and the filter).
"""
import numpy as np
import pytest

from gpu_util import case, dev, host, nrel

CL = [[10, 10, 10], [10, 10, 10]]
PER = [[45, 45, 45], [45, 45, 45]]
MIX = [[45, 10, 10], [45, 10, 10]]


def _state(dims, seed, noise=1e-2, width=3.0):
    from compact_schemes_amd import workloads as wl
    return wl.flow_state(dims, seed=seed, noise=noise, width=width)


def _oracle_rhs(oracle, dims, nbc, st, visc):
    lmx = dims[0] * dims[1] * dims[2]
    ctx = oracle.Context(dims, nbc)
    de, p = np.zeros(5 * lmx), np.zeros(lmx)
    ctx.rhs(int(visc), st["qa"], de, p, st["ssk"], st["xim"], st["etm"], st["zem"], st["yaco"], st["cm"],
            st["sbcc"], st["umf"])
    return de


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(32, 28, 24), (48, 40, 36), (100, 33, 17), (16, 64, 20)])
@pytest.mark.parametrize("nbc", [CL, PER, MIX], ids=["closure", "periodic", "mixed"])
@pytest.mark.parametrize("visc", [True, False], ids=["viscous", "euler"])
def test_fused_rhs_matches_oracle(oracle, dims, nbc, visc):
    from compact_schemes_amd.fused import FusedSolver
    st = _state(dims, seed=3)
    lmx = dims[0] * dims[1] * dims[2]
    sv = FusedSolver(dims, nbc, st, viscous=visc)
    q = dev(st["qa"])
    case(f"fused rhs dims={dims} nbc={nbc[0]} visc={visc}")
    de = host(sv.rhs(q))
    ref = _oracle_rhs(oracle, dims, nbc, st, visc)
    for m in range(5):
        assert nrel(de[m * lmx:(m + 1) * lmx], ref[m * lmx:(m + 1) * lmx]) < 1e-12, f"variable {m}"


@pytest.mark.gpu
@pytest.mark.parametrize("dims,nbc", [((32, 28, 24), CL), ((40, 24, 20), MIX), ((20, 36, 28), PER), ((100, 33, 17), CL)])
@pytest.mark.parametrize("visc", [True, False], ids=["viscous", "euler"])
def test_fused_sweep_orders_agree(oracle, dims, nbc, visc):
    """Both sweep orders (closure and stage write on z, or on x) give the
    oracle's RHS, and each other's to rounding."""
    from compact_schemes_amd.fused import FusedSolver
    st = _state(dims, seed=5)
    lmx = dims[0] * dims[1] * dims[2]
    q = dev(st["qa"])
    res = []
    for order in (0, 1):
        sv = FusedSolver(dims, nbc, st, viscous=visc)
        sv.set_order(order)
        case(f"fused rhs order={order} dims={dims} nbc={nbc[0]} visc={visc}")
        res.append(host(sv.rhs(q)).copy())
    ref = _oracle_rhs(oracle, dims, nbc, st, visc)
    for m in range(5):
        sl = slice(m * lmx, (m + 1) * lmx)
        assert nrel(res[0][sl], ref[sl]) < 1e-12, f"order 0 variable {m}"
        assert nrel(res[1][sl], ref[sl]) < 1e-12, f"order 1 variable {m}"
        assert nrel(res[0][sl], res[1][sl]) < 1e-13, f"orders differ, variable {m}"


@pytest.mark.gpu
@pytest.mark.parametrize("dims,nbc", [((32, 28, 24), CL), ((40, 24, 20), MIX), ((24, 24, 24), PER)])
def test_fused_rk4_step_matches_oracle(oracle, dims, nbc):
    from compact_schemes_amd.fused import FusedSolver
    st = _state(dims, seed=6)
    lmx = dims[0] * dims[1] * dims[2]
    sv = FusedSolver(dims, nbc, st)
    q = dev(st["qa"])
    case(f"fused rk4 dims={dims} nbc={nbc[0]}")
    dt = sv.rk4_step(q, report_dt=True)
    ctx = oracle.Context(dims, nbc)
    qo = st["qa"].copy()
    dto = ctx.rk4_step(1, qo, st["ssk"], st["xim"], st["etm"], st["zem"], st["yaco"], st["cm"], st["sbcc"],
                       st["umf"])
    assert abs(dt - dto) / dto < 1e-14
    qh = host(q)
    for v in range(5):
        sl = slice(v * lmx, (v + 1) * lmx)
        assert nrel(qh[sl] - st["qa"][sl], qo[sl] - st["qa"][sl]) < 1e-10, f"variable {v}"


@pytest.mark.gpu
def test_fused_fixed_dt_and_time_step(oracle):
    from compact_schemes_amd.fused import FusedSolver
    dims = (24, 20, 18)
    st = _state(dims, seed=7)
    sv = FusedSolver(dims, CL, st)
    q = dev(st["qa"])
    case("fused time step")
    dt = sv.time_step(q)
    lmx = dims[0] * dims[1] * dims[2]
    de, _ = oracle.primitives(st["qa"], lmx)
    ref, _ = oracle.time_step(lmx, st["xim"], st["etm"], st["zem"], de, st["yaco"], st["ssk"], st["umf"], 0.95)
    assert abs(dt - ref) / ref < 1e-14
    used = sv.rk4_step(q, dt=0.01, report_dt=True)
    assert used == 0.01


@pytest.mark.gpu
@pytest.mark.parametrize("visc", [True, False], ids=["viscous", "euler"])
@pytest.mark.parametrize("dims, nbc", [((32, 28, 24), CL), ((48, 40, 36), MIX), ((24, 64, 40), PER)])
def test_fused_concurrent_sweeps_bit_identical(dims, nbc, visc):
    """cs_solver_rhs runs the two first gradient sweeps of a viscous solver on
    two streams; the result equals the six sweeps run one after another
    (cs_solver_sweep) bit for bit."""
    import torch
    from compact_schemes_amd.fused import FusedSolver
    st = _state(dims, seed=13)
    sv = FusedSolver(dims, nbc, st, viscous=visc)
    q = dev(st["qa"])
    case(f"concurrent sweeps dims={dims} nbc={nbc[0]} visc={visc}")
    de = sv.rhs(q).clone()
    chained = torch.zeros_like(de)
    for w in (range(6) if visc else range(3, 6)):
        sv.sweep(w, q, chained)
    torch.cuda.synchronize()
    assert torch.equal(de, chained)


@pytest.mark.gpu
@pytest.mark.parametrize("dims, nbc, order, prec", [
    ((32, 64, 24), CL, 1, 0),     # y gradient sweep first, 16 lines x 8 chunks: reduces dt
    ((32, 64, 24), PER, 1, 1),    # the same with FP32 intermediates
    ((24, 48, 40), MIX, 1, 0),    # 16 lines x 6 chunks, not whole waves: a pass of its own
    ((32, 64, 24), CL, 0, 0),     # x gradient sweep first: a pass of its own
    ((32, 64, 24), MIX, 1, -1),   # y sweep tuned to the tiled kernel: a pass of its own
])
def test_fused_step_time_step_reduction(dims, nbc, order, prec):
    """Wherever an RK4 step reduces its time step (inside its first gradient
    sweep, or in a pass of its own), the step's dt equals cs_solver_time_step's
    bit for bit, and the step equals one with that dt given."""
    import torch
    from compact_schemes_amd.fused import FusedSolver
    st = _state(dims, seed=9)
    sv = FusedSolver(dims, nbc, st)
    sv.set_order(order)
    if prec < 0:
        sv.tune(1, 1, 16, 0)     # the tiled y kernel, 16 lines per workgroup
    else:
        sv.set_precision(prec)
    q = dev(st["qa"])
    case(f"fused step dt dims={dims} nbc={nbc[0]} order={order} prec={prec}")
    dt_ts = sv.time_step(q)
    q1 = q.clone()
    dt1 = sv.rk4_step(q1, report_dt=True)
    assert dt1 == dt_ts
    q2 = q.clone()
    sv.rk4_step(q2, dt=dt1)
    assert torch.equal(q1, q2)


@pytest.mark.gpu
def test_fused_rhs_256_matches_oracle(oracle):
    """The benchmark workload itself: 256^3, Mach 0.5 flow with the acoustic
    pulse, non-reflective characteristic BCs on all faces, viscous."""
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    dims = (256, 256, 256)
    lmx = 256 ** 3
    st = wl.flow_state(dims, seed=2)
    sv = FusedSolver(dims, CL, st)
    q = dev(st["qa"])
    case("fused rhs 256^3")
    de = host(sv.rhs(q))
    ref = _oracle_rhs(oracle, dims, CL, st, True)
    for m in range(5):
        assert nrel(de[m * lmx:(m + 1) * lmx], ref[m * lmx:(m + 1) * lmx]) < 1e-12, f"variable {m}"
    # uniform flow: every derivative vanishes, the RHS is zero to rounding
    st0 = wl.flow_state(dims, seed=2, amp=0.0)
    case("fused rhs 256^3 uniform")
    de0 = host(sv.rhs(dev(st0["qa"])))
    assert np.max(np.abs(de0)) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("filt", [False, True], ids=["stages", "stages+filter"])
def test_fused_rk4_step_256_matches_oracle(oracle, filt):
    """The benchmark's timed unit: one RK4 step of the 256^3 viscous workload
    (time step at CFL 0.95, four stages with in-sweep characteristic BCs and
    stage updates, the filter along x, y, z) against the oracle's step.

    Density and energy (unit-size base values) agree to 1e-13 of their size
    after the step.  The change of the state over the step (3e-5 for rho on a
    unit state; the momenta are their own change, their base is 0) agrees to
    1e-10 without the filter.  The filter's matrix (alpha_f = 0.49) has a condition number
    of about 99, so two correct solves of its rows -- the chunked recurrences
    here, Thomas in the oracle -- differ by up to cond * eps * |q|, about 2e-14
    of the unit state, i.e. about 1e-9 of rho's change: with the filter the
    change is held to 3e-9."""
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    dims = (256, 256, 256)
    lmx = 256 ** 3
    st = wl.flow_state(dims, seed=2)
    sv = FusedSolver(dims, CL, st, filt=filt)
    q = dev(st["qa"])
    case(f"fused rk4 step 256^3, filter {filt}")
    dt = sv.rk4_step(q, report_dt=True)
    qh = host(q)
    del q
    sv.close()
    ctx = oracle.Context(dims, CL)
    qo = st["qa"].copy()
    dto = ctx.rk4_step(1, qo, st["ssk"], st["xim"], st["etm"], st["zem"], st["yaco"], st["cm"], st["sbcc"],
                       st["umf"], filt=filt)
    assert abs(dt - dto) / dto < 1e-14
    bar = 3e-9 if filt else 1e-10
    for v in range(5):
        sl = slice(v * lmx, (v + 1) * lmx)
        e_state = nrel(qh[sl], qo[sl])
        e_change = nrel(qh[sl] - st["qa"][sl], qo[sl] - st["qa"][sl])
        print(f"variable {v}: state {e_state:.2e}, change {e_change:.2e}", flush=True)
        if v in (0, 4):
            assert e_state < 1e-13, f"variable {v}"
        assert e_change < bar, f"variable {v}"


@pytest.mark.gpu
def test_fused_errors():
    import compact_schemes_amd as cs
    from compact_schemes_amd.fused import FusedSolver
    dims = (24, 20, 18)
    st = _state(dims, seed=1)
    with pytest.raises(cs.CsError) as e:
        FusedSolver(dims, [[20, 10, 10], [20, 10, 10]], st)          # walls: per-operator API
    assert e.value.code == 4
    with pytest.raises(cs.CsError) as e:
        FusedSolver((24, 20, 8), CL, _state((24, 20, 8), seed=1))
    assert e.value.code == 2


@pytest.mark.gpu
@pytest.mark.parametrize("dims,nbc", [((32, 24, 40), CL), ((40, 32, 48), MIX), ((24, 32, 16), PER)])
@pytest.mark.parametrize("shape", [(1, 16, 8), (1, 8, 8), (5, 16, 8), (5, 8, 8), (1, 16, 16), (1, 8, 16), (5, 16, 16)],
                         ids=lambda s: f"fields{s[0]}-lines{s[1]}-chunk{s[2]}")
def test_fused_filter_matches_oracle(oracle, dims, nbc, shape):
    """The filter sweeps (x, y, z) against the oracle's filter of each
    variable along x, then y, then z, for each launch shape (16-point chunks
    on the y/z axes that take them); the shapes agree to rounding, and the
    three single-axis sweeps run in turn bit for bit."""
    import torch
    from compact_schemes_amd.fused import FusedSolver
    st = _state(dims, seed=8)
    lmx = dims[0] * dims[1] * dims[2]
    sv = FusedSolver(dims, nbc, st)
    q = dev(st["qa"])
    case(f"fused filter dims={dims} nbc={nbc[0]} shape={shape}")
    ref_sv = FusedSolver(dims, nbc, st)                # default launch shapes
    want = torch.zeros_like(q)
    ref_sv.sweep(7, q, want)
    fields, lines, chunk = shape
    for a in range(3):
        ok16 = a > 0 and dims[a] % 16 == 0 and 32 <= dims[a] <= 256
        sv.tune_filter(a, fields, lines, chunk if ok16 else 8)
    out = torch.zeros_like(q)
    sv.sweep(7, q, out)
    got = host(out)
    f0 = st["qa"]
    assert nrel(got - f0, host(want) - f0) < 5e-12     # the solve's FMA contraction differs by shape
    one = q.clone()
    for w in (8, 9, 10):
        sv.sweep(w, one, one)
    assert np.array_equal(host(one), got)
    ctx = oracle.Context(dims, nbc)
    for v in range(5):
        f = st["qa"][v * lmx:(v + 1) * lmx]
        ref = ctx.filter(2, ctx.filter(1, ctx.filter(0, f)))
        assert nrel(got[v * lmx:(v + 1) * lmx] - f, ref - f) < 1e-11, v


@pytest.mark.gpu
@pytest.mark.parametrize("dims,nbc", [((32, 24, 40), CL), ((40, 32, 48), MIX)])
def test_fused_rk4_step_fused_filter_matches_oracle(oracle, dims, nbc):
    test_fused_rk4_step_matches_oracle(oracle, dims, nbc)


def torch_zeros_like(t):
    import torch
    return torch.zeros_like(t)


@pytest.mark.gpu
@pytest.mark.parametrize("dims,nbc", [((32, 24, 40), CL), ((40, 32, 48), MIX)])
def test_fused_rk4_graph_replays_match_eager(dims, nbc):
    """One RK4 step (time step, stages, filter) recorded into a hipGraph
    through torch.cuda.graph: every replay advances the state exactly as an
    eager step of a second solver, with eager work between the replays."""
    import torch
    from compact_schemes_amd.fused import FusedSolver
    # smooth data and half the CFL step: grid-scale noise is not stable next
    # to non-reflective faces over several steps
    st = _state(dims, seed=9, noise=1e-4)
    sv = FusedSolver(dims, nbc, st)
    se = FusedSolver(dims, nbc, st)
    q = dev(st["qa"])
    qe = q.clone()
    dt = 0.5 * sv.time_step(q)
    gs = torch.cuda.Stream()
    case(f"rk4 graph capture dims={dims} nbc={nbc[0]}")
    sv.rk4_step(q.clone(), dt=dt, stream=gs)         # one eager step on the capture stream first
    gs.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=gs):
        sv.rk4_step(q, dt=dt, stream=gs)
    for rep in range(3):
        case(f"rk4 graph replay {rep}")
        graph.replay()
        se.rk4_step(qe, dt=dt)
        torch.cuda.synchronize()
        junk = torch.randn(1000, device="cuda").sum()
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        assert torch.isfinite(qe).all(), rep
        assert torch.equal(q, qe), rep


@pytest.mark.gpu
@pytest.mark.parametrize("nbc", [CL, PER], ids=["closure", "periodic"])
@pytest.mark.parametrize("axis", [1, 2])
@pytest.mark.parametrize("shape", [(0, 16, 0), (0, 8, 1), (1, 16, 0), (2, 16, 0), (2, 8, 1)],
                         ids=["direct16", "direct8-xcd", "tile16", "chunks16", "chunks16-8-xcd"])
def test_fused_launch_shapes_agree(oracle, nbc, axis, shape):
    """Every launch shape of the y/z sweeps (cs_solver_tune: direct kernel,
    row-block tiles, 16-point chunks; 8 or 16 lines per workgroup; XCD
    grouping) gives the oracle's RHS and the default shape's to rounding."""
    from compact_schemes_amd.fused import FusedSolver
    dims = (32, 48, 64)
    lmx = dims[0] * dims[1] * dims[2]
    st = _state(dims, seed=11)
    q = dev(st["qa"])
    ref = _oracle_rhs(oracle, dims, nbc, st, True)
    base = FusedSolver(dims, nbc, st)
    d0 = host(base.rhs(q)).copy()
    sv = FusedSolver(dims, nbc, st)
    sv.tune(axis, *shape)
    case(f"fused launch shape axis={axis} shape={shape} nbc={nbc[0]}")
    de = host(sv.rhs(q))
    for m in range(5):
        sl = slice(m * lmx, (m + 1) * lmx)
        assert nrel(de[sl], ref[sl]) < 1e-12, f"variable {m}"
        assert nrel(de[sl], d0[sl]) < 1e-13, f"variable {m}"
