"""The fused solver on a curvilinear grid (cs_solver_set_metrics): per-point
metric rows, Jacobian and face normals in every sweep, the time step and the
RK stage, against the oracle's calc_fluxes + calc_viscous_shear_stress +
gcbc_go with the same arrays (src/cuda/kernels/physics.hpp:98-191,228-333,
src/cuda/kernels/gcbc.hpp:251-454).  Grid: x = i + s sin(2 pi j/Ny),
y = j + s sin(2 pi k/Nz), z = k + s sin(2 pi i/Nx), analytic cofactor
metrics (every off-diagonal metric term non-zero)."""
import numpy as np
import pytest

from gpu_util import case, dev, host, nrel

CL = [[10, 10, 10], [10, 10, 10]]
PER = [[45, 45, 45], [45, 45, 45]]
MIX = [[45, 10, 10], [45, 10, 10]]


def _state(dims, seed, skew=1.5):
    from compact_schemes_amd import workloads as wl
    return wl.curvilinear_state(dims, seed=seed, skew=skew, noise=1e-2, width=3.0)


def _oracle_rhs(oracle, dims, nbc, st, visc):
    lmx = dims[0] * dims[1] * dims[2]
    ctx = oracle.Context(dims, nbc)
    de, p = np.zeros(5 * lmx), np.zeros(lmx)
    ctx.rhs(int(visc), st["qa"], de, p, st["ssk"], st["xim"], st["etm"], st["zem"], st["yaco"], st["cm"],
            st["sbcc"], st["umf"])
    return de


def test_curvilinear_state_normals():
    """The helper's face normals are unit vectors along grad(xi_a), and a grid
    without skew gives the Cartesian metrics and normals exactly."""
    from compact_schemes_amd import workloads as wl
    dims = (12, 10, 8)
    st = wl.curvilinear_state(dims, skew=1.5)
    for a, c in enumerate(st["cm"]):
        c = c.reshape(2, 3, -1)
        assert np.allclose((c * c).sum(1), 1.0, atol=1e-14)
        assert np.all(c[:, a, :] > 0.5)
    flat = wl.curvilinear_state(dims, skew=0.0)
    cart = wl.flow_state(dims)
    for k in ("xim", "etm", "zem", "yaco"):
        assert np.array_equal(flat[k], cart[k])
    assert all(np.array_equal(a, b) for a, b in zip(flat["cm"], cart["cm"]))


@pytest.mark.gpu
@pytest.mark.parametrize("visc", [True, False])
@pytest.mark.parametrize("dims, nbc", [((40, 36, 32), CL), ((40, 36, 32), PER), ((48, 26, 33), MIX),
                                       ((24, 64, 40), CL)])
def test_curvilinear_rhs_matches_oracle(oracle, dims, nbc, visc):
    from compact_schemes_amd.fused import FusedSolver
    st = _state(dims, seed=12)
    lmx = dims[0] * dims[1] * dims[2]
    sv = FusedSolver(dims, nbc, st, viscous=visc, curvilinear=True)
    q = dev(st["qa"])
    case(f"curvilinear rhs dims={dims} nbc={nbc[0]} visc={visc}")
    de = host(sv.rhs(q))
    ref = _oracle_rhs(oracle, dims, nbc, st, visc)
    errs = [nrel(de[m * lmx:(m + 1) * lmx], ref[m * lmx:(m + 1) * lmx]) for m in range(5)]
    print("curvilinear rhs vs oracle:", " ".join(f"{e:.2e}" for e in errs))
    assert max(errs) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("nbc", [CL, MIX])
def test_curvilinear_rk4_step_matches_oracle(oracle, nbc):
    """One RK4 step (CFL time step, four stages, filter) against the oracle's
    step with the same metric arrays: dt to 1e-14, the change of state to 1e-10."""
    from compact_schemes_amd.fused import FusedSolver
    dims = (40, 36, 32)
    st = _state(dims, seed=13)
    lmx = dims[0] * dims[1] * dims[2]
    sv = FusedSolver(dims, nbc, st, curvilinear=True)
    q = dev(st["qa"])
    case(f"curvilinear rk4 step nbc={nbc[0]}")
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
@pytest.mark.parametrize("nbc", [CL, MIX])
def test_curvilinear_with_cartesian_arrays_is_the_cartesian_solver(nbc):
    """The per-point path fed the constant diagonal metrics and the unit normals
    e_a gives the constant-metric solver's RHS and RK4 step bit for bit (the
    off-diagonal products are exact zeros)."""
    import torch
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    dims = (32, 28, 24)
    st = wl.flow_state(dims, seed=14, noise=1e-2, width=3.0)
    cart = FusedSolver(dims, nbc, st)
    curv = FusedSolver(dims, nbc, st, curvilinear=True)
    q = dev(st["qa"])
    case(f"curvilinear = cartesian nbc={nbc[0]}")
    a = cart.rhs(q).clone()
    b = curv.rhs(q).clone()
    assert torch.equal(a, b)
    qa, qb = q.clone(), q.clone()
    assert cart.rk4_step(qa, report_dt=True) == curv.rk4_step(qb, report_dt=True)
    assert torch.equal(qa, qb)


@pytest.mark.gpu
def test_curvilinear_launch_shapes_agree(oracle):
    """Every launch shape of the y/z sweeps gives the curvilinear RHS to rounding."""
    from compact_schemes_amd.fused import FusedSolver
    dims = (32, 48, 64)
    st = _state(dims, seed=15)
    q = dev(st["qa"])
    base = host(FusedSolver(dims, CL, st, curvilinear=True).rhs(q)).copy()
    for axis in (1, 2):
        for shape in ((0, 8, 0), (1, 16, 1), (2, 16, 0)):
            sv = FusedSolver(dims, CL, st, curvilinear=True)
            sv.tune(axis, *shape)
            case(f"curvilinear launch shape axis={axis} shape={shape}")
            assert nrel(host(sv.rhs(q)), base) < 1e-13


@pytest.mark.gpu
def test_curvilinear_refusals_and_reset():
    """Partial metric arguments are refused; all-NULL metrics return a
    curvilinear solver to the constants."""
    import compact_schemes_amd as cs
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    dims = (24, 20, 18)
    st = wl.flow_state(dims, seed=16, noise=1e-2, width=3.0)
    sv = FusedSolver(dims, CL, st, viscous=False, curvilinear=True)
    q = dev(st["qa"])
    case("curvilinear refusals")
    with pytest.raises(cs.CsError):
        cs.check(cs.lib().cs_solver_set_metrics(sv.h, cs._p(q), None, None, None, None, None, None))
    cs.check(cs.lib().cs_solver_set_metrics(sv.h, None, None, None, None, None, None, None))
    ref = FusedSolver(dims, CL, st, viscous=False)
    import torch
    assert torch.equal(sv.rhs(q).clone(), ref.rhs(q).clone())


def _euler(dims, nbc, seed, filt=True):
    from compact_schemes_amd.fused import FusedSolver
    st = _state(dims, seed)
    return st, FusedSolver(dims, nbc, st, viscous=False, filt=filt, curvilinear=True)


CONFIG5 = [[45, 10, 10], [45, 10, 10]]


@pytest.mark.gpu
@pytest.mark.parametrize("nbc", [CONFIG5, CL, PER], ids=["config5", "closure", "periodic"])
def test_curvilinear_rhs_vjp_matches_finite_differences(nbc):
    """<lam, dR/dq v> by fourth-order central differences of the curvilinear
    RHS against <(dR/dq)^T lam, v> from the fused transposed sweeps, for four
    seeded directions."""
    import torch
    dims = (24, 32, 16)
    st, sv = _euler(dims, nbc, 41)
    rng = np.random.default_rng(42)
    n = st["qa"].size
    q0 = dev(st["qa"])
    lam = dev(rng.standard_normal(n))
    case(f"curvilinear rhs vjp fd nbc={nbc[0]}")
    g = sv.rhs_vjp(q0, lam).clone()

    def R(q):
        return sv.rhs(q).clone()
    eps = 1e-4
    for k in range(4):
        v = dev(rng.standard_normal(n))
        d = (-R(q0 + 2 * eps * v) + 8 * R(q0 + eps * v) - 8 * R(q0 - eps * v) + R(q0 - 2 * eps * v)) / (12 * eps)
        fd = float(torch.dot(lam, d))
        ad = float(torch.dot(g, v))
        assert abs(fd - ad) <= 1e-8 * abs(ad), (k, fd, ad)


@pytest.mark.gpu
@pytest.mark.parametrize("nbc", [CONFIG5, CL], ids=["config5", "closure"])
def test_curvilinear_rk4_vjp_and_tape(nbc):
    """The adjoint of one curvilinear RK4 step (dt fixed, filter on) against
    central differences; the recorded tape reproduces the step and the
    adjoint bit for bit."""
    import torch
    dims = (24, 16, 32)
    st, sv = _euler(dims, nbc, 43)
    rng = np.random.default_rng(44)
    n = st["qa"].size
    q0 = dev(st["qa"])
    dt = sv.time_step(q0)
    lam = dev(rng.standard_normal(n))
    v = dev(rng.standard_normal(n))

    def step(q):
        x = q.clone()
        sv.rk4_step(x, dt=dt)
        return x
    case(f"curvilinear rk4 vjp nbc={nbc[0]}")
    g = sv.rk4_vjp(q0, lam.clone(), dt).clone()
    eps = 1e-6
    fd = float(torch.dot(lam, (step(q0 + eps * v) - step(q0 - eps * v)) / (2 * eps)))
    ad = float(torch.dot(g, v))
    assert abs(fd - ad) <= 1e-7 * abs(ad), (fd, ad)
    tape = sv.new_tape()
    q = q0.clone()
    assert sv.rk4_step_tape(q, tape, dt=dt, report_dt=True) == dt
    assert torch.equal(q, step(q0))
    assert torch.equal(sv.rk4_vjp_tape(tape, lam.clone(), dt).clone(), g)


@pytest.mark.gpu
def test_curvilinear_adjoint_with_cartesian_arrays():
    """Fed the constant diagonal metrics and unit normals, the per-point
    adjoint equals the constant-metric adjoint to rounding (the general
    metric-row Jacobian sums its terms in another order)."""
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    dims = (24, 16, 32)
    st = wl.flow_state(dims, seed=45, noise=1e-2, width=3.0)
    cart = FusedSolver(dims, CONFIG5, st, viscous=False)
    curv = FusedSolver(dims, CONFIG5, st, viscous=False, curvilinear=True)
    rng = np.random.default_rng(46)
    q0 = dev(st["qa"])
    lam = dev(rng.standard_normal(st["qa"].size))
    dt = cart.time_step(q0)
    case("curvilinear adjoint = cartesian")
    assert nrel(host(curv.rhs_vjp(q0, lam)), host(cart.rhs_vjp(q0, lam))) < 1e-13
    assert nrel(host(curv.rk4_vjp(q0, lam.clone(), dt)), host(cart.rk4_vjp(q0, lam.clone(), dt))) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(24, 20, 18), (24, 16, 32)], ids=["ragged", "chunks8"])
def test_curvilinear_per_operator_adjoint(dims):
    """The per-operator adjoint (lines without 8-point chunks, and adjoint
    mode 0) with per-point metrics: the RHS product against fourth-order
    central differences, the RK4-step product against central differences,
    and on 8-point lines against the fused transposed sweeps to 1e-12."""
    import torch
    st, sv = _euler(dims, CONFIG5, 47)
    rng = np.random.default_rng(49)
    n = st["qa"].size
    q0 = dev(st["qa"])
    lam = dev(rng.standard_normal(n))
    v = dev(rng.standard_normal(n))
    dt = sv.time_step(q0)
    case(f"curvilinear per-operator adjoint dims={dims}")
    fused = None
    if dims == (24, 16, 32):
        fused = (sv.rhs_vjp(q0, lam).clone(), sv.rk4_vjp(q0, lam.clone(), dt).clone())
        sv.set_adjoint_mode(0)
    g = sv.rhs_vjp(q0, lam).clone()
    R = lambda q: sv.rhs(q).clone()
    eps = 1e-4
    d = (-R(q0 + 2 * eps * v) + 8 * R(q0 + eps * v) - 8 * R(q0 - eps * v) + R(q0 - 2 * eps * v)) / (12 * eps)
    fd, ad = float(torch.dot(lam, d)), float(torch.dot(g, v))
    assert abs(fd - ad) <= 1e-8 * abs(ad), (fd, ad)
    a = sv.rk4_vjp(q0, lam.clone(), dt).clone()

    def step(q):
        x = q.clone()
        sv.rk4_step(x, dt=dt)
        return x
    e = 1e-6
    fd = float(torch.dot(lam, (step(q0 + e * v) - step(q0 - e * v)) / (2 * e)))
    ad = float(torch.dot(a, v))
    assert abs(fd - ad) <= 1e-7 * abs(ad), (fd, ad)
    if fused is not None:
        assert nrel(host(g), host(fused[0])) < 1e-12
        assert nrel(host(a), host(fused[1])) < 1e-12

@pytest.mark.gpu
@pytest.mark.parametrize("visc", [True, False])
def test_curvilinear_from_coordinates_matches_per_operator(oracle, visc):
    """Self-contained on a curvilinear grid: the metrics come from the
    coordinates on the GPU (cs_metrics, closure rows), the face normals from
    those metrics, and the fused curvilinear RHS equals the per-operator path
    (calc_viscous_shear_stress + calc_fluxes + gcbc_go, layer 1) and the
    oracle on the same arrays."""
    import compact_schemes_amd as cs
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    from compact_schemes_amd.stepper import Stepper
    dims = (40, 32, 36)
    nx, ny, nz = dims
    lmx = nx * ny * nz
    s = 1.5
    tw = 2 * np.pi
    k, j, i = (g.reshape(-1).astype(float) for g in np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx),
                                                                   indexing="ij"))
    xyz = np.concatenate([i + s * np.sin(tw * j / ny), j + s * np.sin(tw * k / nz), k + s * np.sin(tw * i / nx)])
    case(f"curvilinear from coordinates visc={visc}")
    num = cs.Numerics(dims, CL)
    xim, etm, zem, yaco = (host(t) for t in num.metrics(dev(xyz)))
    assert np.all(yaco > 0)
    st = wl.flow_state(dims, seed=17, noise=1e-2, width=3.0)
    st.update(xim=xim, etm=etm, zem=zem, yaco=yaco, cm=wl.face_normals_from_metrics(dims, xim, etm, zem))
    q = dev(st["qa"])
    sv = FusedSolver(dims, CL, st, viscous=visc, curvilinear=True)
    de = host(sv.rhs(q))
    ref = Stepper(dims, CL, visc, st["umf"], st)
    dr = host(ref.rhs(q))
    orc = _oracle_rhs(oracle, dims, CL, st, visc)
    for m in range(5):
        sl = slice(m * lmx, (m + 1) * lmx)
        assert nrel(de[sl], dr[sl]) < 1e-12, f"per-operator, variable {m}"
        assert nrel(de[sl], orc[sl]) < 1e-12, f"oracle, variable {m}"


@pytest.mark.gpu
@pytest.mark.parametrize("nbc", [CL, MIX])
def test_curvilinear_fp32_intermediates(oracle, nbc):
    """FP32 storage of the viscous intermediates on the skewed grid: the RHS
    within 1e-9 of the FP64 oracle, and the FP64 mode unchanged by a round
    trip through the FP32 one."""
    import torch
    from compact_schemes_amd.fused import FusedSolver
    dims = (40, 36, 32)
    st = _state(dims, seed=18)
    lmx = dims[0] * dims[1] * dims[2]
    sv = FusedSolver(dims, nbc, st, curvilinear=True)
    q = dev(st["qa"])
    case(f"curvilinear fp32 intermediates nbc={nbc[0]}")
    d64 = sv.rhs(q).clone()
    sv.set_precision(1)
    d32 = host(sv.rhs(q))
    ref = _oracle_rhs(oracle, dims, nbc, st, True)
    errs = [nrel(d32[m * lmx:(m + 1) * lmx], ref[m * lmx:(m + 1) * lmx]) for m in range(5)]
    print("curvilinear fp32 intermediates vs fp64 oracle:", " ".join(f"{e:.2e}" for e in errs))
    assert max(errs) < 1e-9
    assert max(errs) > 0.0                    # the FP32 kernels ran
    sv.set_precision(0)
    assert torch.equal(sv.rhs(q).clone(), d64)


@pytest.mark.gpu
@pytest.mark.parametrize("dims, nbc, prec", [
    ((32, 40, 24), CL, 0),        # x lines of 4 chunks of 8: the closing x sweep reduces dt
    ((32, 40, 24), PER, 1),       # the same with FP32 intermediates
    ((36, 40, 24), CL, 0),        # ragged x lines: a pass of its own
    ((32, 40, 24), MIX, -1),      # x sweep tuned to the tiled kernel: a pass of its own
])
def test_curvilinear_step_time_step_reduction(dims, nbc, prec):
    """Wherever a curvilinear RK4 step reduces its time step (inside the
    closing x gradient sweep, or in a pass of its own over q, the metric rows
    and J), the step's dt equals cs_solver_time_step's bit for bit, and the
    step equals one with that dt given."""
    import torch
    from compact_schemes_amd.fused import FusedSolver
    st = _state(dims, seed=17)
    sv = FusedSolver(dims, nbc, st, curvilinear=True)
    if prec < 0:
        sv.tune(0, 1, 16, 0)     # the tiled x kernel
    else:
        sv.set_precision(prec)
    q = dev(st["qa"])
    case(f"curvilinear step dt dims={dims} nbc={nbc[0]} prec={prec}")
    dt_ts = sv.time_step(q)
    q1 = q.clone()
    dt1 = sv.rk4_step(q1, report_dt=True)
    assert dt1 == dt_ts
    q2 = q.clone()
    sv.rk4_step(q2, dt=dt1)
    assert torch.equal(q1, q2)
