"""The fused solver with the viscous intermediates stored in FP32
(cs_solver_set_precision 1): the gradients of the first two gradient sweeps,
the stress tensor and the heat flux travel between sweeps as FP32; the
conserved variables, the partial RHS, the boundary rows, the characteristic
BCs and the RK accumulation stay FP64, and every sweep computes in FP64.

These intermediates enter the RHS scaled by the viscosity (1e-4 here), so
their FP32 rounding (6e-8 relative) moves the RHS by about 1e-11 of its size.
Bars: 1e-9 on the RHS against the FP64 oracle; config 4 (512x512x256,
periodic, 50 RK4 steps with the filter) within 1e-10 of the FP64 run,
relative to the state, for every conserved variable."""
import numpy as np
import pytest

from gpu_util import case, dev, host, nrel

CL = [[10, 10, 10], [10, 10, 10]]
PER = [[45, 45, 45], [45, 45, 45]]
MIX = [[45, 10, 10], [45, 10, 10]]


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(32, 28, 24), (48, 40, 36)])
@pytest.mark.parametrize("nbc", [CL, PER, MIX], ids=["closure", "periodic", "mixed"])
def test_fp32_intermediates_rhs_matches_oracle(oracle, dims, nbc):
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    st = wl.flow_state(dims, seed=3, noise=1e-2, width=3.0)
    lmx = dims[0] * dims[1] * dims[2]
    sv = FusedSolver(dims, nbc, st)
    sv.set_precision(1)
    q = dev(st["qa"])
    case(f"fp32 intermediates rhs dims={dims} nbc={nbc[0]}")
    de = host(sv.rhs(q))
    ctx = oracle.Context(dims, nbc)
    ref, p = np.zeros(5 * lmx), np.zeros(lmx)
    ctx.rhs(1, st["qa"], ref, p, st["ssk"], st["xim"], st["etm"], st["zem"], st["yaco"], st["cm"], st["sbcc"],
            st["umf"])
    errs = [nrel(de[m * lmx:(m + 1) * lmx], ref[m * lmx:(m + 1) * lmx]) for m in range(5)]
    print("fp32 intermediates, rhs vs fp64 oracle:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < 1e-9
    sv.set_precision(0)
    de64 = host(sv.rhs(q))
    assert max(nrel(de64[m * lmx:(m + 1) * lmx], ref[m * lmx:(m + 1) * lmx]) for m in range(5)) < 1e-12


@pytest.mark.gpu
def test_fp32_intermediates_launch_shapes_agree():
    """The FP32-intermediate kernels on every y/z launch shape give the same RHS to rounding."""
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    dims = (32, 48, 64)
    lmx = dims[0] * dims[1] * dims[2]
    st = wl.flow_state(dims, seed=4, noise=1e-2, width=3.0)
    q = dev(st["qa"])
    base = FusedSolver(dims, CL, st)
    base.set_precision(1)
    d0 = host(base.rhs(q)).copy()
    for axis in (1, 2):
        for shape in [(0, 16, 0), (1, 16, 0), (2, 16, 0), (2, 8, 1)]:
            sv = FusedSolver(dims, CL, st)
            sv.set_precision(1)
            sv.tune(axis, *shape)
            case(f"fp32 intermediates axis={axis} shape={shape}")
            de = host(sv.rhs(q))
            for m in range(5):
                sl = slice(m * lmx, (m + 1) * lmx)
                assert nrel(de[sl], d0[sl]) < 1e-12, (axis, shape, m)


@pytest.mark.gpu
def test_config4_fp32_intermediates_within_1e10_of_fp64():
    """Config 4: 512x512x256, the RHS workload's flow with periodic faces, 50
    RK4 steps with the filter, dt fixed from the CFL limit of the initial state."""
    import torch
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    dims = (512, 512, 256)
    lmx = dims[0] * dims[1] * dims[2]
    st = wl.flow_state(dims, seed=4)
    q0 = torch.from_numpy(st["qa"]).cuda()
    del st["qa"]
    out = {}
    dt = None
    for prec in (0, 1):
        sv = FusedSolver(dims, PER, st)
        sv.set_precision(prec)
        if dt is None:
            dt = sv.time_step(q0)
        q = q0.clone()
        case(f"config 4: 50 steps, precision {prec}")
        for _ in range(50):
            sv.rk4_step(q, dt=dt)
        torch.cuda.synchronize()
        out[prec] = q
        sv.close()
        del sv
        torch.cuda.empty_cache()
    errs = []
    for v in range(5):
        sl = slice(v * lmx, (v + 1) * lmx)
        d = out[1][sl] - out[0][sl]
        errs.append(float(d.abs().max() / out[0][sl].abs().max()))
    print("config 4, fp32 intermediates vs fp64, max relative to the state:",
          " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert torch.isfinite(out[1]).all()
    assert max(errs) < 1e-10
