"""GPU parity of the fused solver on axes longer than 256 points.

x lines of 257-512 points that split into 8-point chunks run the transposing
x kernel with one line per wave (up to 64 chunks) and the fused x filter;
other long axes run the 16-point-chunk kernels: the exact form (N = 16 P)
with the fused y/z filter, and the ragged form (any N) with the line-solver
filter of an RK4 step.  One long axis at a time, on x, y and z in turn, with
the other two axes short so the oracle stays fast.
Bars as for the other fused tests: 1e-12 norm-wise on the RHS, 1e-10 on the
change of the state over one RK4 step.
"""
import numpy as np
import pytest

from gpu_util import case, dev, host, nrel

CL = [[10, 10, 10], [10, 10, 10]]
PER = [[45, 45, 45], [45, 45, 45]]
MIX = [[45, 10, 10], [45, 10, 10]]

LONG = [(272, 16, 16), (16, 272, 16), (16, 16, 272),      # x: 34 chunks of 8; y, z: 17 chunks of 16
        (264, 20, 16), (20, 264, 16), (16, 20, 264),      # x: 33 chunks of 8; y, z: ragged
        (512, 16, 16), (16, 512, 16), (16, 16, 512),      # x: 64 chunks of 8; y, z: 32 chunks of 16
        (270, 20, 16)]                                    # x: ragged 16-point chunks


def _state(dims, seed, noise=1e-2, width=3.0):
    from compact_schemes_amd import workloads as wl
    return wl.flow_state(dims, seed=seed, noise=noise, width=width)


@pytest.mark.gpu
@pytest.mark.parametrize("dims", LONG, ids=lambda d: "x".join(map(str, d)))
@pytest.mark.parametrize("nbc", [CL, PER], ids=["closure", "periodic"])
@pytest.mark.parametrize("visc", [True, False], ids=["viscous", "euler"])
def test_fused_rhs_long_axis(oracle, dims, nbc, visc):
    from compact_schemes_amd.fused import FusedSolver
    st = _state(dims, seed=11)
    lmx = dims[0] * dims[1] * dims[2]
    sv = FusedSolver(dims, nbc, st, viscous=visc)
    q = dev(st["qa"])
    case(f"fused rhs long axis dims={dims} nbc={nbc[0]} visc={visc}")
    de = host(sv.rhs(q))
    ctx = oracle.Context(dims, nbc)
    ref, p = np.zeros(5 * lmx), np.zeros(lmx)
    ctx.rhs(int(visc), st["qa"], ref, p, st["ssk"], st["xim"], st["etm"], st["zem"], st["yaco"], st["cm"],
            st["sbcc"], st["umf"])
    for m in range(5):
        assert nrel(de[m * lmx:(m + 1) * lmx], ref[m * lmx:(m + 1) * lmx]) < 1e-12, f"variable {m}"


@pytest.mark.gpu
@pytest.mark.parametrize("dims", LONG, ids=lambda d: "x".join(map(str, d)))
@pytest.mark.parametrize("nbc", [CL, MIX], ids=["closure", "mixed"])
def test_fused_rk4_step_long_axis(oracle, dims, nbc):
    from compact_schemes_amd.fused import FusedSolver
    st = _state(dims, seed=12)
    lmx = dims[0] * dims[1] * dims[2]
    sv = FusedSolver(dims, nbc, st)
    q = dev(st["qa"])
    case(f"fused rk4 long axis dims={dims} nbc={nbc[0]}")
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
