"""GPU parity of the compact derivative and filter (HIP library vs CPU oracle).

Bar: FP64, norm-wise relative difference max|x - x_ref| / max|x_ref| <= 1e-12
(the oracle solves with the reference's PCR where the line length is a power
of two, Thomas otherwise; the HIP path with chunked LU recurrences).
"""
import numpy as np
import pytest

from gpu_util import case, dev, host, nrel

TOL = 1e-12
NBC_CL = [[10, 10, 10], [10, 10, 10]]
NBC_PER = [[45, 45, 45], [45, 45, 45]]
SHAPES = [(64, 64, 64), (48, 40, 36), (256, 16, 12), (16, 256, 12), (12, 16, 256), (100, 33, 17),
          (300, 9, 8), (9, 300, 8), (8, 9, 300), (512, 8, 8),
          (700, 8, 8), (8, 1000, 8), (8, 9, 700)]   # up to 63 chunks per line (16 lines per workgroup)


def _field(dims, seed):
    rng = np.random.default_rng(seed)
    lxi, let, lze = dims
    k, j, i = np.meshgrid(np.arange(lze), np.arange(let), np.arange(lxi), indexing="ij")
    f = np.sin(2 * np.pi * (3 * i / lxi + 2 * j / let + 5 * k / lze) + 0.3) + 0.1 * rng.standard_normal(i.shape)
    return f.reshape(-1)


@pytest.mark.gpu
@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("nbc", [NBC_CL, NBC_PER], ids=["closure", "periodic"])
def test_deriv_matches_oracle(oracle, dims, nbc):
    import compact_schemes_amd as cs
    f = _field(dims, 1)
    ctx = oracle.Context(dims, nbc)
    num = cs.Numerics(dims, nbc)
    fin = dev(f)
    for axis in range(3):
        case(f"deriv dims={dims} nbc={nbc[0][0]} axis={axis}")
        out = dev(np.zeros_like(f))
        num.deriv(axis, fin, out, var=3)
        ref = ctx.deriv_field(axis, f, var=3)
        assert nrel(host(out), ref) < TOL
        drva = host(num.buffers(axis)[0])[:, 3]
        assert nrel(drva, ctx.drva(axis)[:, 3]) < TOL


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(64, 32, 16), (40, 36, 20)])
def test_reference_api_fill_mpigo_deriv1d(oracle, dims):
    """fill_buffers -> mpigo -> deriv1d (the reference's calling sequence,
    physics_pc.hpp:99-120) gives the oracle's send/recv and derivative."""
    import compact_schemes_amd as cs
    lmx = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(2)
    f3 = rng.standard_normal(5 * 3 * lmx)          # [5 var][3 comp][lmx]
    num = cs.Numerics(dims, NBC_PER)
    ctx = oracle.Context(dims, NBC_PER)
    fin = dev(f3)
    case(f"fill_buffers dims={dims}")
    num.fill_buffers(fin, 0, 5, 0)                 # nrt = 0: component = axis
    for m in range(5):
        num.mpigo(m)
        for axis in range(3):
            comp = f3[(3 * m + axis) * lmx:(3 * m + axis + 1) * lmx].copy()
            ctx.fill(axis, comp, m)
        ctx.mpigo(m)
    for axis in range(3):
        _, send, recv = num.buffers(axis)
        assert nrel(host(send), ctx.send(axis)) < 1e-14
        assert nrel(host(recv), ctx.recv(axis)) < 1e-14
    out = dev(np.zeros(5 * 3 * lmx))
    for m in range(5):
        for axis in range(3):
            case(f"deriv2d m={m} axis={axis}")
            num.deriv2d(axis, fin[m * 3 * lmx:(m + 1) * 3 * lmx], out[m * 3 * lmx:(m + 1) * 3 * lmx], 1, 1, m, axis)
            comp = f3[(3 * m + axis) * lmx:(3 * m + axis + 1) * lmx].copy()
            ref = np.zeros(lmx)
            ctx.deriv(axis, comp, ref, 1, 1, m)
            got = host(out)[(3 * m + axis) * lmx:(3 * m + axis + 1) * lmx]
            assert nrel(got, ref) < TOL


@pytest.mark.gpu
def test_kat_ramp_256x128x64():
    """tests/test_numerics_serial.cu:41-146 on the GPU: ramp -> constant stride."""
    import compact_schemes_amd as cs
    dims = (256, 128, 64)
    lmx = 256 * 128 * 64
    num = cs.Numerics(dims, NBC_CL)
    f = dev(np.arange(lmx, dtype=np.float64))
    for axis, sol in enumerate((1.0, 256.0, 256.0 * 128)):
        case(f"kat ramp axis={axis}")
        out = dev(np.zeros(lmx))
        for _ in range(10):                      # the reference test launches 10 times
            num.deriv1d(axis, f, out, 0, 0, 0)
        assert np.max(np.abs(host(out) - sol)) / sol < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(64, 64, 64), (48, 40, 36), (300, 12, 12), (12, 300, 12), (12, 12, 300)])
@pytest.mark.parametrize("nbc", [NBC_CL, NBC_PER], ids=["closure", "periodic"])
def test_filter_matches_oracle(oracle, dims, nbc):
    import compact_schemes_amd as cs
    f = _field(dims, 3)
    ctx = oracle.Context(dims, nbc)
    num = cs.Numerics(dims, nbc)
    fin = dev(f)
    for axis in range(3):
        case(f"filter dims={dims} nbc={nbc[0][0]} axis={axis}")
        out = dev(np.zeros_like(f))
        num.filter(axis, fin, out)
        assert nrel(host(out), ctx.filter(axis, f)) < TOL
    case("filter in place")
    num.filter(1, fin, fin)
    assert nrel(host(fin), ctx.filter(1, f)) < TOL


@pytest.mark.gpu
def test_filter_transfer_function_gpu(oracle):
    """Cyclic filter on cos(w i): T(w) cos(w i), T(pi) = 0 (closed form)."""
    import compact_schemes_amd as cs
    n, af = 64, 0.49
    num = cs.Numerics((n, 4, 4), NBC_PER)
    an = oracle.filter_coeffs(5, af)
    i = np.tile(np.arange(n, dtype=float), 16)
    for k in (0, 3, 20, 32):
        w = 2 * np.pi * k / n
        f = np.cos(w * i)
        T = sum(an[m] * np.cos(m * w) for m in range(6)) / (1 + 2 * af * np.cos(w))
        out = dev(np.zeros_like(f))
        case(f"filter transfer k={k}")
        num.filter(0, dev(f), out)
        assert np.max(np.abs(host(out) - T * f)) < 1e-13


@pytest.mark.gpu
def test_errors_are_reported():
    import compact_schemes_amd as cs
    num = cs.Numerics((16, 16, 4), NBC_CL)
    x = dev(np.zeros(16 * 16 * 4))
    with pytest.raises(cs.CsError) as e:
        num.deriv(3, x, x)
    assert e.value.code == 1
    with pytest.raises(cs.CsError) as e:
        num.deriv(2, x, x)                        # 4 points along z
    assert e.value.code == 2
    num2 = cs.Numerics((16, 16, 16), NBC_PER, setup=False)
    y = dev(np.zeros(16 ** 3))
    with pytest.raises(cs.CsError) as e:
        num2.deriv(0, y, y)                       # seam rows before deriv_setup
    assert e.value.code == 5
    with pytest.raises(cs.CsError):
        cs.Numerics((16, 16, 16), [[10, 10, 99], [10, 10, 10]])


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(64, 64, 64), (100, 33, 17), (300, 9, 8), (48, 40, 36)])
@pytest.mark.parametrize("nbc", [NBC_CL, NBC_PER], ids=["closure", "periodic"])
def test_mixed_precision_solves_match_oracle(oracle, dims, nbc):
    """FP32 line solve + one FP64-residual refinement sweep: derivative and
    filter within 1e-11 (norm-wise relative) of the FP64 oracle; without the
    refinement an FP32 solve is ~1e-7 off."""
    import compact_schemes_amd as cs
    f = _field(dims, 7)
    ctx = oracle.Context(dims, nbc)
    num = cs.Numerics(dims, nbc)
    num.set_precision(1)
    fin = dev(f)
    for axis in range(3):
        case(f"mixed deriv dims={dims} nbc={nbc[0][0]} axis={axis}")
        out = dev(np.zeros_like(f))
        num.deriv(axis, fin, out)
        err = nrel(host(out), ctx.deriv_field(axis, f))
        assert err < 1e-11, err
        if nbc is NBC_PER and dims[axis] < 11:
            continue                                  # the cyclic filter needs 11 points
        case(f"mixed filter dims={dims} nbc={nbc[0][0]} axis={axis}")
        num.filter(axis, fin, out)
        err = nrel(host(out), ctx.filter(axis, f))
        assert err < 1e-11, err


@pytest.mark.gpu
def test_numerics_graph_replays_match_eager():
    """The plain-method step (3 derivatives, filter x, y, z) recorded into a
    hipGraph through torch.cuda.graph replays to the eager results, with
    eager work between the replays."""
    import torch
    import compact_schemes_amd as cs
    dims = (40, 32, 24)
    rng = np.random.default_rng(12)
    f = dev(rng.standard_normal(dims[0] * dims[1] * dims[2]))
    num = cs.Numerics(dims, NBC_PER)
    st = torch.cuda.Stream()
    d = [torch.empty_like(f) for _ in range(3)]
    g = torch.empty_like(f)

    def step(d, g):
        for a in range(3):
            num.deriv(a, f, d[a], stream=st)
        num.filter(0, f, g, stream=st)
        num.filter(1, g, g, stream=st)
        num.filter(2, g, g, stream=st)
    de = [torch.empty_like(f) for _ in range(3)]
    ge = torch.empty_like(f)
    step(de, ge)                                  # eager (also builds the operator tables)
    st.synchronize()
    case("numerics graph capture")
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=st):
        step(d, g)
    for rep in range(3):
        case(f"numerics graph replay {rep}")
        for t in d + [g]:
            t.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        junk = torch.randn(1000, device="cuda").sum()
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        for a in range(3):
            assert torch.equal(d[a], de[a]), (rep, a)
        assert torch.equal(g, ge), rep
