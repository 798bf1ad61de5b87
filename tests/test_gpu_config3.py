"""The full batched workload: 96 periodic subdomains with seeded sizes in
[48, 320] (seeds 3..98), the plain method (derivative along x, y, z, then the
filter along x, y, z) on every one of them in batched launches.

After the autotuner has chosen chunk lengths and the grid mode, the captured
graph's replay must be bit-identical (torch.equal) to running cs_deriv /
cs_filter on each subdomain alone, and the three smallest subdomains must
match the CPU oracle to 1e-12 (norm-wise, FP64).
"""
import numpy as np
import pytest

from gpu_util import case, host, nrel

PER = [[45, 45, 45], [45, 45, 45]]


@pytest.mark.gpu
def test_batch96_tuned_graph_bit_identical_and_oracle(oracle):
    import torch
    import compact_schemes_amd as cs
    from compact_schemes_amd import batch as bm
    from compact_schemes_amd.workloads import batch_fields
    case("config 3: generating 96 subdomains")
    dims, f0 = batch_fields()
    assert len(dims) == 96 and all(48 <= n <= 320 for d in dims for n in d)
    npts = sum(d[0] * d[1] * d[2] for d in dims)
    assert len({n for d in dims for n in d}) > 100          # uneven line lengths
    f = [t.clone() for t in f0]
    d = [torch.empty_like(f[s // 3]) for s in range(3 * len(dims))]
    st = torch.cuda.Stream()
    bm.tune_clear()
    try:
        b = bm.Batch(dims, PER)
        b.bind(f, d)
        case(f"config 3: autotune ({npts} points)")
        b.autotune(reps=2, stream=st)
        st.synchronize()
        for s in range(len(f)):
            assert torch.equal(f[s], f0[s]), f"autotune changed field {s}"
        case(f"config 3: capture + replay (grid mode {b.get_persistent()})")
        b.capture(st)
        for s in range(len(f)):
            f[s].copy_(f0[s])
        torch.cuda.synchronize()
        b.replay(st)
        st.synchronize()
        case("config 3: one subdomain at a time")
        for s, dm in enumerate(dims):
            num = cs.Numerics(dm, PER)                 # built after tuning: the tuned tables
            g = f0[s].clone()
            out = torch.empty_like(g)
            for axis in range(3):
                num.deriv(axis, g, out)
                assert torch.equal(out, d[3 * s + axis]), (s, dm, axis)
            for axis in range(3):
                num.filter(axis, g, g)
            assert torch.equal(g, f[s]), (s, dm, "filter")
            del num
        case("config 3: oracle on the three smallest subdomains")
        small = sorted(range(len(dims)), key=lambda s: dims[s][0] * dims[s][1] * dims[s][2])[:3]
        for s in small:
            dm = dims[s]
            ctx = oracle.Context(dm, PER)
            x = host(f0[s])
            for axis in range(3):
                assert nrel(host(d[3 * s + axis]), ctx.deriv_field(axis, x)) < 1e-12, (dm, axis)
            want = ctx.filter(2, ctx.filter(1, ctx.filter(0, x)))
            assert nrel(host(f[s]) - x, want - x) < 1e-11, (dm, "filter")
            assert nrel(host(f[s]), want) < 1e-12, (dm, "filter")
    finally:
        bm.tune_clear()
