"""Batched subdomains (cs_batch_*): every subdomain's derivative and filter
results are bit-identical to the single-subdomain operators; the captured
graph replays the eager sequence; the autotuner's choices keep the identity.
"""
import numpy as np
import pytest

from gpu_util import case

PER = [[45, 45, 45], [45, 45, 45]]
CL = [[10, 10, 10], [10, 10, 10]]


def _dims(seed, count, lo=48, hi=96):
    rng = np.random.default_rng(seed)
    d = [tuple(int(v) for v in rng.integers(lo, hi + 1, size=3)) for _ in range(count)]
    d.append((300, 20, 18))          # 32-point chunks along x
    d.append((18, 272, 20))          # ... along y
    return d


def _fields(dims, seed):
    import torch
    from compact_schemes_amd import workloads as wl
    out = []
    for s, d in enumerate(dims):
        f, _ = wl.periodic_field(d, seed=seed + s, modes=8, kmax=20, noise=1e-3)
        out.append(torch.from_numpy(f).cuda())
    return out


def _unbatched(dims, nbc, fields):
    import torch
    import compact_schemes_amd as cs
    d_ref, f_ref = [], []
    for s, dm in enumerate(dims):
        num = cs.Numerics(dm, nbc[s])
        f = fields[s].clone()
        for axis in range(3):
            out = torch.empty_like(f)
            num.deriv(axis, f, out)
            d_ref.append(out)
        for axis in range(3):
            num.filter(axis, f, f)
        f_ref.append(f)
    torch.cuda.synchronize()
    return d_ref, f_ref


def _batch(dims, nbc, fields):
    import torch
    from compact_schemes_amd.batch import Batch
    b = Batch(dims, nbc)
    f = [t.clone() for t in fields]
    d = [torch.empty_like(f[s // 3]) for s in range(3 * len(dims))]
    b.bind(f, d)
    return b, f, d


@pytest.mark.gpu
@pytest.mark.parametrize("bcs", ["periodic", "mixed"])
def test_batch_bit_identical_to_single_subdomain(bcs):
    import torch
    dims = _dims(3, 10)
    nbc = [PER if (bcs == "periodic" or s % 2 == 0) else CL for s in range(len(dims))]
    fields = _fields(dims, 100)
    case(f"batch {len(dims)} subdomains bcs={bcs}")
    b, f, d = _batch(dims, nbc, fields)
    items, launches = b.info()
    assert min(items) > 0 and min(launches) > 0
    b.step()
    torch.cuda.synchronize()
    d_ref, f_ref = _unbatched(dims, nbc, fields)
    for s in range(len(dims)):
        for axis in range(3):
            assert torch.equal(d[3 * s + axis], d_ref[3 * s + axis]), f"derivative s={s} axis={axis} dims={dims[s]}"
        assert torch.equal(f[s], f_ref[s]), f"filter s={s} dims={dims[s]}"


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 2], ids=["persistent", "per-subdomain"])
def test_batch_grid_modes_agree(mode):
    """Persistent round-robin grid, one launch per subdomain and operator, and
    one workgroup per item: same bits."""
    import torch
    dims = _dims(6, 6)
    nbc = [PER] * len(dims)
    fields = _fields(dims, 400)
    b1, f1, d1 = _batch(dims, nbc, fields)
    b2, f2, d2 = _batch(dims, nbc, fields)
    b2.set_persistent(mode)
    assert b2.get_persistent() == mode
    case(f"batch grid mode {mode}")
    b1.step()
    b2.step()
    torch.cuda.synchronize()
    for s in range(len(dims)):
        assert torch.equal(f1[s], f2[s])
        for axis in range(3):
            assert torch.equal(d1[3 * s + axis], d2[3 * s + axis])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1, 2], ids=["per-item", "persistent", "per-subdomain"])
def test_batch_graph_replays_match_eager(mode):
    import torch
    dims = _dims(4, 6)
    nbc = [PER] * len(dims)
    fields = _fields(dims, 200)
    b, f, d = _batch(dims, nbc, fields)
    be, fe, de = _batch(dims, nbc, fields)
    b.set_persistent(mode)
    st = torch.cuda.Stream()
    case(f"batch graph capture mode {mode}")
    b.capture(st)
    torch.cuda.synchronize()
    for rep in range(3):
        case(f"batch graph replay {rep}")
        b.replay(st)
        be.step(st)
        st.synchronize()
        junk = torch.randn(1000, device="cuda").sum()      # eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        for s in range(len(dims)):
            assert torch.equal(f[s], fe[s]), f"replay {rep} filter s={s}"
            for axis in range(3):
                assert torch.equal(d[3 * s + axis], de[3 * s + axis]), f"replay {rep} deriv s={s} axis={axis}"


@pytest.mark.gpu
def test_batch_autotune_keeps_identity(tmp_path):
    import torch
    from compact_schemes_amd import batch as bm
    dims = _dims(5, 4)
    nbc = [PER] * len(dims)
    fields = _fields(dims, 300)
    bm.tune_clear()
    try:
        b, f, d = _batch(dims, nbc, fields)
        case("batch autotune")
        b.autotune(reps=2)
        torch.cuda.synchronize()
        for s in range(len(f)):                       # tuning reads the bound fields only
            assert torch.equal(f[s], fields[s].to(f[s].device)), s
        # periodic subdomains: seam rows at both ends (start = end = 1), FP64
        choices = {(op, n): bm.tune_get(op, n, 1, 1) for dm in dims for n in dm for op in (0, 1)}
        assert all(k in (8, 16, 32) for k in choices.values())
        for (op, n), k in choices.items():
            assert (n + k - 1) // k <= (16 if k == 32 else 64)      # chunks per line
        for s in range(len(f)):
            f[s].copy_(fields[s])
        b.step()
        torch.cuda.synchronize()
        d_ref, f_ref = _unbatched(dims, nbc, fields)        # new contexts: tuned tables
        for s in range(len(dims)):
            assert torch.equal(f[s], f_ref[s])
            for axis in range(3):
                assert torch.equal(d[3 * s + axis], d_ref[3 * s + axis])
        path = tmp_path / "tune.txt"
        bm.tune_save(path)
        bm.tune_clear()
        bm.tune_load(path)
        assert {(op, n): bm.tune_get(op, n, 1, 1) for (op, n) in choices} == choices
        assert all(bm.tune_get(op, n, 0, 0) == 0 for (op, n) in choices)     # closure ends: not tuned
    finally:
        bm.tune_clear()


@pytest.mark.gpu
def test_batch_errors():
    import compact_schemes_amd as cs
    from compact_schemes_amd.batch import Batch
    b = Batch([(32, 32, 32)], PER)
    with pytest.raises(cs.CsError) as e:
        b.step()
    assert e.value.code == 5
