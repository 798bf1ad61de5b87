"""Intra-GPU multi-block seams (cs_exchange): a periodic domain split into
blocks along x, each block a numerics context whose x faces are
BC_INTER_SUBDOMAINS seams to the other block.  The sequence is the
reference's fill_buffers -> mpigo -> deriv1d (src/physics_pc.hpp:99-120) with
the exchange between contexts standing in for the MPI neighbour ranks.

Checks: the send/recv buffers and the x derivative of every block against
the oracle running the same blocks (its recv filled from the other block's
send); the blocked derivative of a smooth periodic field against the
analytic derivative, to the accuracy of the single-block periodic derivative
(the seam rows are the same interface scheme at every seam); argument checks.
"""
import numpy as np
import pytest

from gpu_util import case, dev, host, nrel

TOL = 1e-12
SEAM, PER = 40, 45


def _blocks_nbc():
    # x faces: seams to the neighbour block; y, z: periodic on the block itself
    return [[SEAM, PER, PER], [SEAM, PER, PER]]


def _split(f, dims, nb):
    nx, ny, nz = dims
    g = f.reshape(nz, ny, nx)
    w = nx // nb
    return [np.ascontiguousarray(g[:, :, b * w:(b + 1) * w]).reshape(-1) for b in range(nb)]


def _exchange_all(nums, axis, var):
    nb = len(nums)
    for b in range(nb):
        nums[b].exchange(1, nums[(b + 1) % nb], 0, axis, var)


@pytest.mark.gpu
@pytest.mark.parametrize("nb", [2, 3])
def test_blocks_match_oracle(oracle, nb):
    import compact_schemes_amd as cs
    dims = (24 * nb, 20, 16)
    bd = (24, 20, 16)
    lmx = bd[0] * bd[1] * bd[2]
    rng = np.random.default_rng(11)
    f = rng.standard_normal(dims[0] * dims[1] * dims[2])
    parts = _split(f, dims, nb)
    nbc = _blocks_nbc()
    nums = [cs.Numerics(bd, nbc) for _ in range(nb)]
    ctxs = [oracle.Context(bd, nbc) for _ in range(nb)]
    var = 2
    case(f"blocks nb={nb}")
    fins = []
    for b in range(nb):
        f3 = np.zeros(5 * 3 * lmx)
        f3[(3 * var) * lmx:(3 * var + 1) * lmx] = parts[b]        # component 0 of variable `var`
        fins.append(dev(f3))
        nums[b].fill_buffers(fins[b], var, var + 1, 1)           # nrt = 1: component 0 on every axis
        ctxs[b].fill(0, parts[b].copy(), var)
    _exchange_all(nums, 0, var)
    for b in range(nb):
        nxt = (b + 1) % nb
        ctxs[nxt].recv(0)[var, 0] = ctxs[b].send(0)[var, 1]
        ctxs[b].recv(0)[var, 1] = ctxs[nxt].send(0)[var, 0]
    for b in range(nb):
        _, send, recv = nums[b].buffers(0)
        assert nrel(host(send)[var], ctxs[b].send(0)[var]) < 1e-14
        assert nrel(host(recv)[var], ctxs[b].recv(0)[var]) < 1e-14
        out = dev(np.zeros(lmx))
        nums[b].deriv1d(0, dev(parts[b]), out, 1, 1, var)
        ref = np.zeros(lmx)
        ctxs[b].deriv(0, parts[b].copy(), ref, 1, 1, var)
        assert nrel(host(out), ref) < TOL, b


@pytest.mark.gpu
def test_blocks_smooth_field_accuracy():
    """A smooth periodic field differentiated on two blocks is as accurate as
    on one periodic block (both carry interface-scheme seam rows)."""
    import compact_schemes_amd as cs
    dims = (64, 16, 8)
    nx, ny, nz = dims
    x = np.arange(nx)
    k = 2 * np.pi * 3 / nx
    f1 = np.sin(k * x + 0.3)
    df1 = k * np.cos(k * x + 0.3)
    f = np.broadcast_to(f1, (nz, ny, nx)).reshape(-1).copy()
    df = np.broadcast_to(df1, (nz, ny, nx)).reshape(-1)
    one = cs.Numerics(dims, [[PER] * 3, [PER] * 3])
    out1 = dev(np.zeros(f.size))
    case("blocks smooth field")
    one.deriv(0, dev(f), out1)
    err1 = nrel(host(out1), df)
    bd = (32, 16, 8)
    parts, dparts = _split(f, dims, 2), _split(df, dims, 2)
    nums = [cs.Numerics(bd, _blocks_nbc()) for _ in range(2)]
    lmx = bd[0] * bd[1] * bd[2]
    fins = []
    for b in range(2):
        f3 = np.zeros(5 * 3 * lmx)
        f3[:lmx] = parts[b]
        fins.append(dev(f3))
        nums[b].fill_buffers(fins[b], 0, 1, 1)
    _exchange_all(nums, 0, 0)
    errs = []
    for b in range(2):
        out = dev(np.zeros(lmx))
        nums[b].deriv1d(0, dev(parts[b]), out, 1, 1, 0)
        errs.append(nrel(host(out), dparts[b]))
    assert max(errs) < 2.0 * err1 + 1e-12, (errs, err1)
    assert max(errs) < 1e-3


@pytest.mark.gpu
def test_exchange_errors():
    import compact_schemes_amd as cs
    a = cs.Numerics((24, 20, 16), _blocks_nbc())
    b = cs.Numerics((24, 20, 12), _blocks_nbc())
    c = cs.Numerics((24, 20, 16), [[10, 45, 45], [10, 45, 45]])
    with pytest.raises(cs.CsError):
        a.exchange(1, b, 0, 0, 0)          # faces of different sizes
    with pytest.raises(cs.CsError):
        a.exchange(1, c, 0, 0, 0)          # closure face is no seam
    with pytest.raises(cs.CsError):
        a.exchange(1, a, 1, 0, 0)          # a face cannot adjoin itself
    a.exchange(0, a, 1, 1, 0)              # the periodic self-wrap of y
