"""Transposed line operators (cs_deriv_adjoint, cs_filter_adjoint).

At 16^3 the forward operator of each axis is assembled densely -- lines are
independent, so applying it to the N 'comb' fields (unit entry at row r of
every line) gives column r of every line's matrix -- and its transpose is
compared with the adjoint applied to the same combs.  At larger, ragged
sizes the dot-product identity <A v, w> = <v, A^T w> is checked.
"""
import numpy as np
import pytest

from gpu_util import case, dev, host

CL = [[10, 10, 10], [10, 10, 10]]
PER = [[45, 45, 45], [45, 45, 45]]
MIX = [[45, 10, 10], [45, 10, 10]]


def _line_index(dims, axis):
    """[n_lines, N] array of flat indices of each line along axis."""
    lxi, let, lze = dims
    idx = np.arange(lxi * let * lze).reshape(lze, let, lxi)
    if axis == 0:
        return idx.reshape(-1, lxi)
    if axis == 1:
        return idx.transpose(0, 2, 1).reshape(-1, let)
    return idx.transpose(1, 2, 0).reshape(-1, lze)


def _assemble(apply, dims, axis):
    """Per-line matrices M[line, i, r] of a line operator along axis."""
    lines = _line_index(dims, axis)
    n = lines.shape[1]
    lmx = dims[0] * dims[1] * dims[2]
    M = np.zeros((lines.shape[0], n, n))
    for r in range(n):
        e = np.zeros(lmx)
        e[lines[:, r]] = 1.0
        out = apply(e)
        M[:, :, r] = out[lines]
    return M


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(16, 16, 16), (16, 20, 24)])
@pytest.mark.parametrize("nbc", [CL, PER, MIX], ids=["closure", "periodic", "mixed"])
@pytest.mark.parametrize("op", ["deriv", "filter"])
def test_adjoint_is_dense_transpose(dims, nbc, op):
    import torch
    import compact_schemes_amd as cs
    num = cs.Numerics(dims, nbc)
    for axis in range(3):
        case(f"dense transpose {op} dims={dims} nbc={nbc[0]} axis={axis}")

        def fwd(e):
            x = dev(e)
            y = torch.empty_like(x)
            (num.deriv if op == "deriv" else num.filter)(axis, x, y)
            return host(y)

        def adj(e):
            x = dev(e)
            y = torch.empty_like(x)
            (num.deriv_adjoint if op == "deriv" else num.filter_adjoint)(axis, x, y)
            return host(y)
        A = _assemble(fwd, dims, axis)
        At = _assemble(adj, dims, axis)
        err = np.max(np.abs(At - np.transpose(A, (0, 2, 1)))) / np.max(np.abs(A))
        assert err < 1e-13, err


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(64, 48, 100), (300, 20, 18), (24, 272, 20), (17, 33, 65)])
@pytest.mark.parametrize("nbc", [CL, PER, MIX], ids=["closure", "periodic", "mixed"])
def test_adjoint_dot_product(dims, nbc):
    import torch
    import compact_schemes_amd as cs
    rng = np.random.default_rng(11)
    lmx = dims[0] * dims[1] * dims[2]
    v, w = dev(rng.standard_normal(lmx)), dev(rng.standard_normal(lmx))
    num = cs.Numerics(dims, nbc)
    for axis in range(3):
        for op in ("deriv", "filter"):
            case(f"dot product {op} dims={dims} nbc={nbc[0]} axis={axis}")
            av, atw = torch.empty_like(v), torch.empty_like(w)
            (num.deriv if op == "deriv" else num.filter)(axis, v, av)
            (num.deriv_adjoint if op == "deriv" else num.filter_adjoint)(axis, w, atw)
            lhs = float(torch.dot(av, w))
            rhs = float(torch.dot(v, atw))
            scale = float(av.norm() * w.norm())
            assert abs(lhs - rhs) <= 1e-13 * scale, (op, axis, lhs, rhs)
