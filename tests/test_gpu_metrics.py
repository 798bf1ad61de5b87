"""Grid metrics from coordinates (cs_metrics) -- xim, etm, zem = cofactors of
the coordinate Jacobian d x_c / d xi_a, yaco = 1 / det.

Pins: a uniform grid and an affine grid give the exact metrics (the compact
derivative with closure rows is exact for linear data, so the result is the
closed-form cofactor matrix); a periodic axis with its period jump gives the
uniform spacing; a smooth curvilinear grid matches the analytic metrics to
the scheme's truncation error; and every metric equals the oracle's
derivatives combined by the same cofactor formula (1e-12).
"""
import numpy as np
import pytest

from gpu_util import case, dev, host, nrel

CL = [[10, 10, 10], [10, 10, 10]]


def _coords(dims, fn):
    nx, ny, nz = dims
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = fn(i.astype(float), j.astype(float), k.astype(float))
    return np.concatenate([np.broadcast_to(v, i.shape).reshape(-1) for v in (x, y, z)])


def _cofactors(M):
    """M[c][a] arrays -> (xim, etm, zem, yaco) (numpy, for the analytic pins)."""
    u, v, w = ([M[c][a] for c in range(3)] for a in range(3))
    cross = lambda p, q: [p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0]]
    cx, ce, cz = cross(v, w), cross(w, u), cross(u, v)
    det = u[0] * cx[0] + u[1] * cx[1] + u[2] * cx[2]
    return np.concatenate(cx), np.concatenate(ce), np.concatenate(cz), 1.0 / det


def _metrics(dims, nbc, xyz, period=None):
    import compact_schemes_amd as cs
    num = cs.Numerics(dims, nbc)
    return [host(t) for t in num.metrics(dev(xyz), period)]


@pytest.mark.gpu
def test_metrics_uniform_and_affine():
    dims = (24, 20, 16)
    lmx = dims[0] * dims[1] * dims[2]
    case("metrics uniform/affine")
    h = (0.5, 2.0, 1.5)
    xim, etm, zem, yaco = _metrics(dims, CL, _coords(dims, lambda i, j, k: (h[0] * i, h[1] * j, h[2] * k)))
    one = np.ones(lmx)
    assert np.allclose(xim, np.concatenate([h[1] * h[2] * one, 0 * one, 0 * one]), rtol=0, atol=1e-13)
    assert np.allclose(etm, np.concatenate([0 * one, h[0] * h[2] * one, 0 * one]), rtol=0, atol=1e-13)
    assert np.allclose(zem, np.concatenate([0 * one, 0 * one, h[0] * h[1] * one]), rtol=0, atol=1e-13)
    assert nrel(yaco, one / (h[0] * h[1] * h[2])) < 1e-14
    A = np.array([[1.0, 0.3, -0.2], [0.1, 0.8, 0.25], [-0.15, 0.05, 1.2]])      # x_c = A[c][a] xi_a + b_c
    xyz = _coords(dims, lambda i, j, k: tuple(A[c, 0] * i + A[c, 1] * j + A[c, 2] * k + c for c in range(3)))
    got = _metrics(dims, CL, xyz)
    want = _cofactors([[A[c, a] * one for a in range(3)] for c in range(3)])
    for g, w in zip(got, want):
        assert nrel(g, w) < 1e-13


@pytest.mark.gpu
def test_metrics_periodic_jump():
    dims = (32, 16, 12)
    lmx = dims[0] * dims[1] * dims[2]
    nbc = [[45, 10, 10], [45, 10, 10]]
    hx = 0.25
    xyz = _coords(dims, lambda i, j, k: (hx * i, 1.0 * j, 1.0 * k))
    case("metrics periodic")
    xim, etm, zem, yaco = _metrics(dims, nbc, xyz, period=[[dims[0] * hx, 0, 0], [0, 0, 0], [0, 0, 0]])
    assert nrel(etm[lmx:2 * lmx], hx * np.ones(lmx)) < 1e-13
    assert nrel(yaco, np.ones(lmx) / hx) < 1e-13


@pytest.mark.gpu
def test_metrics_curvilinear(oracle):
    """x = i + a sin(2 pi j / Ny), y = j + a sin(2 pi k / Nz), z = k + a sin(2 pi i / Nx)."""
    dims = (40, 36, 32)
    nx, ny, nz = dims
    lmx = nx * ny * nz
    a = 0.3
    tw = 2 * np.pi
    fn = lambda i, j, k: (i + a * np.sin(tw * j / ny), j + a * np.sin(tw * k / nz), k + a * np.sin(tw * i / nx))
    xyz = _coords(dims, fn)
    case("metrics curvilinear")
    got = _metrics(dims, CL, xyz)
    # analytic Jacobian
    k, j, i = (g.reshape(-1) for g in np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij"))
    one, zero = np.ones(lmx), np.zeros(lmx)
    M = [[one, a * tw / ny * np.cos(tw * j / ny), zero],
         [zero, one, a * tw / nz * np.cos(tw * k / nz)],
         [a * tw / nx * np.cos(tw * i / nx), zero, one]]
    want = _cofactors(M)
    for g, w in zip(got, want):
        assert nrel(g, w) < 1e-4          # truncation of the 3rd-order closure rows: 3.2e-5 at 32-40 points/wave
    # the oracle's derivatives combined by the same formula
    ctx = oracle.Context(dims, CL)
    Mo = [[ctx.deriv_field(ax, xyz[c * lmx:(c + 1) * lmx].copy()) for ax in range(3)] for c in range(3)]
    for g, w in zip(got, _cofactors(Mo)):
        assert nrel(g, w) < 1e-12
