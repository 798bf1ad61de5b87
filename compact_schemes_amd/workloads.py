"""Seeded synthetic inputs of the benchmark and test workloads (numpy, FP64).

Fields use the reference layout: SoA, variable-major, x fastest,
idx = (k*let + j)*lxi + i.  Grid spacing is 1 (unit Cartesian metrics).
"""
import numpy as np

GAM = 1.4

NBC_PERIODIC = [[45, 45, 45], [45, 45, 45]]
NBC_NON_REFLECTIVE = [[10, 10, 10], [10, 10, 10]]


def grid_index(dims):
    lxi, let, lze = dims
    k, j, i = np.meshgrid(np.arange(lze), np.arange(let), np.arange(lxi), indexing="ij")
    return i.astype(np.float64), j.astype(np.float64), k.astype(np.float64)


def periodic_field(dims, seed=1, modes=8, kmax=20, noise=1e-3):
    """Sum of `modes` plane waves sin(2 pi (kx x + ky y + kz z)/L + phi) with
    wavenumbers in 1..kmax and phases in [0, 2 pi), plus `noise` N(0,1).
    Returns (f, df_dx, df_dy, df_dz) of the noise-free part as flat arrays."""
    rng = np.random.default_rng(seed)
    lxi, let, lze = dims
    x, y, z = grid_index(dims)
    f = np.zeros_like(x)
    g = [np.zeros_like(x) for _ in range(3)]
    for _ in range(modes):
        kx, ky, kz = rng.integers(1, kmax + 1, size=3)
        phi = rng.uniform(0.0, 2.0 * np.pi)
        wx, wy, wz = 2 * np.pi * kx / lxi, 2 * np.pi * ky / let, 2 * np.pi * kz / lze
        arg = wx * x + wy * y + wz * z + phi
        f += np.sin(arg)
        c = np.cos(arg)
        g[0] += wx * c
        g[1] += wy * c
        g[2] += wz * c
    f_noisy = f + noise * rng.standard_normal(f.shape)
    return f_noisy.reshape(-1), [gi.reshape(-1) for gi in g]


def metrics(dims):
    """Unit Cartesian metrics: xim = (1,0,0), etm = (0,1,0), zem = (0,0,1), yaco = 1."""
    lmx = dims[0] * dims[1] * dims[2]
    xim = np.zeros(3 * lmx)
    etm = np.zeros(3 * lmx)
    zem = np.zeros(3 * lmx)
    xim[0:lmx] = 1.0
    etm[lmx:2 * lmx] = 1.0
    zem[2 * lmx:3 * lmx] = 1.0
    return xim, etm, zem, np.ones(lmx)


def face_normals(dims):
    """cm[a] = [2 face][3][face_size]: unit normal e_a on both faces of axis a."""
    lxi, let, lze = dims
    fs = (let * lze, lxi * lze, lxi * let)
    cm = []
    for a in range(3):
        c = np.zeros((2, 3, fs[a]))
        c[:, a, :] = 1.0
        cm.append(c.reshape(-1))
    return cm


def boundary_weights(dims):
    """sbcc: 1 on every boundary point of the subdomain, 0 inside."""
    x, y, z = grid_index(dims)
    lxi, let, lze = dims
    on = (x == 0) | (x == lxi - 1) | (y == 0) | (y == let - 1) | (z == 0) | (z == lze - 1)
    return on.astype(np.float64).reshape(-1)


def flow_state(dims, seed=2, mach=0.5, amp=1e-3, width=8.0, mu=1e-4, noise=0.0):
    """Uniform flow at Mach `mach` in x (carried by the frame velocity umf,
    stored velocity = perturbation), rho0 = 1, p0 = 1/gam (c0 = 1), with a
    Gaussian pressure pulse p' = amp exp(-ln2 r^2/width^2) at a seeded centre
    in the middle half of the box and rho' = p'/c0^2.  `noise` adds seeded
    N(0,1)*noise perturbations to rho, u, v, w, p (exercises every term).
    Returns a dict of flat float64 arrays and scalars."""
    rng = np.random.default_rng(seed)
    lxi, let, lze = dims
    lmx = lxi * let * lze
    x, y, z = grid_index(dims)
    cen = [rng.uniform(0.25 * n, 0.75 * n) for n in dims]
    r2 = (x - cen[0]) ** 2 + (y - cen[1]) ** 2 + (z - cen[2]) ** 2
    pp = amp * np.exp(-np.log(2.0) * r2 / width ** 2)
    p = 1.0 / GAM + pp
    rho = 1.0 + pp
    u = np.zeros_like(rho)
    v = np.zeros_like(rho)
    w = np.zeros_like(rho)
    if noise > 0:
        rho = rho + noise * rng.standard_normal(rho.shape)
        u = u + noise * rng.standard_normal(rho.shape)
        v = v + noise * rng.standard_normal(rho.shape)
        w = w + noise * rng.standard_normal(rho.shape)
        p = p + noise * rng.standard_normal(rho.shape) / GAM
    E = p / (GAM - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    qa = np.concatenate([a.reshape(-1) for a in (rho, rho * u, rho * v, rho * w, E)])
    xim, etm, zem, yaco = metrics(dims)
    return {
        "qa": qa,
        "xim": xim, "etm": etm, "zem": zem, "yaco": yaco,
        "ssk": np.full(lmx, mu),
        "cm": face_normals(dims),
        "sbcc": boundary_weights(dims),
        "umf": (mach, 0.0, 0.0),
        "center": cen,
    }


def batch_fields(count=96, seed0=3, lo=48, hi=320):
    """Subdomain s: sizes drawn with seed seed0+s in [lo, hi]; field = 8 plane
    waves with seeded wavenumbers 1..20 and phases + 1e-3 N(0,1) noise,
    generated on the device."""
    import numpy as np
    import torch
    dims, fields = [], []
    for s in range(count):
        rng = np.random.default_rng(seed0 + s)
        d = tuple(int(v) for v in rng.integers(lo, hi + 1, size=3))
        lxi, let, lze = d
        z, y, x = torch.meshgrid(torch.arange(lze, device="cuda", dtype=torch.float64),
                                 torch.arange(let, device="cuda", dtype=torch.float64),
                                 torch.arange(lxi, device="cuda", dtype=torch.float64), indexing="ij")
        f = torch.zeros_like(x)
        for _ in range(8):
            kx, ky, kz = (int(v) for v in rng.integers(1, 21, size=3))
            phi = float(rng.uniform(0.0, 2.0 * np.pi))
            f += torch.sin(2 * np.pi * (kx * x / lxi + ky * y / let + kz * z / lze) + phi)
        g = torch.Generator(device="cuda")
        g.manual_seed(seed0 + s)
        f += 1e-3 * torch.randn(f.shape, generator=g, device="cuda", dtype=torch.float64)
        dims.append(d)
        fields.append(f.reshape(-1).contiguous())
        del x, y, z
    return dims, fields


def with_spacing(st, dims, spacing):
    """The state `st` on a Cartesian grid with point spacings (dx, dy, dz)
    (negative: the coordinate decreases with the index): metrics are the
    cofactors xim = (dy dz, 0, 0), etm = (0, dz dx, 0), zem = (0, 0, dx dy),
    yaco = 1 / (dx dy dz), and the face normals cm[a] are the unit vectors of
    grad xi_a = yaco * metric_a * e_a.  Returns a new dict."""
    dx, dy, dz = spacing
    lmx = dims[0] * dims[1] * dims[2]
    out = dict(st)
    m = (dy * dz, dz * dx, dx * dy)
    J = 1.0 / (dx * dy * dz)
    xim, etm, zem = np.zeros(3 * lmx), np.zeros(3 * lmx), np.zeros(3 * lmx)
    xim[0:lmx] = m[0]
    etm[lmx:2 * lmx] = m[1]
    zem[2 * lmx:3 * lmx] = m[2]
    out["xim"], out["etm"], out["zem"] = xim, etm, zem
    out["yaco"] = np.full(lmx, J)
    out["cm"] = [c * np.sign(J * m[a]) for a, c in enumerate(face_normals(dims))]
    return out


def face_normals_from_metrics(dims, xim, etm, zem):
    """cm[a] = [2 face][3][face]: the unit normals grad(xi_a)/|grad(xi_a)| on
    the two faces of axis a from metric rows as cs_metrics returns them
    ([3][lmx] each, proportional to grad(xi_a) with the positive factor
    1/yaco when yaco > 0), face points ordered as the gcbc face loops run."""
    lxi, let, lze = dims
    cm = []
    for a, (rows, n) in enumerate(zip((xim, etm, zem), dims)):
        r = np.asarray(rows).reshape(3, lze, let, lxi)
        c = []
        for face in range(2):
            idx = face * (n - 1)
            fr = r[:, :, :, idx] if a == 0 else (r[:, :, idx, :] if a == 1 else r[:, idx, :, :])
            fr = fr.reshape(3, -1)
            c.append(fr / np.sqrt((fr * fr).sum(0)))
        cm.append(np.stack(c).reshape(-1))
    return cm


def curvilinear_state(dims, seed=2, skew=0.3, **flow):
    """flow_state on the periodic sine-skewed grid x = i + s sin(2 pi j / Ny),
    y = j + s sin(2 pi k / Nz), z = k + s sin(2 pi i / Nx) (unit index
    spacing), with its analytic metrics as the reference stores them: xim,
    etm, zem = the cofactor rows of the coordinate Jacobian (d x/d eta x
    d x/d zeta, ...), yaco = 1 / det; cm = the unit normals grad(xi_a) /
    |grad(xi_a)| on the two faces of each axis ([2 face][3][face], face points
    ordered as the gcbc face loops run: x k*let+j, y k*lxi+i, z j*lxi+i)."""
    st = flow_state(dims, seed=seed, **flow)
    nx, ny, nz = dims
    lmx = nx * ny * nz
    x, y, z = (g.reshape(-1) for g in grid_index(dims))
    tw = 2 * np.pi
    one, zero = np.ones(lmx), np.zeros(lmx)
    # M[c][a] = d x_c / d xi_a
    M = [[one, skew * tw / ny * np.cos(tw * y / ny), zero],
         [zero, one, skew * tw / nz * np.cos(tw * z / nz)],
         [skew * tw / nx * np.cos(tw * x / nx), zero, one]]
    u, v, w = ([M[c][a] for c in range(3)] for a in range(3))
    cross = lambda p, q: [p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0]]
    rows = [cross(v, w), cross(w, u), cross(u, v)]
    det = u[0] * rows[0][0] + u[1] * rows[0][1] + u[2] * rows[0][2]
    st["xim"], st["etm"], st["zem"] = (np.concatenate(r) for r in rows)
    st["yaco"] = 1.0 / det
    st["cm"] = face_normals_from_metrics(dims, st["xim"], st["etm"], st["zem"])
    return st
