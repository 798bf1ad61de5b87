"""Run configuration (config.yaml schema), block grid and legacy-VTK output:
host code of the C-ABI library, no GPU.

The schema is the reference's config.yaml:1-18 as read by
driver/canard.cu:37-98, src/domdcomp.hpp:67-107, src/grid.hpp:74-87 and
src/physics_base.hpp:135-145; the grid is grid::generate
(src/grid.hpp:89-110); the VTK file is IOwriter::go_vtk (src/IO.hpp:108-159).
"""
import os

import numpy as np
import pytest

from vtk_util import read_vtk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "run_config.yaml")
REFERENCE_CONFIG = "/root/reference/config.yaml"


def test_config_fixture_values():
    import compact_schemes_amd as cs
    cfg = cs.load_config(FIXTURE)
    assert cfg.nblocks == 0 and list(cfg.nbpc) == [1, 1, 1]
    assert (cfg.lximb, cfg.letmb, cfg.lzemb) == (47, 31, 23)
    dims, nbc = cs.config_dims(cfg)
    assert dims == (48, 32, 24)
    assert nbc == [[45, 10, 10], [45, 10, 10]]
    assert list(cfg.mbcd) == [-1] * 6
    assert (cfg.doml0, cfg.doml1, cfg.domh, cfg.span) == (0.5, 1.5, 0.25, 0.2)
    assert (cfg.amach.x, cfg.amach.y, cfg.amach.z) == (0.5, -0.1, 0.0)
    assert (cfg.cfl, cfg.tmax, cfg.ndata, cfg.nsteps, cfg.viscous, cfg.re) == (0.8, 0.5, 3, 7, 0, 2500.0)
    assert cfg.pulse_amp == 1e-4
    # not in the file: the defaults
    assert cfg.alpha_f == 0.49 and cfg.filter == 1 and cfg.pulse_width == 8.0


def test_config_defaults_are_the_driver_constants():
    """cfl 0.95, tmax 1, two outputs: driver/canard.cu:117-127."""
    import compact_schemes_amd as cs
    cfg = cs.config_defaults()
    assert (cfg.cfl, cfg.tmax, cfg.ndata, cfg.nsteps) == (0.95, 1.0, 2, 0)


@pytest.mark.skipif(not os.path.exists(REFERENCE_CONFIG), reason="reference checkout not present")
def test_reference_config_file_parses():
    """The reference's own config.yaml: two ranks along x on block 0, a
    2048 x 64 x 64 block, non-reflective faces, Mach 0.3 along x."""
    import compact_schemes_amd as cs
    cfg = cs.load_config(REFERENCE_CONFIG)
    assert list(cfg.nbpc) == [2, 1, 1]
    assert cs.config_dims(cfg) == ((2048, 64, 64), [[10, 10, 10], [10, 10, 10]])
    assert (cfg.doml0, cfg.doml1, cfg.domh, cfg.span) == (0.5, 0.5, 0.5, 0.1)
    assert (cfg.amach.x, cfg.amach.y, cfg.amach.z) == (0.3, 0.0, 0.0)


BAD = [
    ("driver:\n  - nblocks: 0\n", "missing domdcomp.nbpc"),
    ("driver:\n  - nblocks: x\n", ":2: not an integer"),
    ("driver\n", ":1: expected a section name"),
    ("  - nblocks: 0\n", ":1: entry outside a section"),
]


@pytest.mark.parametrize("text,msg", BAD)
def test_config_errors_name_the_line(tmp_path, text, msg):
    import compact_schemes_amd as cs
    p = tmp_path / "bad.yaml"
    p.write_text(text)
    with pytest.raises(cs.CsError, match=msg):
        cs.load_config(str(p))


def test_config_rejects_unknown_bc_and_short_lists(tmp_path):
    import compact_schemes_amd as cs
    text = open(FIXTURE).read()
    p = tmp_path / "c.yaml"
    p.write_text(text.replace("[45, 10, 10, 45, 10, 10]", "[45, 10, 11, 45, 10, 10]"))
    with pytest.raises(cs.CsError, match="unknown boundary code 11"):
        cs.load_config(str(p))
    p.write_text(text.replace("[45, 10, 10, 45, 10, 10]", "[45, 10, 10]"))
    with pytest.raises(cs.CsError, match="nbbc needs 6"):
        cs.load_config(str(p))
    with pytest.raises(cs.CsError, match="cannot open"):
        cs.load_config(str(tmp_path / "missing.yaml"))


def test_grid_spans_and_spacing():
    """x runs from -doml0 to doml1, y from -domh to domh, z from span/2 down
    to -span/2, each with constant spacing along its own index only."""
    import compact_schemes_amd as cs
    cfg = cs.load_config(FIXTURE)
    (nx, ny, nz), _ = cs.config_dims(cfg)
    xyz = cs.config_grid(cfg).reshape(3, nz, ny, nx)
    x, y, z = xyz
    assert np.allclose(x[:, :, 0], -0.5, rtol=0, atol=1e-15) and np.allclose(x[:, :, -1], 1.5, rtol=0, atol=1e-15)
    assert np.allclose(y[:, 0, :], -0.25, rtol=0, atol=1e-15) and np.allclose(y[:, -1, :], 0.25, rtol=0, atol=1e-15)
    assert np.allclose(z[0], 0.1, rtol=0, atol=1e-15) and np.allclose(z[-1], -0.1, rtol=0, atol=1e-15)
    assert np.allclose(np.diff(x, axis=2), 2.0 / 47, rtol=0, atol=1e-14)
    assert np.allclose(np.diff(y, axis=1), 0.5 / 31, rtol=0, atol=1e-14)
    assert np.allclose(np.diff(z, axis=0), -0.2 / 23, rtol=0, atol=1e-14)
    assert np.all(np.diff(x, axis=0) == 0) and np.all(np.diff(x, axis=1) == 0)
    assert np.all(np.diff(y, axis=0) == 0) and np.all(np.diff(z, axis=1) == 0)


@pytest.mark.parametrize("precision", [32, 64])
def test_vtk_roundtrip(tmp_path, precision):
    import compact_schemes_amd as cs
    dims = (5, 4, 3)
    n = 60
    rng = np.random.default_rng(3)
    xyz = rng.standard_normal(3 * n)
    rho, p = rng.uniform(0.5, 2, n), rng.uniform(0.1, 1, n)
    path = str(tmp_path / "out.vtk")
    cs.write_vtk(path, dims, xyz, {"density": rho, "p": p}, precision=precision)
    d, pts, fields = read_vtk(path)
    assert d == dims and list(fields) == ["density", "p"]
    cast = (lambda a: a.astype(np.float32).astype(np.float64)) if precision == 32 else (lambda a: a)
    assert np.array_equal(pts, cast(xyz.reshape(3, n).T))       # points interleaved x, y, z
    assert np.array_equal(fields["density"], cast(rho)) and np.array_equal(fields["p"], cast(p))


def test_vtk_rejects_bad_arguments(tmp_path):
    import compact_schemes_amd as cs
    with pytest.raises(cs.CsError, match="precision"):
        cs.write_vtk(str(tmp_path / "a.vtk"), (2, 2, 2), np.zeros(24), {}, precision=16)
    with pytest.raises(cs.CsError, match="names"):
        cs.write_vtk(str(tmp_path / "a.vtk"), (2, 2, 2), np.zeros(24), {"two words": np.zeros(8)})
    with pytest.raises(cs.CsError, match="cannot open"):
        cs.write_vtk(str(tmp_path / "no" / "a.vtk"), (2, 2, 2), np.zeros(24), {})
