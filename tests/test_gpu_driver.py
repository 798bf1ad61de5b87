"""The run driver end to end on a small block: configuration -> grid ->
initial pulse -> RK4 steps -> VTK files, against the same steps through the
library's Python binding (same solver configuration, same initial state)."""
import os
import subprocess

import numpy as np
import pytest

from gpu_util import case, dev, host
from vtk_util import read_vtk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CANARD = os.path.join(ROOT, "compact_schemes_amd", "driver", "_bin", "canard")

CONFIG = """driver:
  - nblocks: 0
domdcomp:
  - nbpc: [1, 1, 1]
  - lximb: [39]
  - letmb: [27]
  - lzemb: [23]
  - nbbc: [45, 10, 10, 45, 10, 10]
  - mbcd: [-1, -1, -1, -1, -1, -1]
grid:
  - doml0: 0.5
  - doml1: 1.5
  - domh: 0.6
  - span: 1.0
physics:
  - amach1: 0.5
  - amach2: 0.0
  - amach3: 0.1
run:
  - tmax: 100.0
  - ndata: 1
  - nsteps: 4
  - re: 1.0e3
  - pulse_amp: 1.0e-2
  - pulse_width: 4.0
"""


@pytest.mark.gpu
def test_driver_run_matches_library_steps(tmp_path):
    import torch
    import compact_schemes_amd as cs
    from compact_schemes_amd import workloads as wl
    from compact_schemes_amd.fused import FusedSolver
    cfgp = tmp_path / "config.yaml"
    cfgp.write_text(CONFIG)
    case("driver: 40x28x24 block, 4 steps")
    r = subprocess.run([CANARD, str(cfgp), str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "4 steps" in r.stdout
    files = sorted(f for f in os.listdir(tmp_path) if f.endswith(".vtk"))
    assert files == ["output_data0.vtk", "output_data1.vtk"]     # t = 0 and the end

    cfg = cs.load_config(str(cfgp))
    dims, nbc = cs.config_dims(cfg)
    lmx = dims[0] * dims[1] * dims[2]
    dx, dy, dz = 2.0 / 39, 1.2 / 27, -1.0 / 23
    st = wl.with_spacing(wl.flow_state(dims, seed=1, mu=1e-3), dims, (dx, dy, dz))
    st["umf"] = (0.5, 0.0, 0.1)
    sv = FusedSolver(dims, nbc, st)
    q = torch.empty(5 * lmx, dtype=torch.float64, device="cuda")
    cs.init_pulse(dims, (0.5 * (dims[0] - 1), 0.5 * (dims[1] - 1), 0.5 * (dims[2] - 1)), 1e-2, 4.0, q)
    sb = torch.empty(lmx, dtype=torch.float64, device="cuda")
    cs.boundary_weights(dims, sb)
    sv.sbcc = sb
    q0 = host(q).copy()
    for _ in range(4):
        sv.rk4_step(q)
    qh = host(q)
    xyz = cs.config_grid(cfg)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    for name, qq in (("output_data0.vtk", q0), ("output_data1.vtk", qh)):
        d, pts, fields = read_vtk(str(tmp_path / name))
        assert d == dims
        assert np.array_equal(pts, f32(xyz.reshape(3, lmx).T))
        rho = qq[:lmx]
        u = qq[lmx:2 * lmx] / rho + 0.5
        w = qq[3 * lmx:4 * lmx] / rho + 0.1
        p = 0.4 * (qq[4 * lmx:] - 0.5 * (qq[lmx:2 * lmx] ** 2 + qq[2 * lmx:3 * lmx] ** 2 + qq[3 * lmx:4 * lmx] ** 2) / rho)
        assert np.array_equal(fields["density"], f32(rho)), name
        for k, ref in (("u", u), ("w", w), ("p", p)):
            assert np.max(np.abs(fields[k] - ref)) <= 1e-6 * np.max(np.abs(ref)), (name, k)
    assert not np.array_equal(q0, qh)
