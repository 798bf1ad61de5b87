"""GPU parity of the C++ operator classes (include/compact_schemes.hpp), which
keep the reference's host API (numerics_pc, physics, gcbc, gcbc_go).

The programs under tests/cpp are written the way the reference's own tests
are (tests/test_numerics_serial.cu, tests/test_physics.cu); they run as child
processes and talk to this test through files.
"""
import os
import subprocess

import numpy as np
import pytest

from gpu_util import case, nrel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_bin")


def _run(args, timeout=120):
    exe = os.path.join(BIN, args[0])
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    r = subprocess.run([exe] + [str(a) for a in args[1:]], capture_output=True, text=True, timeout=timeout)
    print(r.stdout, r.stderr, flush=True)
    return r


@pytest.mark.gpu
def test_numerics_pc_ramp_kat():
    """deriv1d<Axis> of the linear-index ramp on 256x128x64 with closures on
    every face: constant derivative 1, lxi, lxi*let to 1e-12 (the reference
    asserts 1e-6 in float)."""
    case("c++ mirror ramp KAT")
    r = _run(["test_numerics_mirror"])
    assert r.returncode == 0 and "PASS" in r.stdout


CL = [10, 10, 10, 10, 10, 10]
PER = [45, 45, 45, 45, 45, 45]
MIX = [45, 10, 10, 45, 10, 10]


@pytest.mark.gpu
@pytest.mark.parametrize("dims,nbc,visc", [((32, 28, 24), CL, 1), ((24, 20, 18), MIX, 1), ((20, 24, 16), PER, 0),
                                           ((40, 16, 20), CL, 0)])
def test_rhs_through_cpp_classes_matches_oracle(oracle, tmp_path, dims, nbc, visc):
    from compact_schemes_amd import workloads as wl
    st = wl.flow_state(dims, seed=11, noise=1e-2, width=3.0)
    lmx = dims[0] * dims[1] * dims[2]
    for k in ("qa", "xim", "etm", "zem", "yaco", "ssk"):
        np.ascontiguousarray(st[k], dtype=np.float64).tofile(tmp_path / k)
    for a in range(3):
        np.ascontiguousarray(st["cm"][a], dtype=np.float64).tofile(tmp_path / f"cm{a}")
    case(f"c++ mirror rhs dims={dims} nbc={nbc} visc={visc}")
    r = _run(["test_rhs_mirror", tmp_path, *dims, visc, *nbc, *st["umf"]])
    assert r.returncode == 0, r.stderr
    de = np.fromfile(tmp_path / "de", dtype=np.float64)
    ctx = oracle.Context(dims, [nbc[:3], nbc[3:]])
    ref, p = np.zeros(5 * lmx), np.zeros(lmx)
    ctx.rhs(visc, st["qa"], ref, p, st["ssk"], st["xim"], st["etm"], st["zem"], st["yaco"], st["cm"],
            wl.boundary_weights(dims), st["umf"])
    for m in range(5):
        assert nrel(de[m * lmx:(m + 1) * lmx], ref[m * lmx:(m + 1) * lmx]) < 1e-12, f"variable {m}"
