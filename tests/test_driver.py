"""The run driver's argument and configuration errors (host code only: it
exits before touching a GPU)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CANARD = os.path.join(ROOT, "compact_schemes_amd", "driver", "_bin", "canard")


def _run(args, env=None):
    if not os.path.exists(CANARD):
        pytest.fail("the driver is not built (run __graft_entry__.build())")
    e = dict(os.environ)
    e.pop("CANARD_ROOT", None)
    e.update(env or {})
    return subprocess.run([CANARD] + args, capture_output=True, text=True, env=e, timeout=60)


def test_usage_without_configuration():
    r = _run([])
    assert r.returncode != 0 and "usage: canard" in r.stderr


def test_configuration_errors_exit_with_the_message(tmp_path):
    bad = tmp_path / "config.yaml"
    bad.write_text("driver:\n  - nblocks: zero\n")
    r = _run([str(bad)])
    assert r.returncode != 0 and ":2: not an integer" in r.stderr
    r = _run([], env={"CANARD_ROOT": str(tmp_path / "nowhere")})
    assert r.returncode != 0 and "cannot open run configuration" in r.stderr
