"""The C-ABI library loads and exports exactly what include/compact_schemes.h
declares (no GPU needed: nothing is called that touches a device)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "compact_schemes.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cs_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    import compact_schemes_amd as cs
    assert declared() == sorted(cs.EXPORTS)


def test_library_exports_every_declared_symbol():
    import compact_schemes_amd as cs
    if not os.path.exists(cs.LIB_PATH):
        pytest.fail("libcompact_schemes_amd.so is not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", cs.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(cs_[a-z0-9_]+)$", out, flags=re.M))
    missing = [s for s in declared() if s not in exported]
    assert not missing, f"not exported: {missing}"
    lib = cs.lib()                                   # loads, binds every signature
    assert lib.cs_version().decode().startswith("compact_schemes_amd")


def test_cpp_header_compiles():
    """The header-only C++ mirror of the reference classes compiles against the C header."""
    src = os.path.join(ROOT, "include", "compact_schemes.hpp")
    if not os.path.exists("/opt/rocm/include/hip/hip_runtime_api.h"):
        pytest.skip("HIP headers not installed")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
                        "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", "-x", "c++", src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_mirror_programs_build():
    """The C++ programs written against the mirror classes link the library."""
    bins = [os.path.join(ROOT, "tests", "cpp", "_bin", n) for n in ("test_numerics_mirror", "test_rhs_mirror")]
    missing = [b for b in bins if not os.path.exists(b)]
    assert not missing, f"not built (run __graft_entry__.build()): {missing}"
    out = subprocess.run(["ldd", bins[0]], capture_output=True, text=True).stdout
    assert "libcompact_schemes_amd.so" in out and "not found" not in out.split("libcompact_schemes_amd.so")[1].split("\n")[0]
