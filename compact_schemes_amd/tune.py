"""The process-wide cache of launch shapes (csrc/tune.hip).

Three kinds of entries, each keyed by what the kernels' speed depends on:
line kernels (operator, line length, start/end rows, precision) -> chunk
length; fused right-hand-side sweeps (axis, line length, seam, precision,
metrics: 0 constant, 1 per point) -> (variant, lines, xcd); fused filter
sweeps (axis, line length, seam) -> (fields, lines, chunk).  The autotuners
(Batch.autotune, FusedSolver.autotune) fill it; operator tables and solvers built afterwards
read it; save/load keep it across runs as text.
"""
import ctypes

from . import check, lib

DERIV, FILTER = 0, 1


def set_line(op, n, K, start=0, end=0, precision=0):
    """K = 8, 16 or 32 points per chunk (0 clears); start/end: 0 = closure
    rows, 1 = seam halo rows at that end of the line."""
    check(lib().cs_tune_set(int(op), int(n), int(start), int(end), int(precision), int(K)))


def get_line(op, n, start=0, end=0, precision=0):
    k = ctypes.c_int()
    check(lib().cs_tune_get(int(op), int(n), int(start), int(end), int(precision), ctypes.byref(k)))
    return k.value


def set_sweep(axis, n, seam, precision, variant, lines, xcd, metrics=0):
    """metrics: 0 = solvers with constant metrics, 1 = per-point metrics."""
    check(lib().cs_tune_set_sweep(int(axis), int(n), int(seam), int(precision), int(metrics), int(variant),
                                  int(lines), int(xcd)))


def get_sweep(axis, n, seam, precision=0, metrics=0):
    """(variant, lines, xcd) or None."""
    found = ctypes.c_int()
    out = (ctypes.c_int * 3)()
    check(lib().cs_tune_get_sweep(int(axis), int(n), int(seam), int(precision), int(metrics), ctypes.byref(found),
                                  out))
    return tuple(out) if found.value else None


def set_filter_sweep(axis, n, seam, fields, lines, chunk):
    check(lib().cs_tune_set_filter_sweep(int(axis), int(n), int(seam), int(fields), int(lines), int(chunk)))


def get_filter_sweep(axis, n, seam):
    """(fields, lines, chunk) or None."""
    found = ctypes.c_int()
    out = (ctypes.c_int * 3)()
    check(lib().cs_tune_get_filter_sweep(int(axis), int(n), int(seam), ctypes.byref(found), out))
    return tuple(out) if found.value else None


def clear():
    check(lib().cs_tune_clear())


def save(path):
    check(lib().cs_tune_save(str(path).encode()))


def load(path):
    check(lib().cs_tune_load(str(path).encode()))
