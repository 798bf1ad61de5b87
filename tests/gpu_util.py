"""Helpers shared by the GPU parity tests."""
import sys

import numpy as np


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def nrel(x, ref):
    """Norm-wise relative difference max|x - ref| / max|ref| (the parity measure)."""
    x = np.asarray(x)
    ref = np.asarray(ref)
    den = np.max(np.abs(ref))
    if den == 0:
        return float(np.max(np.abs(x)))
    return float(np.max(np.abs(x - ref)) / den)


def case(msg):
    """Name the case on stdout (flushed) before a GPU step runs."""
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()
