"""Helpers shared by the GPU tests (no GPU work at import time)."""
import numpy as np


def interior_global(sim, U):
    return sim.gather(U)


def print_case(*args):
    print("CASE:", *args, flush=True)


def rel_linf_vars(a, b):
    """max over variables of max|a-b|/max|b| for arrays shaped [4, ...] or [3, ...]."""
    out = 0.0
    for k in range(a.shape[0]):
        den = np.max(np.abs(b[k]))
        num = np.max(np.abs(a[k] - b[k]))
        out = max(out, num / den if den > 0 else num)
    return out
