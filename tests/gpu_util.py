"""Helpers shared by the GPU tests (no GPU work at import time)."""
import numpy as np


def interior_global(sim, U):
    return sim.gather(U)


def print_case(*args):
    print("CASE:", *args, flush=True)


def rel_linf_vars(a, b):
    """Relative L-infinity per conserved variable, max over variables:
    max|a_k - b_k| / scale_k, arrays shaped [nvars, ...] (rho, momenta..., E).
    scale_k = max|b_k|, except that a momentum's scale is floored at the
    flow's momentum scale sqrt(max rho * max E) (~ rho*c): a momentum that is
    identically zero or orders of magnitude below rho*c (the transverse
    momentum of a planar problem) is otherwise measured against its own
    round-off."""
    out = 0.0
    nv = a.shape[0]
    floor = np.sqrt(np.max(np.abs(b[0])) * np.max(np.abs(b[nv - 1])))
    for k in range(nv):
        den = np.max(np.abs(b[k]))
        if 0 < k < nv - 1:
            den = max(den, floor)
        num = np.max(np.abs(a[k] - b[k]))
        out = max(out, num / den if den > 0 else num)
    return out


# How a 2D RK2 step runs on the GPU: the one-launch step kernel (U1 kept on
# chip; the default), the two stage kernels with rows fetched into LDS, or
# the two register-path stage kernels (used for state buffers of 4 GiB and
# more).
STEP_PATHS = ["one_launch", "two_stage", "register"]


def configure(sim, path):
    """Select one of STEP_PATHS on a Sim2D handle."""
    if path == "one_launch":
        sim.set_stage_kernel("auto")
        sim.set_step_kernel("auto")
        assert sim.one_launch
    elif path == "two_stage":
        sim.set_stage_kernel("lds_ring")
        sim.set_step_kernel("two_stage")
        assert not sim.one_launch
    else:
        sim.set_stage_kernel("register")
        sim.set_step_kernel("two_stage")
        assert not sim.one_launch
