"""ctypes binding of ``libnest_hip.so`` (C ABI declared in include/nest_hip.h).

The library is the product: every compute call below goes to the HIP
kernels. There is no CPU fallback — if the library is missing the import of a
solver raises ``NestLibraryMissing``.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnest_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "nest_hip.h")

NEST_OK = 0
STATUS = {0: "OK", 1: "INVALID", 2: "ALLOC", 3: "HIP", 4: "NONFINITE", 5: "UNSUPPORTED"}
BC = {"periodic": 0, "outflow": 1, "reflecting": 2}
LIMITER = {"minmod": 0, "mc": 1, "none": 2}
MODE = {"exact": 0, "fast": 1}


class NestLibraryMissing(RuntimeError):
    pass


class NestError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"nest_hip error {code} ({STATUS.get(code, '?')}): {msg}")
        self.code = code


_vp = C.c_void_p
_d, _i, _ll, _sz = C.c_double, C.c_int, C.c_longlong, C.c_size_t
_pd, _pi = C.POINTER(C.c_double), C.POINTER(C.c_int)

# name -> (restype, argtypes)
SIGNATURES = {
    "nest_hip_last_error": (C.c_char_p, []),
    "nest_hip_version": (_i, []),
    "nest_hip_device_count": (_i, []),
    "nest_hip_device_copy": (_i, [_vp, _vp, _ll, _vp]),
    "nest_hip_solver1d_create": (_i, [_i, _d, _d, _i, _i, _i, C.POINTER(_vp)]),
    "nest_hip_solver1d_destroy": (_i, [_vp]),
    "nest_hip_solver1d_upload": (_i, [_vp, _pd]),
    "nest_hip_solver1d_download": (_i, [_vp, _pd]),
    "nest_hip_solver1d_compute_dt": (_i, [_vp, _d, _pd]),
    "nest_hip_solver1d_rk2_step": (_i, [_vp, _d]),
    "nest_hip_solver1d_run": (_i, [_vp, _d, _d, _pi]),
    "nest_hip_solver1d_set_stream": (_i, [_vp, _vp]),
    "nest_hip_solver1d_set_step_kernel": (_i, [_vp, _i]),
    "nest_hip_solver1d_stage": (_i, [_vp, _i, _d]),
    "nest_hip_solver1d_exchange": (_i, [_vp, _i]),
    "nest_hip_solver1d_start": (_i, [_vp, _d, _d]),
    "nest_hip_solver1d_advance": (_i, [_vp, _i]),
    "nest_hip_solver1d_status": (_i, [_vp, _pd, _pd, _pi, _pi]),
    "nest_hip_euler1d_step_noalloc": (_i, [_vp, _vp, _i, _d, _d, _i, _i, _i, _d, _d, _d, _vp, _vp]),
    "euler1d_gpu_create": (_vp, [_i, _i, _d, _d]),
    "euler1d_gpu_destroy": (None, [_vp]),
    "euler1d_gpu_upload": (None, [_vp, _pd]),
    "euler1d_gpu_download": (None, [_vp, _pd]),
    "euler1d_gpu_compute_dt": (_d, [_vp, _d]),
    "euler1d_gpu_rk2_step": (None, [_vp, _d]),
    "euler1d_gpu_run": (_i, [_vp, _d, _d]),
    "euler1d_gpu_run_timed": (_d, [_vp, _d, _d, _pi]),
    "nest_hip_mesh2d_create": (_i, [_i, _i, _i, _i, _d, _d, _d, _i, _i, _i, _i, C.POINTER(_vp)]),
    "nest_hip_batch2d_create": (_i, [_i, _pi, _pi, _pd, _pd, _d, _i, _i, _i, C.POINTER(_vp)]),
    "nest_hip_sim2d_destroy": (_i, [_vp]),
    "nest_hip_sim2d_info": (_i, [_vp, C.POINTER(_sz), _pi, _pi, _pi]),
    "nest_hip_sim2d_patch_info": (_i, [_vp, _i, C.POINTER(_ll), _pi, _pi, _pi]),
    "nest_hip_sim2d_set_stream": (_i, [_vp, _vp]),
    "nest_hip_sim2d_bind": (_i, [_vp, _vp, _vp]),
    "nest_hip_sim2d_device_ptrs": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp)]),
    "nest_hip_sim2d_set_tiling": (_i, [_vp, _i, _i, _i]),
    "nest_hip_sim2d_set_stage_tiling": (_i, [_vp, _i, _i, _i]),
    "nest_hip_sim2d_set_stage_kernel": (_i, [_vp, _i]),
    "nest_hip_sim2d_set_step_kernel": (_i, [_vp, _i]),
    "nest_hip_sim2d_get_step_kernel": (_i, [_vp, _pi]),
    "nest_hip_sim2d_get_step_tiling": (_i, [_vp, _pi, _pi]),
    "nest_hip_sim2d_get_tiling": (_i, [_vp, _pi, _pi]),
    "nest_hip_sim2d_set_problem_rows": (_i, [_vp, _i, _pi]),
    "nest_hip_sim2d_upload": (_i, [_vp, _pd]),
    "nest_hip_sim2d_download": (_i, [_vp, _pd]),
    "nest_hip_sim2d_upload_patch": (_i, [_vp, _i, _pd]),
    "nest_hip_sim2d_download_patch": (_i, [_vp, _i, _pd]),
    "nest_hip_sim2d_fill_halos": (_i, [_vp]),
    "nest_hip_sim2d_exchange": (_i, [_vp, _i]),
    "nest_hip_sim2d_compute_dt": (_i, [_vp, _d, _pd]),
    "nest_hip_sim2d_stage": (_i, [_vp, _i, _d, _i]),
    "nest_hip_sim2d_rk2_step": (_i, [_vp, _d]),
    "nest_hip_sim2d_start": (_i, [_vp, _d, _d, _d]),
    "nest_hip_sim2d_advance": (_i, [_vp, _i, _i]),
    "nest_hip_sim2d_status": (_i, [_vp, _pd, _pd, _pi, _pi]),
    "nest_hip_sim2d_advance_per_patch": (_i, [_vp, _i, _d]),
    "nest_hip_sim2d_synchronize": (_i, [_vp]),
    "nest_hip_sim2d_advance_timed": (_i, [_vp, _i, _pd, _pd]),
    "nest_hip_sim2d_global_size": (_i, [_vp, C.POINTER(C.c_longlong)]),
    "nest_hip_sim2d_pack": (_i, [_vp, _vp, _vp, _i]),
    "nest_hip_sim2d_set_adjoint_tiling": (_i, [_vp, _i, _i, _i, _i]),
    "nest_hip_sim2d_adjoint_stage": (_i, [_vp, _vp, _vp, _d, _vp, _d, _vp, _d]),
    "nest_hip_sim2d_rk2_vjp": (_i, [_vp, _vp, _vp, _d]),
    "nest_hip_sim2d_rollout": (_i, [_vp, _i, _d, _i]),
    "nest_hip_sim2d_rollout_adjoint": (_i, [_vp, _vp]),
    "nest_hip_sim2d_rollout_info": (_i, [_vp, _pi, _pi, _pd]),
}

_lib = None


def header_symbols(path: str = HEADER) -> list[str]:
    """Function names declared in include/nest_hip.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b((?:nest_hip|euler1d_gpu)_\w+)\s*\(", text)
    return sorted(set(names))


def load(path: str = LIB_PATH):
    """Load the library (once) and attach signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NestLibraryMissing(
            f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(no CPU fallback exists for the HIP path)")
    # One HIP runtime per process: the torch wheel bundles its own HIP / HSA
    # runtime (same soname), which cannot open the GPU once another copy has.
    # Loading torch first (when installed) makes this library bind to it, so
    # both can share the device and its streams.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return load().nest_hip_last_error().decode()


def check(rc: int) -> None:
    if rc != NEST_OK:
        raise NestError(rc, last_error())
