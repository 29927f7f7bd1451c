"""The C-ABI library loads without a GPU and exports every symbol that
include/nest_hip.h declares (no compute calls here)."""
import ctypes as C
import os
import subprocess

import pytest

from nest_amd import _lib


def test_header_declares_expected_entry_points():
    syms = _lib.header_symbols()
    for s in ("nest_hip_mesh2d_create", "nest_hip_sim2d_advance", "euler1d_gpu_create", "euler1d_gpu_run_timed"):
        assert s in syms
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    for s in _lib.header_symbols():
        assert hasattr(L, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in _lib.header_symbols() if s not in exported]
    assert not missing, missing


def test_library_is_gfx950_code_object():
    """The embedded offload bundle targets gfx950 (our own build's fatbin)."""
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_error_reporting_without_device():
    """Argument validation happens before any device call."""
    L = _lib.load()
    h = C.c_void_p()
    rc = L.nest_hip_solver1d_create(1, 0.1, 1.4, 0, 0, 0, C.byref(h))
    assert rc == 1 and "n >= 2" in _lib.last_error()
    rc = L.nest_hip_mesh2d_create(8, 8, 8, 1, 0.1, 0.1, 1.4, 0, 0, 0, 0, C.byref(h))
    assert rc == 1 and "2 interior cells" in _lib.last_error()
    assert L.nest_hip_sim2d_info(None, None, None, None, None) == 1
    rc = L.nest_hip_mesh2d_create(8, 8, 1, 1, 0.1, 0.1, 1.4, 3, 0, 0, 0, C.byref(h))
    assert rc == 1 and "2 (reflecting)" in _lib.last_error()


def test_invalid_arguments_rejected_before_device_work():
    """Every entry point below rejects its bad argument with NEST_ERR_INVALID
    and a message, before any device call (so the checks run on a host with
    no GPU): no exception and no crash crosses the C ABI."""
    L = _lib.load()
    h = C.c_void_p()
    buf = (C.c_double * 8)()
    other = (C.c_double * 8)()
    cases = [
        ("1d gamma", lambda: L.nest_hip_solver1d_create(16, 0.1, 1.0, 0, 0, 0, C.byref(h)), "gamma > 1"),
        ("1d bc", lambda: L.nest_hip_solver1d_create(16, 0.1, 1.4, 3, 0, 0, C.byref(h)), "bc in {0,1,2}"),
        ("1d limiter", lambda: L.nest_hip_solver1d_create(16, 0.1, 1.4, 0, 7, 0, C.byref(h)), "limiter"),
        ("1d out", lambda: L.nest_hip_solver1d_create(16, 0.1, 1.4, 0, 0, 0, None), "out is null"),
        ("1d kernel on null handle", lambda: L.nest_hip_solver1d_set_step_kernel(None, 3), "null handle"),
        ("2d dx", lambda: L.nest_hip_mesh2d_create(8, 8, 1, 1, 0.0, 0.1, 1.4, 0, 0, 0, 0, C.byref(h)), ""),
        ("batch empty", lambda: L.nest_hip_batch2d_create(0, None, None, None, None, 1.4, 0, 0, 0, C.byref(h)), ""),
        ("noalloc aliasing", lambda: L.nest_hip_euler1d_step_noalloc(buf, buf, 4, 0.25, 1.4, 0, 0, 0, 0.4, 0.0, 1.0,
                                                                     None, None), "distinct"),
        ("noalloc n", lambda: L.nest_hip_euler1d_step_noalloc(buf, other, 1, 0.25, 1.4, 0, 0, 0, 0.4, 0.0, 1.0,
                                                              None, None), "n must be in"),
        ("noalloc cfl", lambda: L.nest_hip_euler1d_step_noalloc(buf, other, 4, 0.25, 1.4, 0, 0, 0, 0.0, 0.0, 1.0,
                                                                None, None), "cfl > 0"),
    ]
    for name, call, msg in cases:
        h.value = None
        rc = call()
        err = _lib.last_error()
        assert rc == 1, (name, rc, err)
        assert err and msg in err, (name, err)
        assert not h.value, name  # no handle handed out


def test_adjoint_rollout_rejects_reflecting_before_device_work():
    from nest_amd.adjoint import Rollout
    for make in (lambda: Rollout.mesh(16, 16, bc="reflecting"), lambda: Rollout.batch(2, 16, 16, bc="reflecting")):
        with pytest.raises(ValueError, match="reflecting"):
            make()


def test_no_cpu_fallback_when_library_missing(tmp_path):
    with pytest.raises(_lib.NestLibraryMissing):
        saved = _lib._lib
        _lib._lib = None
        try:
            _lib.load(str(tmp_path / "absent.so"))
        finally:
            _lib._lib = saved


def test_problems_are_decomposition_independent():
    from nest_amd import problems as P
    import numpy as np
    a = P.uniform01(2, np.arange(100, dtype=np.uint64))
    b = np.concatenate([P.uniform01(2, np.arange(0, 37, dtype=np.uint64)),
                        P.uniform01(2, np.arange(37, 100, dtype=np.uint64))])
    assert np.array_equal(a, b) and (a >= 0).all() and (a < 1).all()
    sizes = P.ensemble_sizes()
    assert len(sizes) == 512 and min(sizes) >= 64 and max(sizes) <= 512
