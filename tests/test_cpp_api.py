"""C++20 host API: CPU tests of core/patch/pipeline and the plan tables, and
the exec::hip driver (tests/cpp/sod2d_pipeline.cpp) on a GPU. Binaries are
built by `make -C tests/cpp` (part of __graft_entry__.build())."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "_bin")


def _build():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp"), "_bin/test_core_api", "_bin/test_plan"],
                   check=True)


def test_core_patch_pipeline_api():
    _build()
    r = subprocess.run([os.path.join(BIN, "test_core_api")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed checks" in r.stdout


def test_adjoint_building_blocks_match_dual_numbers():
    """Reverse-mode prim / limiter / HLLE+passive face functions of the
    adjoint kernels, compiled for the host, against dual-number JVPs of the
    oracle numerics: <w, J v> == <J^T w, v> to 1e-12 over every branch."""
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp"), "_bin/test_adjoint_face"], check=True)
    r = subprocess.run([os.path.join(BIN, "test_adjoint_face")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS" in r.stdout


def test_backend_hip_header_compiles_and_links():
    """exec::hip drivers (2D mesh and 1D patch vectors) build against
    include/nest/backend_hip.hpp and link libnest_hip.so (no GPU needed to
    build)."""
    for name in ("sod2d_pipeline", "sod1d_pipeline"):
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp"), "_bin/" + name], check=True)
        assert os.path.exists(os.path.join(BIN, name))


@pytest.mark.gpu
def test_exec_hip_pipeline_matches_oracle():
    """Per-patch pipeline (cfl_reduce, exchange, rk2_stage<1>, exchange,
    rk2_stage<2>), the same with a user's host Reduce stage giving dt, and
    the fused pipeline on exec::hip, bitwise vs the oracle."""
    exe = os.path.join(BIN, "sod2d_pipeline")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp/_bin/sod2d_pipeline not built")
    print("CASE: exec::hip sod2d pipeline", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
def test_exec_hip_1d_pipeline_matches_oracle():
    """exec::hip on 1D patch vectors (hip::context<1>): per-stage pipeline
    with device dt, with a user's host Reduce stage giving dt, and fused
    device-dt steps, on 4 and 3 periodic patches and one outflow patch:
    bitwise vs the oracle, dt included, and the handed-out halos equal an
    exchange's."""
    exe = os.path.join(BIN, "sod1d_pipeline")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp/_bin/sod1d_pipeline not built")
    print("CASE: exec::hip sod1d pipeline", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr
