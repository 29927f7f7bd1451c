"""Device tables (halo plan + work list) checked on the CPU: compiled from
tests/cpp/test_plan.cpp with the library's own plan builder and the oracle."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_plan_tables_bounds_and_fill_semantics(tmp_path):
    exe = str(tmp_path / "test_plan")
    subprocess.run(["g++", "-std=c++20", "-O2", "-ffp-contract=off", "-Wall", "-Wno-unknown-pragmas",
                    os.path.join(HERE, "cpp", "test_plan.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
