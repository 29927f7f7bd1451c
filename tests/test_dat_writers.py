"""CPU: nest_amd.dat writes .dat files byte-identical to the reference's own
example writers (examples/sod1d/sod1d.cpp:50-91, examples/sod2d/sod2d.cpp:
58-152).

* against golden files the reference writers produced
  (tests/golden/gen_dat_golden.py): 1D with the p >= 0 clamp active in one
  cell, 2D on an uneven 3 x 2 patch partition (patch output order), and the
  x-slice;
* where oracle/_ref is built (the reference compiled from its source, build
  container only), against the reference writers live on further seeded
  states and partitions.
"""
import ctypes as C
import os

import numpy as np
import pytest

from nest_amd import dat

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "dat")
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref")


def _read(path):
    with open(path, "rb") as f:
        return f.read()


def test_writers_match_reference_golden_files(tmp_path):
    z = np.load(os.path.join(GOLD, "inputs.npz"))
    p1 = dat.write_sod1d(str(tmp_path), z["G1"], float(z["t1"]), int(z["step1"]), x_min=float(z["x_min1"]),
                         x_max=float(z["x_max1"]), gamma=float(z["gamma"]))
    assert _read(p1) == _read(os.path.join(GOLD, "sod_0007.dat"))
    p2 = dat.write_sod2d(str(tmp_path), z["G2"], float(z["t2"]), int(z["step2"]), npx=int(z["npx"]),
                         npy=int(z["npy"]), x_min=float(z["x_min2"]), x_max=float(z["x_max2"]),
                         y_min=float(z["y_min2"]), y_max=float(z["y_max2"]), gamma=float(z["gamma"]))
    assert _read(p2) == _read(os.path.join(GOLD, "sod2d_0003.dat"))
    p3 = dat.write_sod2d_xslice(str(tmp_path), z["G2"], float(z["t2"]), int(z["step2"]),
                                x_min=float(z["x_min2"]), x_max=float(z["x_max2"]), gamma=float(z["gamma"]))
    assert _read(p3) == _read(os.path.join(GOLD, "sod2d_xslice_0003.dat"))


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "libref_io2d.so")),
                    reason="reference writers not built here (oracle/_ref, build container only)")
@pytest.mark.parametrize("seed, nx, ny, npx, npy", [(1, 16, 12, 2, 3), (2, 9, 17, 4, 1), (3, 25, 10, 5, 4)])
def test_writers_match_reference_live(tmp_path, seed, nx, ny, npx, npy):
    dp, i, d, s = C.POINTER(C.c_double), C.c_int, C.c_double, C.c_char_p
    l1 = C.CDLL(os.path.join(REF, "libref_io1d.so"))
    l2 = C.CDLL(os.path.join(REF, "libref_io2d.so"))
    l1.ref_write_sod1d.argtypes = [dp, i, i, d, d, d, d, i, s]
    l2.ref_write_sod2d.argtypes = [dp, i, i, i, i, d, d, d, d, d, d, i, s]
    r = np.random.default_rng(seed)
    n = nx * ny
    G1 = np.stack([r.uniform(0.1, 2, n), r.uniform(-1, 1, n), r.uniform(0.5, 3, n)])
    G2 = np.stack([r.uniform(0.1, 2, (nx, ny)), r.uniform(-1, 1, (nx, ny)), r.uniform(-1, 1, (nx, ny)),
                   r.uniform(0.5, 3, (nx, ny))])
    ref, ours = tmp_path / "ref", tmp_path / "ours"
    ref.mkdir()
    ours.mkdir()
    t, gamma = 0.1 * seed + 1e-7, 1.4 + 0.1 * seed
    a, b = np.ascontiguousarray(G1), np.ascontiguousarray(G2)
    assert l1.ref_write_sod1d(a.ctypes.data_as(dp), n, npx, 0.0, 2.0, gamma, t, seed, str(ref).encode()) == 0
    assert l2.ref_write_sod2d(b.ctypes.data_as(dp), nx, ny, npx, npy, -1.0, 1.0, 0.0, 0.5, gamma, t, seed,
                              str(ref).encode()) == 0
    dat.write_sod1d(str(ours), G1, t, seed, x_min=0.0, x_max=2.0, gamma=gamma)
    dat.write_sod2d(str(ours), G2, t, seed, npx=npx, npy=npy, x_min=-1.0, x_max=1.0, y_min=0.0, y_max=0.5,
                    gamma=gamma)
    dat.write_sod2d_xslice(str(ours), G2, t, seed, x_min=-1.0, x_max=1.0, gamma=gamma)
    for name in sorted(os.listdir(ref)):
        assert _read(ours / name) == _read(ref / name), name
