"""Golden .dat files from the reference's own writers (run in the build
container only, where /root/reference is mounted and oracle/_ref is built).

oracle/_ref/libref_io{1d,2d}.so are the reference's examples/sod1d/sod1d.cpp
and examples/sod2d/sod2d.cpp compiled from their sources with `main` renamed
(recipe: oracle/Makefile, target `ref`). This script feeds them small seeded
states and stores, under tests/golden/dat/:
  * inputs.npz — the states and parameters;
  * the files the reference writers produced (sod_0007.dat,
    sod2d_0003.dat, sod2d_xslice_0003.dat).
tests/test_dat_writers.py checks nest_amd.dat against them byte for byte.
"""
import ctypes as C
import os
import shutil
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "dat")
_dp, _i, _d, _s = C.POINTER(C.c_double), C.c_int, C.c_double, C.c_char_p


def states():
    r = np.random.default_rng(20)
    n = 37
    rho = r.uniform(0.1, 2.0, n)
    u = r.uniform(-1.0, 1.0, n)
    p = r.uniform(0.01, 2.0, n)
    E = p / 0.4 + 0.5 * rho * u * u
    E[5] = 0.4 * rho[5] * u[5] * u[5]  # E below the kinetic energy: the p >= 0 clamp
    G1 = np.stack([rho, rho * u, E])
    nx, ny = 11, 9
    rho = r.uniform(0.1, 2.0, (nx, ny))
    u = r.uniform(-1.0, 1.0, (nx, ny))
    v = r.uniform(-1.0, 1.0, (nx, ny))
    p = r.uniform(0.01, 2.0, (nx, ny))
    G2 = np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])
    return G1, G2


def main():
    G1, G2 = states()
    params = dict(n1=37, npatch1=3, x_min1=-0.5, x_max1=1.5, t1=0.123456789, step1=7,
                  npx=3, npy=2, x_min2=0.0, x_max2=1.0, y_min2=0.0, y_max2=0.1, t2=0.0625, step2=3, gamma=1.4)
    os.makedirs(OUT, exist_ok=True)
    l1 = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_io1d.so"))
    l2 = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_io2d.so"))
    l1.ref_write_sod1d.argtypes = [_dp, _i, _i, _d, _d, _d, _d, _i, _s]
    l2.ref_write_sod2d.argtypes = [_dp, _i, _i, _i, _i, _d, _d, _d, _d, _d, _d, _i, _s]
    tmp = tempfile.mkdtemp()
    try:
        a = np.ascontiguousarray(G1)
        assert l1.ref_write_sod1d(a.ctypes.data_as(_dp), 37, 3, -0.5, 1.5, 1.4, 0.123456789, 7, tmp.encode()) == 0
        b = np.ascontiguousarray(G2)
        assert l2.ref_write_sod2d(b.ctypes.data_as(_dp), 11, 9, 3, 2, 0.0, 1.0, 0.0, 0.1, 1.4, 0.0625, 3,
                                  tmp.encode()) == 0
        for name in ("sod_0007.dat", "sod2d_0003.dat", "sod2d_xslice_0003.dat"):
            shutil.copy(os.path.join(tmp, name), os.path.join(OUT, name))
    finally:
        shutil.rmtree(tmp)
    np.savez(os.path.join(OUT, "inputs.npz"), G1=G1, G2=G2, **{k: np.array(v) for k, v in params.items()})
    print("wrote", sorted(os.listdir(OUT)))


if __name__ == "__main__":
    main()
