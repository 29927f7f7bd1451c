"""A NaN in the state stops a device-dt run with the non-finite status, in
both numerics modes, on every 2D step path and both 1D step kernels.

The fast kernels are compiled with NaNs assumed absent (hardware min / max
without NaN quieting), so their dt and wave-speed checks test the bit
pattern; this pins that the check still fires. The NaN reaches the check
through the wave-speed slots (u64 maxima of the speeds' bit patterns, in
which a NaN ranks above every finite speed): the run records done = 2 and
the status call returns NEST_ERR_NONFINITE.
"""
import numpy as np
import pytest

from nest_amd import Sim2D, Solver1D, problems as P
from nest_amd._lib import NestError
from tests.gpu_util import STEP_PATHS, configure, print_case

pytestmark = pytest.mark.gpu
NONFINITE = 4


def _expect_nonfinite(status):
    with pytest.raises(NestError) as e:
        status()
    assert e.value.code == NONFINITE


@pytest.mark.parametrize("path", STEP_PATHS)
@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("when", ["initial", "after_steps"])
def test_2d_nan_stops_device_dt_run(mode, path, when):
    nx, ny, npx, npy = 96, 80, 2, 2
    print_case("2d nonfinite", mode, path, when)
    sim = Sim2D.mesh(nx, ny, npx, npy, mode=mode)
    configure(sim, path)
    G = P.sod2d(nx, ny)
    if when == "initial":
        G[3, 40, 33] = np.nan  # energy of one interior cell
        sim.upload(sim.scatter(G))
        sim.start(cfl=0.4)
        sim.advance(3)
    else:
        sim.upload(sim.scatter(G))
        sim.start(cfl=0.4)
        sim.advance(2)
        assert sim.status()["done"][0] == 0
        U = sim.download()
        Gd = sim.gather(U)
        Gd[0, 17, 61] = np.nan  # density of one interior cell
        sim.upload(sim.scatter(Gd))
        sim.advance(3)
    _expect_nonfinite(sim.status)
    sim.close()


@pytest.mark.parametrize("kernel", ["tile", "lanes", "pairs"])
@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_1d_nan_stops_device_dt_run(mode, kernel):
    n = 500
    print_case("1d nonfinite", mode, kernel)
    s = Solver1D(n, 1.0 / n, bc="outflow", mode=mode)
    s.set_step_kernel(kernel)
    U = P.sod1d(n)
    U.reshape(3, -1)[2, 2 + 123] = np.nan  # energy of one interior cell
    s.upload(U)
    s.start(cfl=0.4, t_end=0.2)
    s.advance(4)
    _expect_nonfinite(s.status)
