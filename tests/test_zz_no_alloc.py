"""GPU: the 1D step entry point that allocates nothing
(nest_hip_euler1d_step_noalloc), through the C ABI, against the existing
entry points it is a variant of (nest_hip_solver1d_start +
nest_hip_solver1d_advance(1) of a solver handle): the same output state,
ghost cells included, and the same dt, bit for bit, in both numerics modes
and for every boundary condition; two calls in a row, the second on a
smaller state (the shared scratch array keeps the first call's contents,
which the second must not see); sizes beyond the limit are refused before
any device work.
"""
import ctypes as C

import numpy as np
import pytest

from nest_amd import Solver1D, _lib, problems as P
from tests.gpu_util import print_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _noalloc(L, U, n, bc, lim, mode, cfl, t, t_end):
    dev_in = torch.from_numpy(U).cuda()
    dev_out = torch.full_like(dev_in, -7.0)
    dt = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = L.nest_hip_euler1d_step_noalloc(C.c_void_p(dev_in.data_ptr()), C.c_void_p(dev_out.data_ptr()), n, 1.0 / n,
                                         1.4, _lib.BC[bc], _lib.LIMITER[lim], _lib.MODE[mode], cfl, t, t_end,
                                         C.c_void_p(dt.data_ptr()), None)
    _lib.check(rc)
    torch.cuda.synchronize()
    return dev_out.cpu().numpy(), float(dt.item())


def _handle(U, n, bc, lim, mode, cfl, t_end):
    s = Solver1D(n, 1.0 / n, bc=bc, limiter=lim, mode=mode)
    s.upload(U)
    s.start(cfl, t_end)
    s.advance(1)
    out = s.download()
    st = s.status()
    s.close()
    return out, st["dt"]


@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("bc", ["periodic", "outflow", "reflecting"])
def test_noalloc_step_equals_solver_handle(mode, bc):
    L = _lib.load()
    for n, lim, t_end in ((1000, "minmod", float("inf")), (257, "mc", 1e-4)):
        print_case("noalloc 1d step", mode, bc, n, lim, t_end)
        U = P.sod1d(n)
        got, dt = _noalloc(L, U, n, bc, lim, mode, 0.4, 0.0, t_end)
        want, dt_want = _handle(U, n, bc, lim, mode, 0.4, t_end)
        assert dt == dt_want and dt > 0.0, (dt, dt_want)
        assert np.array_equal(got, want), (n, np.abs(got - want).max())
        if np.isfinite(t_end):
            assert dt == min(dt, t_end)


def test_noalloc_limits_and_errors():
    L = _lib.load()
    print_case("noalloc limits")
    a = torch.zeros(3 * 8, dtype=torch.float64, device="cuda")
    b = torch.zeros_like(a)
    for n in (1, (1 << 28) + 1):
        rc = L.nest_hip_euler1d_step_noalloc(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), n, 0.1, 1.4, 0, 0,
                                             0, 0.4, 0.0, 1.0, None, None)
        assert rc == 1 and str(1 << 28) in _lib.last_error()
    rc = L.nest_hip_euler1d_step_noalloc(C.c_void_p(a.data_ptr()), C.c_void_p(a.data_ptr()), 4, 0.1, 1.4, 0, 0, 0,
                                         0.4, 0.0, 1.0, None, None)
    assert rc == 1 and "distinct" in _lib.last_error()
    torch.cuda.synchronize()
    assert torch.count_nonzero(b) == 0  # nothing launched


def test_noalloc_two_calls_second_smaller():
    """Two calls in a row on one GPU: n = 1000 (time 0, no end), then
    n = 300 at a later time with an end time that clamps dt; each equals the
    handle path from the same input."""
    L = _lib.load()
    print_case("noalloc two calls")
    U1 = P.sod1d(1000)
    got1, dt1 = _noalloc(L, U1, 1000, "outflow", "minmod", "exact", 0.4, 0.0, float("inf"))
    want1, dtw1 = _handle(U1, 1000, "outflow", "minmod", "exact", 0.4, float("inf"))
    assert dt1 == dtw1 and np.array_equal(got1, want1)
    U2 = P.sod1d(300, right=(0.25, 0.0, 0.3))
    t0, t_end = 0.05, 0.05 + 1e-5
    got2, dt2 = _noalloc(L, U2, 300, "periodic", "minmod", "exact", 0.4, t0, t_end)
    # the handle path from the same state: its clock starts at 0, so give it
    # the same remaining time
    want2, dtw2 = _handle(U2, 300, "periodic", "minmod", "exact", 0.4, t_end - t0)
    assert dt2 == t_end - t0 and dtw2 == t_end - t0
    assert np.array_equal(got2, want2)
