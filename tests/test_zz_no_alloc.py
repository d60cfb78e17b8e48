"""pdes_rk4_noalloc (the RK4 driver that allocates nothing, its buffers in a
fixed per-GPU device array) against pdes_rk4 (context-allocated buffers),
through the C ABI: bit-identical states on two calls in a row, the second on
a smaller state, and the size limit rejected with a clear error."""
import ctypes as C

import numpy as np
import pytest

from pdes_amd import euler, problems
from pdes_amd._lib import PdesError, check, lib

pytestmark = pytest.mark.gpu


def _device_state(eqn):
    qd = C.c_void_p()
    check(lib().pdes_state_ptr(eqn.ctx, C.byref(qd), None))
    return qd


def _rk4_both(n, p, flux, nsteps, h):
    m, op, ph = problems.make_problem("vortex", n, p, "diage", flux)
    q0 = problems.vortex_fourier_state(m, ph, seed=1, amp=1e-2)
    eqn = euler.EulerData(m, op, ph)
    try:
        eqn.set_q(q0)
        euler.rk4(eqn, nsteps, h)
        ref = eqn.get_q()
        # the caller's state: a device buffer of numDof doubles (here the
        # context's own state array, reloaded with q0)
        eqn.set_q(q0)
        check(lib().pdes_rk4_noalloc(eqn.ctx, _device_state(eqn), nsteps, h))
        got = eqn.get_q()
    finally:
        eqn.free()
    return ref, got, q0


@pytest.mark.parametrize("flux", ["ir", "roe"])
def test_rk4_noalloc_matches_rk4_on_two_calls(flux):
    h = 2e-4
    for n in (64, 24):   # 393,216 then 55,296 doubles of state
        print(f"case noalloc rk4 n={n} p=2 diage {flux}", flush=True)
        ref, got, q0 = _rk4_both(n, 2, flux, 3, h)
        assert not np.array_equal(ref, q0)
        assert np.array_equal(got, ref), f"n={n}: max diff {np.abs(got - ref).max()}"


def test_rk4_noalloc_rejects_a_state_above_its_array():
    m, op, ph = problems.make_problem("vortex", 150, 2, "diage", "roe")   # 2,160,000 doubles > 2^21
    eqn = euler.EulerData(m, op, ph)
    try:
        with pytest.raises(PdesError) as ei:
            check(lib().pdes_rk4_noalloc(eqn.ctx, _device_state(eqn), 1, 1e-4))
        assert ei.value.code == -1 and "exceeds" in str(ei.value)
    finally:
        eqn.free()
