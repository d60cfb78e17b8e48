"""Shared test helpers (problem setup, tolerance checks)."""
import numpy as np

from pdes_amd import abi, problems


def setup(kind, n, p, op_kind="omega", flux="roe", **kw):
    m, op, ph = problems.make_problem(kind, n, p, op_kind, flux, **kw)
    return m, op, ph, abi.make_descriptors(m, op, ph)


def state_for(kind, m, ph, seed=0):
    """Smooth test states with no exactly-zero velocity component (a zero
    normal velocity sits on the |Un| kink of the Roe dissipation, where the
    complex step takes a one-sided derivative)."""
    if kind == "vortex":
        return problems.vortex_fourier_state(m, ph, seed=seed, amp=1e-2)
    q = problems.freestream_bump_state(m, ph, seed=seed)
    rng = np.random.default_rng(seed + 100)
    q[..., 1:m.dim + 1] += 1e-3 * rng.uniform(0.5, 1.0, size=q[..., 1:m.dim + 1].shape)
    return q


def assert_parity(got, ref, scale, rtol=1e-12, what=""):
    """|got - ref| <= rtol * (|ref| + scale) entrywise, where scale is the sum of
    the magnitudes of the contributions to each entry (floating-point scale)."""
    err = np.abs(got - ref)
    bound = rtol * (np.abs(ref) + scale) + 1e-300
    worst = np.max(err / bound)
    assert worst <= 1.0, f"{what}: max err/bound = {worst:.3g}, max abs err {err.max():.3g}"
    return worst
