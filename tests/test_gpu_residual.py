"""GPU parity of the fused residual kernel against the CPU oracle, through
the C ABI.  Tolerance: |r_gpu - r_cpu| <= 1e-12 (|r_cpu| + s), s = sum of the
magnitudes of the contributions to the entry (the entry's floating-point
scale), i.e. 1e-12 relative in FP64."""
import numpy as np
import pytest

from pdes_amd import abi, euler, problems
from helpers import assert_parity, boundary_flux_sum, setup, state_for

pytestmark = pytest.mark.gpu

CASES = [("vortex", 5, 1, "omega", "roe"), ("vortex", 5, 2, "omega", "roe"),
         ("vortex", 4, 3, "omega", "roe"), ("vortex", 4, 4, "omega", "roe"),
         ("vortex", 5, 1, "diage", "roe"), ("vortex", 5, 2, "diage", "roe"),
         ("vortex", 4, 3, "diage", "roe"), ("vortex", 4, 4, "diage", "roe"),
         ("vortex", 5, 1, "diage", "ir"), ("vortex", 5, 2, "diage", "ir"),
         ("vortex", 4, 3, "diage", "ir"), ("vortex", 4, 4, "diage", "ir"),
         ("cube", 2, 1, "omega", "roe"), ("cube", 2, 2, "omega", "roe"),
         ("vortex", 4, 1, "omega", "es"), ("vortex", 4, 2, "omega", "es"),
         ("vortex", 3, 3, "omega", "es"), ("vortex", 3, 4, "omega", "es"),
         ("cube", 2, 1, "omega", "es"), ("cube", 2, 2, "omega", "es")]


@pytest.mark.parametrize("kind,n,p,opk,flux", CASES)
def test_residual_parity(oracle, kind, n, p, opk, flux):
    print(f"case {kind} n={n} p={p} {opk} {flux}", flush=True)
    m, op, ph, b = setup(kind, n, p, opk, flux)
    q = state_for(kind, m, ph)
    ref = oracle.residual(b, q)
    scale = oracle.residual_scale(b, q)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    euler.evalResidual(m, op, eqn, None)
    assert_parity(eqn.res, ref, scale, 1e-12, f"{kind} p={p} {opk} {flux}")
    # strong-form norm (calcNorm) matches
    nrm = euler.calcNorm(eqn)
    ref_n = np.sqrt((ref ** 2 * eqn.Minv).sum())
    assert abs(nrm - ref_n) <= 1e-12 * ref_n
    eqn.free()


@pytest.mark.parametrize("kind,n,p,opk,flux", [CASES[1], CASES[5], CASES[9], CASES[11], CASES[13], CASES[15],
                                              CASES[19]])
def test_jv_matches_complex_step(oracle, kind, n, p, opk, flux):
    """evaldRdqProduct on the device vs the oracle's complex step (h=1e-20)."""
    print(f"case {kind} n={n} p={p} {opk} {flux}", flush=True)
    m, op, ph, b = setup(kind, n, p, opk, flux)
    q = state_for(kind, m, ph)
    v = np.random.default_rng(7).standard_normal(q.shape)
    ref = oracle.jv_complex_step(b, q, v)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    got = euler.evaldRdqProduct(eqn, v)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err
    eqn.free()


def test_residual_deterministic():
    """Owner-pull without atomics: two evaluations are bitwise identical."""
    m, op, ph, b = setup("vortex", 16, 2, "diage", "ir")
    q = state_for("vortex", m, ph)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    euler.evalResidual(m, op, eqn)
    r1 = eqn.res.copy()
    euler.evalResidual(m, op, eqn)
    assert np.array_equal(r1, eqn.res)
    eqn.free()


def test_positivity_error():
    """checkDensity/checkPressure: a negative density raises PDES_ERR_PHYS."""
    from pdes_amd._lib import PdesError
    m, op, ph, b = setup("vortex", 4, 1, "omega", "roe")
    q = state_for("vortex", m, ph)
    q[3, 1, 0] = -0.5
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    with pytest.raises(PdesError) as ei:
        euler.evalResidual(m, op, eqn)
    assert ei.value.code == -4
    # the flag is cleared: a valid state evaluates again
    q[3, 1, 0] = 1.0
    eqn.set_q(q)
    euler.evalResidual(m, op, eqn)
    eqn.free()


@pytest.mark.parametrize("p,flux", [(2, "roe"), (4, "roe"), (2, "ir"), (4, "ir")])
def test_full_size_freestream_and_conservation(oracle, p, flux):
    """Config-1 size (256x256, 131072 triangles), properties that hold at any
    size: (a) uniform flow with freestream BC gives a zero residual; (b) on the
    vortex state with vortex BC the residual summed over all nodes equals
    minus the boundary-flux integral sum_b sum_k B_k F_Roe(q_b, q_bc; n_k):
    the volume terms sum to zero per element (Q_d 1 = 0, S skew) and each
    interior face adds equal and opposite contributions to its two elements."""
    print(f"full size p={p} {flux}", flush=True)
    m, op, ph, b = setup("vortex", 256, p, "diage", flux, Ma=0.5, aoa_deg=5.0)
    ph_fs = abi.Physics(**{**ph.__dict__, "bc_type": [abi.PDES_BC_FREESTREAM] * m.numBC})
    eqn = euler.EulerData(m, op, ph_fs)
    eqn.set_q(problems.freestream(m.coords.shape[:-1], 2, ph_fs))
    euler.evalResidual(m, op, eqn)
    assert np.abs(eqn.res).max() < 1e-12
    eqn.free()
    q = problems.vortex_fourier_state(m, ph, seed=1, amp=1e-2)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    euler.evalResidual(m, op, eqn)
    tot = eqn.res.sum(axis=(0, 1))
    mag = np.abs(eqn.res).sum(axis=(0, 1))
    eqn.free()
    bsum = boundary_flux_sum(oracle, m, op, ph, q)
    print("residual total", tot, "boundary flux", bsum, "sum |res|", mag)
    assert np.all(np.abs(tot + bsum) <= 1e-13 * mag + 1e-12 * np.abs(bsum))
    assert np.abs(bsum).max() > 1e-6 * mag.max()


RK_VARIANTS = [(2, "roe", 6), (4, "roe", 4), (2, "ir", 6), (4, "ir", 4)]


@pytest.mark.parametrize("p,flux,n", RK_VARIANTS)
def test_rk4_matches_oracle(oracle, p, flux, n):
    """rk4 on the device (4 residuals per step with the stage update fused into
    the residual kernel's epilogue, rk4.jl:144-344) against the same RK4
    driven by the oracle residual, for the four diagonal-E variants the
    benchmark steps and every launch variant of the RK-stage kernel."""
    from pdes_amd._lib import check, lib
    m, op, ph, b = setup("vortex", n, p, "diage", flux)
    q0 = state_for("vortex", m, ph)
    h = 1e-3
    eqn = euler.EulerData(m, op, ph)
    q = q0.copy()
    f = lambda x: eqn.Minv * oracle.residual(b, x)
    for _ in range(3):
        k1 = f(q)
        k2 = f(q + h / 2 * k1)
        k3 = f(q + h / 2 * k2)
        k4 = f(q + h * k3)
        q = q + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    first = None
    for var in range(8 if flux == "roe" else 4):
        print(f"case p={p} {flux} n={n} rk variant {var}", flush=True)
        check(lib().pdes_set_launch_variant(eqn.ctx, 0, var))
        eqn.set_q(q0)
        euler.rk4(eqn, 3, h)
        got = eqn.get_q()
        assert np.abs(got - q).max() < 1e-12 * np.abs(q).max(), var
        if first is None:
            first = got
        else:
            assert np.array_equal(got, first), f"RK variant {var} differs from variant 0"
    eqn.free()


def _per_node_metric_problem(kind, n, p, opk, flux, seed=5):
    """Perturbs every node's dxidx by a factor 1 + 1e-3 U(0, 1), so the
    context stores one metric per node (curved-element layout) and the kernels
    take their per-node paths."""
    m, op, ph = problems.make_problem(kind, n, p, opk, flux, Ma=0.5, aoa_deg=2.0)
    rng = np.random.default_rng(seed)
    m.dxidx = np.ascontiguousarray(m.dxidx * (1 + 1e-3 * rng.random(m.dxidx.shape[:2])[:, :, None, None]))
    return m, op, ph, abi.make_descriptors(m, op, ph)


PER_NODE_CASES = [("vortex", 4, 2, "diage", "ir"), ("vortex", 3, 4, "diage", "ir"), ("vortex", 4, 2, "diage", "roe"),
                  ("vortex", 4, 2, "omega", "roe"), ("vortex", 3, 2, "omega", "es"), ("cube", 2, 2, "omega", "roe"),
                  ("cube", 1, 2, "omega", "es")]


@pytest.mark.parametrize("kind,n,p,opk,flux", PER_NODE_CASES)
def test_per_node_metric_parity(oracle, kind, n, p, opk, flux):
    """Per-node metrics (the curved-mesh layout the ABI accepts): residual,
    complex-step J v and the element-wise FP64 Jacobian against the oracle,
    which always reads dxidx per node.  The split form takes node j's metric
    for the pair (j, k), j > k, as calcVolumeIntegralsSplitFormLinear does
    (euler_funcs.jl:240-288)."""
    from pdes_amd import linsolve
    print(f"case {kind} n={n} p={p} {opk} {flux} per-node metric", flush=True)
    m, op, ph, b = _per_node_metric_problem(kind, n, p, opk, flux)
    q = state_for(kind, m, ph)
    ref = oracle.residual(b, q)
    scale = oracle.residual_scale(b, q)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    euler.evalResidual(m, op, eqn, None)
    assert_parity(eqn.res, ref, scale, 1e-12, f"{kind} p={p} {opk} {flux} per-node metric")
    v = np.random.default_rng(7).standard_normal(q.shape)
    jv_ref = oracle.jv_complex_step(b, q, v)
    jv = euler.evaldRdqProduct(eqn, v)
    assert np.abs(jv - jv_ref).max() < 1e-12 * np.abs(jv_ref).max()
    linsolve.evalJacobian(m, op, eqn, fp64_blocks=True)
    got = linsolve.getJacobianValues(eqn, fp64_blocks=True)
    rowptr, colind = oracle.sparsity_bsr(b)
    jref = oracle.jacobian_cs(b, q, rowptr, colind)
    assert np.abs(got - jref).max() <= 1e-12 * np.abs(jref).max()
    eqn.free()


VARIANT_CASES = [("vortex", 6, 2, "diage", "ir"), ("vortex", 6, 2, "diage", "roe"), ("vortex", 4, 4, "diage", "roe"),
                 ("vortex", 5, 3, "diage", "roe"), ("vortex", 5, 1, "diage", "roe"), ("vortex", 5, 4, "omega", "roe"),
                 ("vortex", 5, 2, "omega", "roe"), ("cube", 2, 2, "omega", "roe"), ("cube", 2, 1, "omega", "roe")]


@pytest.mark.parametrize("kind,n,p,opk,flux", VARIANT_CASES)
def test_every_launch_variant_parity(oracle, kind, n, p, opk, flux):
    """The autotuner's launch variants all give the oracle's residual, with
    the same bits: 0-3 (elements per workgroup x waves per SIMD) and, for the
    weak form, 4-7, the same shapes with Q^T F on the FP64 matrix cores (the
    matrix instruction accumulates its products in the vector loop's order,
    each fused, so the result is the vector loop's; Omega p=2/tet shapes keep
    the product in an LDS array of its own, the others lay it over the flux
    planes).  The split form runs variant v & 3 for v >= 4.  autotune picks
    a valid variant."""
    from pdes_amd._lib import check, lib
    print(f"case {kind} n={n} p={p} {opk} {flux}", flush=True)
    m, op, ph, b = setup(kind, n, p, opk, flux)
    q = state_for(kind, m, ph)
    ref = oracle.residual(b, q)
    scale = oracle.residual_scale(b, q)
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q)
    first = None
    for var in range(8):
        check(lib().pdes_set_launch_variant(eqn.ctx, var, var))
        euler.evalResidual(m, op, eqn)
        assert_parity(eqn.res, ref, scale, 1e-12, f"variant {var}")
        # the launch shape moves data between LDS slots only: same bits
        if first is None:
            first = eqn.res.copy()
        else:
            assert np.array_equal(eqn.res, first), f"variant {var} differs from variant 0"
    chosen = euler.autotune(eqn, 2)
    assert all(0 <= v < (8 if flux == "roe" else 4) for v in chosen)
    assert lib().pdes_set_launch_variant(eqn.ctx, 8, 0) != 0
    eqn.free()
