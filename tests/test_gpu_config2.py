"""3D Newton-Krylov in mixed precision (the unit cube split into tets,
SBP-Omega p=2, Roe flux, freestream rho = 1, M = 0.5, 2 degrees angle of
attack, freestream plus a seeded smooth 1e-2 bump): residual, Newton and
GMRES in FP64, Jacobian blocks and the ILU(0) preconditioner in FP32
(newtonInner, src/NonlinearSolvers/newton.jl:137-304, with the PETSc KSP
options of src/input/read_input.jl:560-581).

Small cubes: the converged state against the oracle's Newton with a sparse
direct solve, and against the device run with FP64 blocks.  32^3 cube
(196,608 tets): both precisions reach ||R|| <= 1e-10 and agree."""
import time

import numpy as np
import pytest

from pdes_amd import abi, euler, linsolve, problems

pytestmark = pytest.mark.gpu


def _minv(m, op):
    return np.repeat((1.0 / (op.w[None, :] / m.jac))[:, :, None], m.dim + 2, axis=2)


def _solve(m, op, ph, q0, fp64, restart, pc=abi.PDES_PC_ILU0, krylov_reltol=1e-2, verbose=False):
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q0)
    opts = linsolve.NewtonOptions(res_abstol=1e-10, krylov_reltol=krylov_reltol, pc_kind=pc,
                                  fp64_blocks=fp64, gmres_restart=restart, krylov_itermax=3000)
    opts.itermax = 1
    its = kry = 0
    t0 = time.perf_counter()
    rep = None
    for _ in range(30):
        rep = linsolve.newton(m, op, eqn, opts)
        its += rep.iterations
        kry += rep.total_krylov_iters
        if verbose:
            print(f"fp64_blocks={fp64} step {its}: ||R|| {rep.res_norm:.3e}, {rep.total_krylov_iters} Krylov, "
                  f"{time.perf_counter() - t0:.1f} s", flush=True)
        if rep.converged:
            break
    q = eqn.get_q()
    eqn.free()
    return q, rep, its, kry


@pytest.mark.parametrize("n", [2, 3])
def test_config2_small_cube_matches_oracle_and_fp64(oracle, n):
    print(f"case cube n={n} omega p=2 roe mixed precision", flush=True)
    m, op, ph = problems.make_problem("cube", n, 2, "omega", "roe", Ma=0.5, aoa_deg=2.0)
    b = abi.make_descriptors(m, op, ph)
    q0 = problems.freestream_bump_state(m, ph, seed=2)
    q_ref, norms = oracle.newton_direct(b, q0, _minv(m, op), res_abstol=1e-10)
    assert norms[-1] < 1e-10
    q_mix, rep_mix, its_mix, kry_mix = _solve(m, op, ph, q0, False, 30)
    q_64, rep_64, its_64, kry_64 = _solve(m, op, ph, q0, True, 30)
    print("oracle Newton", len(norms) - 1, "mixed", its_mix, kry_mix, "fp64", its_64, kry_64, flush=True)
    for rep in (rep_mix, rep_64):
        assert rep.converged and rep.res_norm <= 1e-10
    scale = np.abs(q_ref).max()
    assert np.abs(q_mix - q_ref).max() <= 1e-9 * scale
    assert np.abs(q_mix - q_64).max() <= 1e-10 * scale
    assert abs(its_mix - its_64) <= 1
    # the inexact Newton with a 1e-2 forcing term takes more steps than the
    # direct solve, but not many
    assert its_mix <= 2 * (len(norms) - 1) + 2


def test_config2_full_cube_mixed_vs_fp64():
    """32^3 x 6 tets (196,608 elements, 9.8 M DOFs).  GMRES(60): with
    element-block ILU(0) GMRES(30) stagnates on the second Newton system at
    this size."""
    t0 = time.perf_counter()
    m, op, ph = problems.make_problem("cube", 32, 2, "omega", "roe", Ma=0.5, aoa_deg=2.0)
    q0 = problems.freestream_bump_state(m, ph, seed=2)
    print(f"case cube n=32 omega p=2 roe: mesh {time.perf_counter() - t0:.1f} s", flush=True)
    q_mix, rep_mix, its_mix, kry_mix = _solve(m, op, ph, q0, False, 60, verbose=True)
    q_64, rep_64, its_64, kry_64 = _solve(m, op, ph, q0, True, 60, verbose=True)
    for rep in (rep_mix, rep_64):
        assert rep.converged and rep.res_norm <= 1e-10
    d = np.abs(q_mix - q_64).max()
    print("max |q_mixed - q_fp64|", d, "iterations", its_mix, its_64, flush=True)
    assert d <= 1e-12
    assert abs(its_mix - its_64) <= 1
