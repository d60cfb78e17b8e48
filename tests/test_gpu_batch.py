"""Batched steady solves (config 3 shape, small sizes) and their isolation
guarantees, through the C ABI."""
import numpy as np
import pytest

from pdes_amd import euler, linsolve, problems
from pdes_amd.batch import BatchSolver

pytestmark = pytest.mark.gpu

OPTS = linsolve.NewtonOptions(res_abstol=1e-10, krylov_reltol=1e-4, itermax=20)


def small_instances(count=7, seed=11):
    return problems.config3_instances(count=count, seed=seed, nmin=3, nmax=9)


def test_packed_residual_parity(oracle):
    """The packed residual (per-instance inflow Mach in the boundary states)
    equals the oracle residual of each instance."""
    from pdes_amd import abi
    from helpers import assert_parity
    inst = small_instances()
    bs = BatchSolver([(m, op, ph) for (m, op, ph, q0) in inst])
    for i, (m, op, ph, q0) in enumerate(inst):
        bs.set_state(i, q0)
    for i, (m, op, ph, q0) in enumerate(inst):
        b = abi.make_descriptors(m, op, ph)
        r = bs.eval_residual(i)
        print(i, op.kind, op.degree, m.numEl, "Min", ph.vortex_Min, flush=True)
        assert_parity(r, oracle.residual(b, q0), oracle.residual_scale(b, q0), 1e-12, f"instance {i}")
    bs.free()


def test_batch_converges_and_matches_single_context():
    inst = small_instances()
    bs = BatchSolver([(m, op, ph) for (m, op, ph, q0) in inst])
    print("groups", bs.num_groups(), flush=True)
    for i, (m, op, ph, q0) in enumerate(inst):
        bs.set_state(i, q0)
    reps = bs.solve(OPTS)
    for i, (m, op, ph, q0) in enumerate(inst):
        r = reps[i]
        print(i, op.kind, op.degree, m.numEl, "it", r.iterations, "norm", r.res_norm, flush=True)
        assert r.converged == 1 and r.res_norm < 1e-10
        # same problem through the single-context Newton path
        eqn = euler.EulerData(m, op, ph)
        eqn.set_q(q0)
        rep1 = linsolve.newton(m, op, eqn, OPTS)
        assert rep1.converged
        q1 = eqn.get_q()
        qb = bs.get_state(i)
        # the two paths reduce in different orders (packed chunked reductions
        # against one context's), so their Newton iterates round differently;
        # both states solve the same discrete problem to the tolerance, and
        # agree to what that tolerance fixes on these coarse meshes
        eqn.set_q(qb)
        euler.evalResidual_device(eqn)
        assert euler.calcNorm(eqn) < 1e-10
        diff = np.abs(qb - q1).max() / np.abs(q1).max()
        print(i, "relative state difference", diff, flush=True)
        assert diff <= 1e-7
        eqn.free()
    bs.free()


def test_batch_result_independent_of_mix():
    """Each instance gives the same bits alone and inside the mixed batch."""
    inst = small_instances(count=6, seed=5)
    bs = BatchSolver([(m, op, ph) for (m, op, ph, q0) in inst])
    for i, (m, op, ph, q0) in enumerate(inst):
        bs.set_state(i, q0)
    reps = bs.solve(OPTS)
    mixed = [bs.get_state(i) for i in range(len(inst))]
    bs.free()
    for i, (m, op, ph, q0) in enumerate(inst):
        alone = BatchSolver([(m, op, ph)])
        alone.set_state(0, q0)
        r = alone.solve(OPTS)[0]
        assert r.iterations == reps[i].iterations
        assert r.total_krylov_iters == reps[i].total_krylov_iters
        assert np.array_equal(alone.get_state(0), mixed[i]), f"instance {i} differs"
        alone.free()


def test_batch_failure_isolated():
    """checkDensity/checkPressure failing in one instance retires it (-1)
    without stopping the others."""
    inst = small_instances(count=4, seed=9)
    bs = BatchSolver([(m, op, ph) for (m, op, ph, q0) in inst])
    for i, (m, op, ph, q0) in enumerate(inst):
        q = q0.copy()
        if i == 2:
            q[0, 0, 0] = -1.0
        bs.set_state(i, q)
    reps = bs.solve(OPTS)
    assert reps[2].converged == -1
    for i in (0, 1, 3):
        assert reps[i].converged == 1 and reps[i].res_norm < 1e-10
    bs.free()


@pytest.mark.parametrize("pc_kind,globalize", [(2, False), (1, False), (2, True)])
def test_repack_mid_solve_is_bit_identical(pc_kind, globalize):
    """pdes_batch_advance / export / import: three same-kind solves advanced a
    few passes, exported at GMRES restarts, moved (in another order, beside a
    fresh fourth solve) into a new batch and finished there, give the bits and
    the Newton / Krylov counts of the same solves run uninterrupted."""
    opts = linsolve.NewtonOptions(res_abstol=1e-10, krylov_reltol=1e-2, itermax=30, pc_kind=pc_kind,
                                  globalize_euler=globalize, euler_tau=1.0)
    probs = []
    for i, (n, Min) in enumerate([(10, 0.5), (7, 0.7), (12, 0.4), (6, 0.6)]):
        m, op, ph = problems.make_problem("vortex", n, 2, "omega", "roe", vortex_Min=Min)
        probs.append((m, op, ph, problems.vortex_noise_state(m, ph, seed=20 + i)))
    ref = BatchSolver([(m, op, ph) for (m, op, ph, q) in probs])
    for i, p in enumerate(probs):
        ref.set_state(i, p[3])
    ref_reps = ref.solve(opts)
    ref_q = [ref.get_state(i) for i in range(len(probs))]
    ref.free()

    first = BatchSolver([(m, op, ph) for (m, op, ph, q) in probs[:3]])
    for i in range(3):
        first.set_state(i, probs[i][3])
    # one pass at a time until some solve is caught in the middle of a
    # linear solve (with easy systems a GMRES solve can end within one pass)
    for _ in range(12):
        reps, active, done = first.advance(opts, 1)
        assert not done
        states = {i: first.export(i) for i in range(3) if active[i]}
        if any(st[0].in_solve for st in states.values()):
            break
    finished = {i: (first.get_state(i), reps[i]) for i in range(3) if not active[i]}
    moved = sum(1 for i in states if states[i][0].in_solve)
    print("moved mid-solve:", moved, "finished before the move:", len(finished), flush=True)
    assert moved >= 1
    first.free()
    order = [i for i in (2, 0, 1) if i in states] + [3]
    second = BatchSolver([probs[i][:3] for i in order])
    for k, i in enumerate(order):
        if i == 3:
            second.set_state(k, probs[3][3])
        else:
            second.import_(k, *states[i])
    while True:
        reps2, active2, done2 = second.advance(opts, 4)
        if done2:
            break
    for k, i in enumerate(order):
        finished[i] = (second.get_state(k), reps2[k])
    second.free()
    for i in range(4):
        q, r = finished[i]
        assert r.converged == 1 and r.res_norm <= 1e-10
        assert r.iterations == ref_reps[i].iterations and r.total_krylov_iters == ref_reps[i].total_krylov_iters, i
        assert np.array_equal(q, ref_q[i]), f"instance {i} differs after the move"


def test_batch_ilu_sweep_kinds_bit_identical(monkeypatch):
    """A batched ILU(0) solve of uneven instances (diagonal-E p=2 and p=3,
    Omega p=2) gives the same bits, Newton and Krylov counts with the
    dependency-driven sweeps (PDES_ILU_DEP: chunks of same-level elements in
    ticket order, compacted to the instances still solving at each GMRES
    cycle) as with the level-by-level sweeps (PDES_ILU_LEVELS)."""
    opts = linsolve.NewtonOptions(res_abstol=1e-10, krylov_reltol=1e-2, itermax=30, pc_kind=2)
    probs = []
    for i, (n, p, opk, flux) in enumerate([(14, 2, "diage", "ir"), (6, 2, "diage", "ir"), (9, 3, "diage", "ir"),
                                           (8, 2, "omega", "roe"), (11, 2, "diage", "ir")]):
        m, op, ph = problems.make_problem("vortex", n, p, opk, flux, vortex_Min=0.4 + 0.05 * i)
        probs.append((m, op, ph, problems.vortex_noise_state(m, ph, seed=40 + i)))
    out = {}
    for kind in ("PDES_ILU_DEP", "PDES_ILU_LEVELS"):
        monkeypatch.setenv(kind, "1")
        monkeypatch.delenv("PDES_ILU_LEVELS" if kind == "PDES_ILU_DEP" else "PDES_ILU_DEP", raising=False)
        bs = BatchSolver([(m, op, ph) for (m, op, ph, q) in probs])
        for i, p in enumerate(probs):
            bs.set_state(i, p[3])
        reps = bs.solve(opts)
        out[kind] = ([bs.get_state(i) for i in range(len(probs))],
                     [(r.converged, r.iterations, r.total_krylov_iters) for r in reps])
        bs.free()
        print(kind, out[kind][1], flush=True)
    (qd, rd), (ql, rl) = out["PDES_ILU_DEP"], out["PDES_ILU_LEVELS"]
    assert rd == rl
    assert all(r[0] == 1 for r in rd)
    for i in range(len(probs)):
        assert np.array_equal(qd[i], ql[i]), f"instance {i} differs between the sweep kinds"


@pytest.mark.parametrize("pc_kind,globalize,forced", [(2, False, "PDES_ILU_DEP"), (2, False, "PDES_ILU_LEVELS"),
                                                      (1, False, None), (2, True, None)])
def test_batch_graph_replays_match_direct_launches(monkeypatch, pc_kind, globalize, forced):
    """The batched GMRES captures each block of 8 Arnoldi steps into a HIP
    graph once its launch configuration repeats, and replays it in later
    cycles.  Replayed and directly launched blocks (a batch created with
    PDES_NO_GRAPHS) give the same bits, Newton and Krylov counts, with
    block-Jacobi and both ILU(0) sweep kinds, with and without the
    pseudo-transient term; the graphed run did replay graphs."""
    opts = linsolve.NewtonOptions(res_abstol=1e-10, krylov_reltol=1e-5, itermax=30, pc_kind=pc_kind,
                                  globalize_euler=globalize, euler_tau=1.0)
    probs = []
    for i, (n, Min) in enumerate([(14, 0.5), (9, 0.7), (12, 0.4)]):
        m, op, ph = problems.make_problem("vortex", n, 2, "diage", "ir", vortex_Min=Min)
        probs.append((m, op, ph, problems.vortex_noise_state(m, ph, seed=30 + i)))
    if forced:
        monkeypatch.setenv(forced, "1")
    out = {}
    for graphs in (True, False):
        if graphs:
            monkeypatch.delenv("PDES_NO_GRAPHS", raising=False)
        else:
            monkeypatch.setenv("PDES_NO_GRAPHS", "1")
        bs = BatchSolver([(m, op, ph) for (m, op, ph, q) in probs])
        for i, p in enumerate(probs):
            bs.set_state(i, p[3])
        reps = bs.solve(opts)
        out[graphs] = ([bs.get_state(i) for i in range(len(probs))],
                       [(r.converged, r.iterations, r.total_krylov_iters) for r in reps], bs.graph_stats())
        bs.free()
        print("graphs" if graphs else "direct", out[graphs][1], "captures/replays", out[graphs][2], flush=True)
    (qg, rg, sg), (qd, rd, sd) = out[True], out[False]
    assert rg == rd
    assert all(r[0] == 1 for r in rg)
    assert sg[1] > 0 and sd == (0, 0)
    for i in range(len(probs)):
        assert np.array_equal(qg[i], qd[i]), f"instance {i} differs between graph replays and direct launches"
