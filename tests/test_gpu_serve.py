"""Served solves: admission by HBM footprint, concurrent lanes, and results
that do not depend on what a request was batched or overlapped with."""
import numpy as np
import pytest

from pdes_amd import abi, linsolve, problems
from pdes_amd.batch import BatchSolver
from pdes_amd.serve import SolveServer, device_memory, estimate_bytes

pytestmark = pytest.mark.gpu

OPTS = linsolve.NewtonOptions(res_abstol=1e-10, krylov_reltol=1e-2, itermax=30, pc_kind=abi.PDES_PC_ILU0)


def _requests():
    rng = np.random.default_rng(21)
    out = []
    for i in range(8):
        p = 1 + i % 2
        Ma = float(rng.uniform(0.3, 0.8))
        if i % 3 != 2:
            m, op, ph = problems.make_problem("vortex", int(rng.integers(4, 9)), p, "omega", "roe", vortex_Min=Ma)
            q0 = problems.vortex_noise_state(m, ph, seed=i)
        else:
            m, op, ph = problems.make_problem("cube", 2, p, "omega", "roe", Ma=Ma, aoa_deg=float(rng.uniform(-5, 5)))
            q0 = problems.freestream_bump_state(m, ph, seed=i)
        out.append((m, op, ph, q0))
    return out


def test_served_results_equal_solo_solves():
    """Requests queue behind a budget of about two requests, are admitted as
    memory frees up, overlap on two lanes, and each result equals (bit for
    bit) the same request solved alone."""
    reqs = _requests()
    sizes = [estimate_bytes(m, op, ph, OPTS) for (m, op, ph, q0) in reqs]
    budget = 2 * max(sizes)
    srv = SolveServer(budget, lanes=2, opts=OPTS)
    tickets = [srv.submit(m, op, ph, q0) for (m, op, ph, q0) in reqs]
    for t in tickets:
        assert t.wait(120.0), "request did not finish"
    srv.shutdown()
    assert srv.peak_bytes <= budget
    for i, (t, (m, op, ph, q0)) in enumerate(zip(tickets, reqs)):
        assert t.error is None, t.error
        print(i, m.dim, op.degree, m.numEl, "batch", t.batch_size, "it", t.report.iterations, "norm", t.report.res_norm,
              flush=True)
        assert t.report.converged == 1 and t.report.res_norm <= 1e-10
        solo = BatchSolver([(m, op, ph)])
        solo.set_state(0, q0)
        r = solo.solve(OPTS)[0]
        assert r.iterations == t.report.iterations
        assert np.array_equal(solo.get_state(0), t.q), f"request {i} differs from its solo solve"
        solo.free()


def test_request_above_budget_is_rejected():
    m, op, ph = problems.make_problem("vortex", 4, 1, "omega", "roe")
    srv = SolveServer(1024, lanes=1, opts=OPTS)
    with pytest.raises(ValueError):
        srv.submit(m, op, ph, problems.vortex_noise_state(m, ph))
    srv.shutdown()


def test_footprint_estimate_covers_the_allocations():
    """pdes_estimate_bytes against the device memory a batch really takes
    (free memory before creation minus free memory after a solve)."""
    m, op, ph = problems.make_problem("cube", 6, 2, "omega", "roe", Ma=0.5, aoa_deg=2.0)
    est = estimate_bytes(m, op, ph, OPTS)
    # a small solve of the same kind first: the runtime allocates the
    # Jacobian kernels' scratch (private segment) at their first launch, once
    # per process, and that is not part of a solve's footprint
    mw, opw, phw = problems.make_problem("cube", 1, 2, "omega", "roe", Ma=0.5, aoa_deg=2.0)
    warm = BatchSolver([(mw, opw, phw)])
    warm.set_state(0, problems.freestream_bump_state(mw, phw))
    warm.solve(OPTS)
    warm.free()
    free0, _ = device_memory()
    bs = BatchSolver([(m, op, ph)])
    bs.set_state(0, problems.freestream_bump_state(m, ph))
    bs.solve(OPTS)
    free1, _ = device_memory()
    bs.free()
    used = free0 - free1
    print("estimate", est, "measured", used, "ratio", used / est)
    assert 0.6 * est <= used <= 1.15 * est


def _minv(m, op):
    return np.repeat((1.0 / (op.w[None, :] / m.jac))[:, :, None], m.dim + 2, axis=2)


def test_served_results_match_oracle_direct_solves(oracle):
    """Every served result (a mix of 2D vortex and 3D freestream requests at
    sizes the oracle solves in seconds, p in {1, 2}, Mach and angle of attack
    drawn as in the serving workload) against the oracle's Newton with a
    sparse direct solve: both reach ||R|| <= 1e-10 and the states agree to
    1e-9 of their magnitude."""
    reqs = _requests()
    srv = SolveServer(4 * max(estimate_bytes(m, op, ph, OPTS) for (m, op, ph, q0) in reqs), lanes=2, opts=OPTS)
    tickets = [srv.submit(m, op, ph, q0) for (m, op, ph, q0) in reqs]
    for t in tickets:
        assert t.wait(120.0), "request did not finish"
    srv.shutdown()
    for i, (t, (m, op, ph, q0)) in enumerate(zip(tickets, reqs)):
        print(f"request {i}: dim {m.dim} p={op.degree} numEl {m.numEl}", flush=True)
        assert t.error is None, t.error
        assert t.report.converged == 1 and t.report.res_norm <= 1e-10
        b = abi.make_descriptors(m, op, ph)
        q_ref, norms = oracle.newton_direct(b, q0, _minv(m, op), res_abstol=1e-10)
        assert norms[-1] < 1e-10
        assert np.abs(t.q - q_ref).max() <= 1e-9 * np.abs(q_ref).max()


def _full_size_requests(count, seed):
    """Requests at the serving workload's sizes: 70% 2D vortex grids with
    edges in 32..128, 30% 3D freestream cubes with 8..24 cells per edge,
    p in {1, 2}, M in [0.3, 0.8], angle of attack in [-5, 5] degrees."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        p = int(rng.integers(1, 3))
        Ma = float(rng.uniform(0.3, 0.8))
        aoa = float(rng.uniform(-5.0, 5.0))
        if rng.random() < 0.7:
            n = int(rng.integers(32, 129))
            m, op, ph = problems.make_problem("vortex", n, p, "omega", "roe", Ma=Ma, aoa_deg=aoa, vortex_Min=Ma)
            q0 = problems.vortex_noise_state(m, ph, seed=seed * 100 + i)
        else:
            n = int(rng.integers(8, 25))
            m, op, ph = problems.make_problem("cube", n, p, "omega", "roe", Ma=Ma, aoa_deg=aoa)
            q0 = problems.freestream_bump_state(m, ph, seed=seed * 100 + i)
        out.append((m, op, ph, q0))
    return out


def test_served_full_size_requests_converge_and_equal_solo():
    """At the serving workload's full request sizes: every served request
    converges to ||R|| <= 1e-10 and equals, bit for bit, the same request
    solved alone."""
    reqs = _full_size_requests(6, 41)
    budget = 3 * max(estimate_bytes(m, op, ph, OPTS) for (m, op, ph, q0) in reqs)
    srv = SolveServer(budget, lanes=2, opts=OPTS)
    tickets = [srv.submit(m, op, ph, q0) for (m, op, ph, q0) in reqs]
    for i, t in enumerate(tickets):
        assert t.wait(300.0), "request did not finish"
        print(f"served {i}: numEl {reqs[i][0].numEl} p={reqs[i][1].degree} latency {t.latency:.2f} s", flush=True)
    srv.shutdown()
    for i, (t, (m, op, ph, q0)) in enumerate(zip(tickets, reqs)):
        assert t.error is None, t.error
        assert t.report.converged == 1 and t.report.res_norm <= 1e-10
        solo = BatchSolver([(m, op, ph)])
        solo.set_state(0, q0)
        r = solo.solve(OPTS)[0]
        print(f"solo {i}: iterations {r.iterations} (served {t.report.iterations})", flush=True)
        assert r.iterations == t.report.iterations
        assert np.array_equal(solo.get_state(0), t.q), f"request {i} differs from its solo solve"
        solo.free()


def test_request_joins_a_running_batch():
    """Continuous batching: a request of the same kind submitted while another
    is being solved joins that running batch (a repack at a GMRES restart)
    and is admitted before the first one finishes; both results equal their
    solo solves bit for bit, and each is handed back as soon as it converges."""
    import time
    mA, opA, phA = problems.make_problem("vortex", 64, 2, "omega", "roe", vortex_Min=0.5)
    qA = problems.vortex_noise_state(mA, phA, seed=31)
    mB, opB, phB = problems.make_problem("vortex", 8, 2, "omega", "roe", vortex_Min=0.7)
    qB = problems.vortex_noise_state(mB, phB, seed=32)
    srv = SolveServer(1 << 36, lanes=2, opts=OPTS, passes=1)
    tA = srv.submit(mA, opA, phA, qA)
    t_wait = time.perf_counter()
    while tA.t_admit == 0.0 and time.perf_counter() - t_wait < 60:
        time.sleep(0.01)
    time.sleep(0.2)
    assert not tA.done(), "the first request finished before the second could join"
    tB = srv.submit(mB, opB, phB, qB)
    assert tA.wait(120.0) and tB.wait(120.0)
    srv.shutdown()
    print("A admitted", tA.t_admit - tA.t_submit, "done", tA.t_done - tA.t_submit, "B admitted",
          tB.t_admit - tA.t_submit, "done", tB.t_done - tA.t_submit, "repacks", srv.repacks, flush=True)
    assert tB.t_admit < tA.t_done
    assert srv.repacks >= 1
    assert tB.t_done < tA.t_done          # the small request leaves as soon as it converges
    for t, (m, op, ph, q0) in ((tA, (mA, opA, phA, qA)), (tB, (mB, opB, phB, qB))):
        assert t.error is None and t.report.converged == 1
        solo = BatchSolver([(m, op, ph)])
        solo.set_state(0, q0)
        r = solo.solve(OPTS)[0]
        assert r.iterations == t.report.iterations
        assert np.array_equal(solo.get_state(0), t.q)
        solo.free()
