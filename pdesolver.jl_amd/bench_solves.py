"""Batched steady solves: solves/s of many independent uneven 2D vortex
problems, packed into one batch, against the plain one-solve-at-a-time device
path on the same problems.

Workload: 96 instances (seed 3), grid edge count uniform in 8..128, Omega p in
{1,2,3} with the Roe weak form or diagonal-E p in {2,3} with the IR split form
and IRSLF faces, inflow Mach uniform in [0.3, 0.95], exact state times
(1 + U(-1e-3, 1e-3)).  Each solve is Newton to ||R||_{M^-1} <= 1e-10 with FP32
Jacobian blocks and an FP32 preconditioner (element-block ILU(0) by default,
or element block-Jacobi) under FP64 GMRES(30).

Both paths build their device objects before the timed region; the timed
region is state upload + Newton solve (+ nothing else) for every instance.

    python pdesolver.jl_amd/bench_solves.py [--count 96] [--nmax 128]
    python pdesolver.jl_amd/bench_solves.py --single 18 [--steps 1]   (one instance alone)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

from pdes_amd import abi, euler, linsolve, problems  # noqa: E402
from pdes_amd.batch import BatchSolver  # noqa: E402


PCS = {"bjacobi": abi.PDES_PC_BJACOBI, "ilu0": abi.PDES_PC_ILU0}


def newton_options(pc: str = "ilu0", restart: int = 30):
    return linsolve.NewtonOptions(res_abstol=1e-10, krylov_reltol=1e-2, itermax=50, pc_kind=PCS[pc],
                                  gmres_restart=restart)


def retirement(inst, reps):
    """When each instance of the batch converged (seconds from the start of
    the batched solve), and the instance that retired last."""
    lat = [r.elapsed_s for r in reps]
    last = int(np.argmax(lat))
    m, op, ph, q = inst[last]
    return {"batched_retire_p50_s": float(np.percentile(lat, 50)),
            "batched_retire_p99_s": float(np.percentile(lat, 99)),
            "batched_last": f"instance {last}: {op.kind} p={op.degree} {m.numEl} elements, "
                            f"{reps[last].iterations} Newton / {reps[last].total_krylov_iters} Krylov, "
                            f"retired at {lat[last]:.1f} s"}


def run(count: int = 96, seed: int = 3, nmin: int = 8, nmax: int = 128, check_tol: float = 1e-8,
        only: str = "both", pc: str = "ilu0", restart: int = 30):
    t0 = time.perf_counter()
    inst = problems.config3_instances(count=count, seed=seed, nmin=nmin, nmax=nmax)
    t_gen = time.perf_counter() - t0
    opts = newton_options(pc, restart)
    ndof = [int(q.size) for (m, op, ph, q) in inst]

    if only == "batched":
        bs = BatchSolver([(m, op, ph) for (m, op, ph, q) in inst])
        t0 = time.perf_counter()
        for i, (m, op, ph, q0) in enumerate(inst):
            bs.set_state(i, q0)
        reps = bs.solve(opts)
        t_batch = time.perf_counter() - t0
        bs.free()
        return {"instances": count, "batched_s": t_batch, "batched_solves_per_s": count / t_batch,
                "converged_batched": sum(1 for r in reps if r.converged == 1),
                "krylov_iters_batched": int(sum(r.total_krylov_iters for r in reps)),
                **retirement(inst, reps)}
    # one solve at a time: a context per instance, created up front
    ctxs = [euler.EulerData(m, op, ph) for (m, op, ph, q) in inst]
    seq_lat = []
    seq_reps = []
    t0 = time.perf_counter()
    for e, (m, op, ph, q0) in zip(ctxs, inst):
        t1 = time.perf_counter()
        e.set_q(q0)
        seq_reps.append(linsolve.newton(m, op, e, opts))
        seq_lat.append(time.perf_counter() - t1)
        print(f"sequential {len(seq_lat)}/{count}: {op.kind} p={op.degree} {m.numEl} elements, "
              f"{seq_reps[-1].iterations} Newton / {seq_reps[-1].total_krylov_iters} Krylov, {seq_lat[-1]:.2f} s",
              file=sys.stderr, flush=True)
    t_seq = time.perf_counter() - t0
    q_seq = [e.get_q() for e in ctxs]
    for e in ctxs:
        e.free()

    # batched
    print("batched solve", file=sys.stderr, flush=True)
    bs = BatchSolver([(m, op, ph) for (m, op, ph, q) in inst])
    t0 = time.perf_counter()
    for i, (m, op, ph, q0) in enumerate(inst):
        bs.set_state(i, q0)
    reps = bs.solve(opts)
    t_batch = time.perf_counter() - t0
    q_bat = [bs.get_state(i) for i in range(len(inst))]
    groups = bs.num_groups()
    bs.free()

    worst = 0.0
    for a, b in zip(q_seq, q_bat):
        worst = max(worst, float(np.abs(a - b).max() / np.abs(a).max()))
    conv_seq = sum(1 for r in seq_reps if r.converged == 1 and r.res_norm <= 1e-10)
    conv_bat = sum(1 for r in reps if r.converged == 1 and r.res_norm <= 1e-10)
    return {
        "workload": f"{count} 2D vortex instances (seed {seed}), n in {nmin}..{nmax}, Omega p1-3 Roe / "
                    "diagonal-E p2-3 IR split, Min in [0.3,0.95], Newton to ||R|| <= 1e-10",
        "preconditioner": pc,
        "gmres_restart": restart,
        "instances": count,
        "total_dofs": int(sum(ndof)),
        "groups": groups,
        "batched_solves_per_s": count / t_batch,
        "sequential_solves_per_s": count / t_seq,
        "speedup": t_seq / t_batch,
        "batched_s": t_batch,
        "sequential_s": t_seq,
        "sequential_latency_p50_s": float(np.percentile(seq_lat, 50)),
        "sequential_latency_p99_s": float(np.percentile(seq_lat, 99)),
        "converged_batched": conv_bat,
        "converged_sequential": conv_seq,
        "newton_iters_batched": int(sum(r.iterations for r in reps)),
        "newton_iters_sequential": int(sum(r.iterations for r in seq_reps)),
        "krylov_iters_batched": int(sum(r.total_krylov_iters for r in reps)),
        "krylov_iters_sequential": int(sum(r.total_krylov_iters for r in seq_reps)),
        **retirement(inst, reps),
        "max_rel_state_diff": worst,
        "states_agree": worst <= check_tol,
        "setup_s": t_gen,
    }


def run_single(index: int, count: int = 96, seed: int = 3, nmin: int = 8, nmax: int = 128, pc: str = "ilu0",
               restart: int = 30, steps: int = 50):
    """One instance of the workload solved alone on a single context, one
    Newton step per call so that each step prints (at most `steps` steps)."""
    inst = problems.config3_instances(count=count, seed=seed, nmin=nmin, nmax=nmax)
    m, op, ph, q0 = inst[index]
    eqn = euler.EulerData(m, op, ph)
    eqn.set_q(q0)
    opts = newton_options(pc, restart)
    opts.itermax = 1
    t0 = time.perf_counter()
    its = kry = 0
    rep = None
    for _ in range(steps):
        rep = linsolve.newton(m, op, eqn, opts)
        its += rep.iterations
        kry += rep.total_krylov_iters
        print(f"instance {index} step {its}: ||R|| {rep.res_norm:.3e}, {kry} Krylov, "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        if rep.converged:
            break
    t = time.perf_counter() - t0
    eqn.free()
    return {"instance": index, "operator": f"{op.kind} p={op.degree}", "numEl": int(m.numEl),
            "newton_iterations": its, "krylov_iterations": kry, "converged": int(rep.converged),
            "res_norm": rep.res_norm, "solve_s": t}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--single", type=int, default=-1, help="solve only this instance, alone, step by step")
    ap.add_argument("--steps", type=int, default=50, help="with --single: at most this many Newton steps")
    ap.add_argument("--count", type=int, default=96)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--nmin", type=int, default=8)
    ap.add_argument("--nmax", type=int, default=128)
    ap.add_argument("--only", choices=["both", "batched"], default="both")
    ap.add_argument("--pc", choices=sorted(PCS), default="ilu0")
    ap.add_argument("--restart", type=int, default=30)
    a = ap.parse_args()
    if a.single >= 0:
        print(json.dumps(run_single(a.single, a.count, a.seed, a.nmin, a.nmax, a.pc, a.restart, a.steps)), flush=True)
        return
    print(json.dumps(run(a.count, a.seed, a.nmin, a.nmax, only=a.only, pc=a.pc, restart=a.restart)), flush=True)


if __name__ == "__main__":
    main()
