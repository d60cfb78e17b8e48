"""Served steady solves: Poisson arrivals of mixed 2D/3D requests into a
`SolveServer`, p50/p99 latency and solves/s per arrival rate.

Requests (seed 4): 70% 2D isentropic-vortex quarter annuli with edge count in
nmin2..nmax2, 30% 3D freestream unit cubes (6 tets per cell) with nmin3..nmax3
cells per edge; Omega p in {1,2} with the Roe flux; freestream / inflow Mach
uniform in [0.3, 0.8], angle of attack uniform in [-5, 5] degrees, rho = 1.
2D requests start from the exact vortex times (1 + U(-1e-3, 1e-3)), 3D ones
from the freestream plus a smooth 1e-2 bump.  Every request is Newton to
||R|| <= 1e-10 with FP32 Jacobian blocks and ILU(0) under FP64 GMRES.
Meshes are built before the clock starts; a feeder thread submits each
request at its arrival time.

    python pdesolver.jl_amd/bench_serve.py [--requests 48] [--rates 20,400]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

from pdes_amd import abi, linsolve, problems  # noqa: E402
from pdes_amd.serve import SolveServer  # noqa: E402


def make_requests(count, seed=4, nmin2=32, nmax2=128, nmin3=8, nmax3=24):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        p = int(rng.integers(1, 3))
        Ma = float(rng.uniform(0.3, 0.8))
        aoa = float(rng.uniform(-5.0, 5.0))
        if rng.uniform() < 0.7:
            n = int(rng.integers(nmin2, nmax2 + 1))
            m, op, ph = problems.make_problem("vortex", n, p, "omega", "roe", Ma=Ma, aoa_deg=aoa, vortex_Min=Ma)
            q0 = problems.vortex_noise_state(m, ph, seed=seed * 1000 + i)
        else:
            n = int(rng.integers(nmin3, nmax3 + 1))
            m, op, ph = problems.make_problem("cube", n, p, "omega", "roe", Ma=Ma, aoa_deg=aoa)
            q0 = problems.freestream_bump_state(m, ph, seed=seed * 1000 + i)
        out.append((m, op, ph, q0))
    return out


def run_rate(reqs, rate, budget, lanes, seed=4):
    opts = linsolve.NewtonOptions(res_abstol=1e-10, krylov_reltol=1e-2, itermax=50, pc_kind=abi.PDES_PC_ILU0)
    srv = SolveServer(budget, lanes=lanes, opts=opts)
    rng = np.random.default_rng(seed + int(rate))
    gaps = rng.exponential(1.0 / rate, size=len(reqs))
    tickets = [None] * len(reqs)

    def feeder():
        t0 = time.perf_counter()
        at = 0.0
        for i, (m, op, ph, q0) in enumerate(reqs):
            at += gaps[i]
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets[i] = srv.submit(m, op, ph, q0)

    th = threading.Thread(target=feeder)
    t_start = time.perf_counter()
    th.start()
    th.join()
    done = 0
    for t in tickets:
        while not t.wait(30.0):
            print(f"rate {rate}: waiting, {done}/{len(reqs)} done", file=sys.stderr, flush=True)
        done += 1
    t_end = max(t.t_done for t in tickets)
    srv.shutdown()
    lat = np.array([t.latency for t in tickets])
    ok = [t for t in tickets if t.error is None and t.report is not None and t.report.converged == 1
          and t.report.res_norm <= 1e-10]
    return {
        "rate_per_s": rate,
        "requests": len(reqs),
        "converged": len(ok),
        "errors": sum(1 for t in tickets if t.error is not None),
        "solves_per_s": len(ok) / (t_end - t_start),
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "mean_batch_size": float(np.mean([t.batch_size for t in tickets])),
        "batches": srv.batches,
        "repacks": srv.repacks,
        "peak_admitted_bytes": srv.peak_bytes,
    }


def run(requests=48, rates=(20.0, 400.0), budget_gb=96.0, lanes=4, seed=4, nmax2=128, nmax3=24):
    t0 = time.perf_counter()
    reqs = make_requests(requests, seed=seed, nmax2=nmax2, nmax3=nmax3)
    gen = time.perf_counter() - t0
    out = {"workload": f"{requests} requests (seed {seed}): 70% 2D vortex annulus n in 32..{nmax2}, 30% 3D cube n in "
                       f"8..{nmax3}, Omega p1-2 Roe, M in [0.3,0.8], aoa in [-5,5] deg, Newton to ||R|| <= 1e-10, "
                       "ILU(0)",
           "hbm_budget_bytes": int(budget_gb * 2**30), "lanes": lanes, "setup_s": gen, "rates": []}
    for r in rates:
        print(f"rate {r}/s", file=sys.stderr, flush=True)
        out["rates"].append(run_rate(reqs, r, int(budget_gb * 2**30), lanes, seed))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--rates", default="20,400")
    ap.add_argument("--budget-gb", type=float, default=96.0)
    ap.add_argument("--lanes", type=int, default=4)
    ap.add_argument("--nmax2", type=int, default=128)
    ap.add_argument("--nmax3", type=int, default=24)
    a = ap.parse_args()
    rates = tuple(float(x) for x in a.rates.split(","))
    print(json.dumps(run(a.requests, rates, a.budget_gb, a.lanes, nmax2=a.nmax2, nmax3=a.nmax3)), flush=True)


if __name__ == "__main__":
    main()
