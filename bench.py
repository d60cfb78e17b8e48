"""Headline benchmark: DOF-residual evaluations per second of the Euler DG-SBP
residual on the 256x256 quarter-annulus isentropic-vortex problem (131,072
triangles), SBP diagonal-E at p=2 and p=4, with (a) Roe face flux and weak-form
Q^T F volume integral and (b) Ismail-Roe entropy-stable split form with the
IRSLF face flux.  A step is one classical RK4 step (4 residual evaluations plus
the Minv-scaled stage updates) of each of the four variants; the default run
times 250 steps, i.e. 1,000 residual evaluations per variant.

    python bench.py [--gpus 1] [--steps 250] [--warmup 5]

Prints one JSON line.  The state is resident in HBM before the timed region.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
# The batched solves run their 5 operator groups concurrently, each on its own
# stream, and the served requests a stream per lane group; HIP maps a process's
# streams onto GPU_MAX_HW_QUEUES hardware queues and streams sharing a queue
# run in turn.  Give the process at least 16 queues (set before the first GPU
# call, which HIP reads it at); the residual timing uses one stream either way.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"
for _p in (os.path.join(ROOT, "pdesolver.jl_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
WARMUP_MIN_MS = 400.0


def residual_bytes_per_launch(m, op):
    """Algorithmic HBM bytes of one residual launch, each array touched once:
    q read + res write (ndof doubles per node), the metric dxidx (per node, or
    once per element when every node of the element has the same metric, as on
    these straight-sided triangles), the element->face records (16 B per face
    slot), the face normals once per interface and the boundary normals and
    coordinates once per boundary face.  SBP tables are excluded (cache
    resident)."""
    Np, Nf, dim, nd = op.numnodes, op.face.numnodes, m.dim, m.dim + 2
    affine = bool(np.all(m.dxidx == m.dxidx[:, :1]))
    metric = 8 * dim * dim * (1 if affine else Np)
    per_el = 8 * (nd * Np + nd * Np) + metric + 16 * (dim + 1)
    faces = 8 * dim * Nf * m.numInterfaces + 8 * 2 * dim * Nf * m.numBoundaryFaces
    return per_el * m.numEl + faces


# Algorithmic FP64 flops of the 2D Ismail-Roe split-form residual, counted from
# the kernel's arithmetic with fma = 2 and div = sqrt = log = 1:
#   per node pair (each unordered pair once):
#     direction n = sum_d 2 S_d[j][k] dxidx_d                       6
#     IR flux F*(q_j, q_k) . n                                      59
#       (two logavg on the series branch 2 x 12, sums, reciprocals and
#        rho_hat 8, velocity average 4, p1_hat 1, |v|^2 3, h_hat 8,
#        one direction 11)
#     res_j -= F*, res_k += F*                                      8
#   per node: IR z-variables                                        18
#   per face node (each interface once): IRSLF flux with the neighbour's
#     z-variables 18, the one-direction IR flux 59 and its two log ratios 8,
#     two entropy-variable conversions without the entropy 36, the entropy
#     jump from the log ratios 5, jump and average 12, lambda_max A0 dv 61,
#     and the two-sided face integration 8                          207
IR_FLOPS_PER_PAIR = 6 + 59 + 8
IR_FLOPS_PER_NODE = 18
IRSLF_FLOPS_PER_FACE_NODE = 199 + 8


def residual_flops_per_launch(m, op):
    """Algorithmic flops of one split-form (IR + IRSLF) residual launch."""
    Np, Nf = op.numnodes, op.face.numnodes
    pairs = Np * (Np - 1) // 2
    face_nodes = Nf * (m.numInterfaces + m.numBoundaryFaces)
    return (pairs * IR_FLOPS_PER_PAIR + Np * IR_FLOPS_PER_NODE) * m.numEl + face_nodes * IRSLF_FLOPS_PER_FACE_NODE


def traffic_for(p, flux):
    """Measured HBM bytes per launch of the residual kernel for this variant,
    read from the committed counter summary (FETCH_SIZE + WRITE_SIZE per
    dispatch, scaled by the same counters on a copy kernel of known size);
    None when no summary is present."""
    path = os.path.join(ROOT, "profiles", "euler_residual_256x256_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    key = f"| p{p}_{flux} |"
    with open(path) as f:
        for line in f:
            if line.startswith(key):
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                return float(cells[-1])
    return None


def variants():
    return [(2, "roe"), (4, "roe"), (2, "ir"), (4, "ir")]


def make(p, flux):
    from pdes_amd import problems
    m, op, ph = problems.make_problem("vortex", 256, p, "diage", flux)
    q = problems.vortex_fourier_state(m, ph, seed=1, amp=1e-2)
    return m, op, ph, q


def dt_for(m, q, ph, p):
    # dt = 0.05 h / ((|u| + a) (2p + 1)) with h the smallest element size
    h = np.sqrt(2.0 * (1.0 / m.jac[:, 0]).min() * 2.0)
    rho = q[..., 0]
    u = np.sqrt(q[..., 1] ** 2 + q[..., 2] ** 2) / rho
    pr = (ph.gamma - 1) * (q[..., 3] - 0.5 * rho * u * u)
    a = np.sqrt(ph.gamma * pr / rho)
    return 0.05 * h / ((u + a).max() * (2 * p + 1))


def host_cpu():
    """Host CPU model (lscpu's "Model name", from /proc/cpuinfo), the cores
    this process may use, and the thread budget of this host share
    (OMP_NUM_THREADS when set)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = len(os.sched_getaffinity(0))
    threads = min(nproc, int(os.environ.get("OMP_NUM_THREADS", nproc) or nproc))
    return model, nproc, max(1, threads)


def cpu_baseline(budget_s=16.0):
    """The CPU oracle (a C++ restatement of the reference residual, -O3)
    timed on this host on a bounded sample: residual evaluations of each
    variant on the full mesh until the per-variant time budget is spent, once
    on one thread (the reference runs one rank per core) and once on all the
    threads of this host share (OpenMP over elements and element-disjoint face
    groups).  The all-threads figure is the reported value."""
    import oracle_py
    from pdes_amd import abi
    oracle_py.build()
    model, nproc, threads = host_cpu()
    L = oracle_py.lib()
    runs = {}
    for nthr, budget in ((1, 0.6 * budget_s), (threads, 0.4 * budget_s)):
        L.oracle_set_threads(nthr)
        total_dof = 0
        total_t = 0.0
        sample = []
        for p, flux in variants():
            m, op, ph, q = make(p, flux)
            b = abi.make_descriptors(m, op, ph)
            n = 0
            t0 = time.perf_counter()
            while True:
                oracle_py.residual(b, q)
                n += 1
                if time.perf_counter() - t0 > budget / 4:
                    break
            dt = time.perf_counter() - t0
            total_dof += q.size * n
            total_t += dt
            sample.append(f"p{p}-{flux}:{n}")
        runs[nthr] = (total_dof / total_t, ",".join(sample))
    L.oracle_set_threads(1)
    return {"value": runs[threads][0], "unit": "DOF-residual evals/s", "cores": threads, "kind": "port",
            "sample": "full 256x256 mesh, residual evaluations per variant " + runs[threads][1],
            "single_thread": {"value": runs[1][0], "cores": 1, "sample": runs[1][1]},
            "host": {"cpu_model": model, "nproc": nproc, "threads_used": threads}}


def solves_metrics(seq_sample=12, serve_requests=32, rates=(20.0, 400.0)):
    """The other two metrics, steady solves/s at ||R|| <= 1e-10 and p50/p99
    solve latency, with inputs built before each clock starts:

    * batched: the 96 uneven 2D vortex solves (seed 3; n in 8..128, Omega p1-3
      Roe or diagonal-E p2-3 IR split form, inflow Mach 0.3-0.95, FP32 blocks,
      ILU(0), GMRES(30)) in one batch; solves/s and the p50/p99 of the times
      at which instances retire;
    * one at a time: the plain single-context device path (pdes_newton_solve)
      on a seeded sample of seq_sample of those instances, timed per solve;
      its solves/s is the sample's count over its total time;
    * served: Poisson arrivals of serve_requests mixed 2D/3D requests (seed 4;
      70% 2D vortex n 32..128, 30% 3D cube n 8..24, Omega p1-2, M 0.3-0.8)
      into the continuously batching SolveServer at each rate; solves/s and
      p50/p99 latency from submit to result."""
    from pdes_amd import euler, linsolve, problems
    from pdes_amd.batch import BatchSolver
    import bench_serve
    import bench_solves
    opts = bench_solves.newton_options("ilu0", 30)
    t0 = time.perf_counter()
    inst = problems.config3_instances()
    gen = time.perf_counter() - t0
    print(f"solves: batched 96 instances (built in {gen:.1f} s)", file=sys.stderr, flush=True)
    bs = BatchSolver([(m, op, ph) for (m, op, ph, q) in inst])
    t0 = time.perf_counter()
    for i, (m, op, ph, q0) in enumerate(inst):
        bs.set_state(i, q0)
    reps = bs.solve(opts)
    t_batch = time.perf_counter() - t0
    bs.free()
    lat_b = [r.elapsed_s for r in reps]
    conv_b = sum(1 for r in reps if r.converged == 1 and r.res_norm <= 1e-10)
    rng = np.random.default_rng(7)
    sample = sorted(int(i) for i in rng.choice(len(inst), size=seq_sample, replace=False))
    seq_lat, conv_s = [], 0
    for i in sample:
        m, op, ph, q0 = inst[i]
        e = euler.EulerData(m, op, ph)
        t1 = time.perf_counter()
        e.set_q(q0)
        r = linsolve.newton(m, op, e, opts)
        seq_lat.append(time.perf_counter() - t1)
        conv_s += int(r.converged == 1 and r.res_norm <= 1e-10)
        e.free()
        print(f"solves: one at a time, instance {i}: {seq_lat[-1]:.2f} s", file=sys.stderr, flush=True)
    seq_rate = len(sample) / sum(seq_lat)
    out = {"batched": {"workload": "96 2D vortex instances (seed 3), n 8..128, Omega p1-3 Roe / diagonal-E p2-3 "
                                   "IR split, inflow Mach 0.3-0.95, Newton to ||R|| <= 1e-10, FP32 blocks, ILU(0), "
                                   "GMRES(30)",
                       "solves_per_s": len(inst) / t_batch, "seconds": t_batch, "converged": conv_b,
                       "retire_p50_s": float(np.percentile(lat_b, 50)),
                       "retire_p99_s": float(np.percentile(lat_b, 99))},
           "one_at_a_time": {"sample": sample, "solves_per_s": seq_rate, "seconds": sum(seq_lat),
                             "converged": conv_s, "latency_p50_s": float(np.percentile(seq_lat, 50)),
                             "latency_p99_s": float(np.percentile(seq_lat, 99))},
           "speedup_batched_vs_one_at_a_time": (len(inst) / t_batch) / seq_rate,
           "gpu_max_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES")}
    t0 = time.perf_counter()
    reqs = bench_serve.make_requests(serve_requests)
    print(f"solves: served requests built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    served = []
    for rate in rates:
        r = bench_serve.run_rate(reqs, rate, 96 << 30, 4)
        served.append({k: r[k] for k in ("rate_per_s", "requests", "converged", "solves_per_s", "latency_p50_s",
                                         "latency_p99_s", "mean_batch_size", "batches", "repacks")})
        print(f"solves: served at {rate}/s: {served[-1]}", file=sys.stderr, flush=True)
    out["served"] = {"workload": f"{serve_requests} Poisson-arrival requests (seed 4): 70% 2D vortex n 32..128, "
                                 "30% 3D cube n 8..24, Omega p1-2 Roe, M 0.3-0.8, aoa +-5 deg, ILU(0), "
                                 "HBM budget 96 GiB", "rates": served}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=250)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solves", action="store_true", help="skip the steady-solve metrics")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="few copy-kernel launches only (for counter runs)")
    ap.add_argument("--variants", default="all", help="comma list like p2_roe,p4_ir (default all)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X; "
              "--gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)
    from pdes_amd import euler

    results = {}
    total_dof_evals = 0
    total_ms = 0.0
    kern = {}
    sel = None if args.variants == "all" else set(args.variants.split(","))
    for p, flux in variants():
        if sel is not None and f"p{p}_{flux}" not in sel:
            continue
        m, op, ph, q = make(p, flux)
        eqn = euler.EulerData(m, op, ph)
        eqn.set_q(q)
        h = dt_for(m, q, ph, p)
        tuned = euler.autotune(eqn, 10)
        eqn.set_q(q)
        # warm-up: the requested steps, then more until at least
        # WARMUP_MIN_MS of device time has run, so that a short timed region
        # starts at steady clocks.  The timed steps continue from the warmed
        # state (the perturbed vortex evolves smoothly at this small dt, and a
        # step's work does not depend on the state), so no host upload leaves
        # the GPU idle between warm-up and timing.
        warm_ms = euler.rk4(eqn, max(args.warmup, 1), h, timed=True)
        while warm_ms < WARMUP_MIN_MS:
            warm_ms += euler.rk4(eqn, 25, h, timed=True)
        ms = euler.rk4(eqn, args.steps, h, timed=True)
        dof_evals = 4 * args.steps * q.size
        total_dof_evals += dof_evals
        total_ms += ms
        # residual kernel alone, HIP events on the context stream
        kms = euler.time_residual(eqn, 50)
        # the same evaluation through the host-buffer boundary: q uploaded from
        # host memory, residual, res copied back (PCIe-inclusive; not `value`)
        t0 = time.perf_counter()
        for _ in range(5):
            eqn.set_q(q)
            euler.evalResidual(m, op, eqn)
        host_ms = (time.perf_counter() - t0) / 5 * 1e3
        nbytes = residual_bytes_per_launch(m, op)
        nflops = residual_flops_per_launch(m, op) if flux == "ir" else None
        kern[(p, flux)] = (kms, nbytes, nflops)
        r = {"dof_evals_per_s": dof_evals / (ms / 1e3), "ms_per_rk4_step": ms / args.steps,
             "residual_kernel_ms": kms, "residual_alg_bytes": nbytes,
             "residual_GBs": nbytes / (kms / 1e3) / 1e9, "numDof": int(q.size), "launch_variants": tuned,
             "host_boundary_ms": host_ms, "host_boundary_dof_evals_per_s": q.size / (host_ms / 1e3)}
        if nflops is not None:
            r["residual_alg_flops"] = nflops
            r["residual_TFLOPs"] = nflops / (kms / 1e3) / 1e12
        results[f"p{p}_{flux}"] = r
        eqn.free()
    # dominant kernel by time in the step: the split form is FP64-bound (priced
    # in flops against the FP64 peak), the weak form HBM-bound (priced in bytes)
    (pd, fd), (kms, nbytes, nflops) = max(kern.items(), key=lambda kv: kv[1][0])
    if nflops is not None:
        achieved = nflops / (kms / 1e3) / 1e12
        roof = {"kernel": f"euler_residual p={pd} {fd} split form", "bound": "valu_fp64", "achieved": achieved,
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                "pipe": "FP64 vector ALU (its 78.6 TFLOP/s peak equals the FP64 MFMA peak; no MFMA is issued)",
                "traffic": traffic_for(pd, fd), "alg_bytes": nbytes}
    else:
        achieved = nbytes / (kms / 1e3) / 1e9
        roof = {"kernel": f"euler_residual p={pd} {fd} weak form", "bound": "hbm", "achieved": achieved,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic_for(pd, fd)}
    line = {
        "metric": "DOF-residual evals/s",
        "other_metrics": "steady solves/s at ||R|| <= 1e-10 and p50/p99 solve latency: steady_solves",
        "value": total_dof_evals / (total_ms / 1e3),
        "unit": "DOF-residual evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2D isentropic vortex, 256x256 quarter annulus (131072 triangles), "
                               "SBP diagonal-E p=2 and p=4, Roe weak form and Ismail-Roe split form, "
                               "residuals inside RK4 stages",
                   "variants": results},
        "roofline": roof,
    }
    # plain copy kernel, 1 GiB read + 1 GiB write: attainable bandwidth, and the
    # reference point that scales the HBM byte counters of the residual kernel
    copy_bytes = 1 << 30
    cms = euler.copy_bandwidth(copy_bytes, 3 if args.calibrate_copy else 20)
    line["copy_reference"] = {"bytes_read": copy_bytes, "bytes_written": copy_bytes, "kernel_ms": cms,
                              "GBs": 2 * copy_bytes / (cms / 1e3) / 1e9}
    if not args.no_solves and args.variants == "all" and not args.calibrate_copy:
        line["steady_solves"] = solves_metrics()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
