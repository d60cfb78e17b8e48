"""3D Newton-Krylov, mixed precision: the unit cube meshed as n^3 hexes split
into 6 tetrahedra each (n = 32: 196,608 elements), SBP-Omega p=2 with the Roe
flux and freestream BCs (rho = 1, M = 0.5, 2 degrees angle of attack), initial
state freestream plus a seeded smooth 1e-2 bump.  Residual, Newton and
GMRES(30) in FP64; Jacobian blocks and the preconditioner in FP32; the same
solve with FP64 blocks for comparison, to ||R|| <= 1e-10.  The default
restart length is 60: with element-block ILU(0), GMRES(30) stagnates on the
second Newton system at this size (relative residual 0.86 after 600
iterations) while GMRES(60) reaches 1e-2 in about 110.

    python pdesolver.jl_amd/bench_newton3d.py [--n 32] [--pc ilu0] [--restart 60]
                                              [--precision both|mixed|fp64] [--steps 50]
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

PCS = {"bjacobi": abi.PDES_PC_BJACOBI, "ilu0": abi.PDES_PC_ILU0}


def run(n=32, p=2, pc="ilu0", krylov_reltol=1e-2, restart=60, precisions=(False, True), max_steps=50):
    t0 = time.perf_counter()
    m, op, ph = problems.make_problem("cube", n, p, "omega", "roe", Ma=0.5, aoa_deg=2.0)
    q0 = problems.freestream_bump_state(m, ph, seed=2)
    setup = time.perf_counter() - t0
    out = {"workload": f"unit cube {n}^3 x 6 tets ({m.numEl} elements), Omega p={p}, Roe, freestream M=0.5 aoa=2 deg, "
                       f"Newton to ||R|| <= 1e-10, GMRES({restart}) rtol {krylov_reltol}, {pc}",
           "numDof": int(q0.size), "setup_s": setup}
    states = {}
    for fp64 in precisions:
        eqn = euler.EulerData(m, op, ph)
        eqn.set_q(q0)
        opts = linsolve.NewtonOptions(res_abstol=1e-10, krylov_reltol=krylov_reltol, pc_kind=PCS[pc],
                                      fp64_blocks=fp64, gmres_restart=restart, krylov_itermax=3000)
        # one Newton step per call so that every step reports its progress;
        # the device state carries over, so this is newtonInner step by step
        opts.itermax = 1
        t1 = time.perf_counter()
        its = kry = 0
        norm0 = None
        rep = None
        for _ in range(max_steps):
            rep = linsolve.newton(m, op, eqn, opts)
            norm0 = rep.res_norm0 if norm0 is None else norm0
            its += rep.iterations
            kry += rep.total_krylov_iters
            print(f"fp64_blocks={fp64} step {its}: ||R|| {rep.res_norm:.3e}, {rep.total_krylov_iters} Krylov, "
                  f"{time.perf_counter() - t1:.1f} s", file=sys.stderr, flush=True)
            if rep.converged:
                break
        dt = time.perf_counter() - t1
        states[fp64] = eqn.get_q()
        out["fp64_blocks" if fp64 else "mixed"] = {
            "converged": int(rep.converged), "newton_iterations": its, "krylov_iterations": kry,
            "res_norm0": norm0, "res_norm": rep.res_norm, "solve_s": dt}
        eqn.free()
    if len(states) == 2:
        d = states[False] - states[True]
        w = (op.w[None, :] / m.jac)[:, :, None]          # H-norm weights (mass matrix)
        out["max_abs_state_diff"] = float(np.abs(d).max())
        out["H_norm_state_diff"] = float(np.sqrt((d * d * w).sum()))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--p", type=int, default=2)
    ap.add_argument("--pc", choices=sorted(PCS), default="ilu0")
    ap.add_argument("--restart", type=int, default=60)
    ap.add_argument("--precision", choices=["both", "mixed", "fp64"], default="both",
                    help="FP32 blocks (mixed), FP64 blocks, or both and their state difference")
    ap.add_argument("--steps", type=int, default=50, help="at most this many Newton steps per solve")
    a = ap.parse_args()
    prec = {"both": (False, True), "mixed": (False,), "fp64": (True,)}[a.precision]
    print(json.dumps(run(a.n, a.p, a.pc, restart=a.restart, precisions=prec, max_steps=a.steps)), flush=True)


if __name__ == "__main__":
    main()
