"""Writes the golden residual fixtures: for a few small seeded problems, the
input state, the oracle's residual, the per-entry floating-point scale and
the oracle's residual norm and complex-step Jv, as compressed numpy archives.

    python tests/golden/make_golden.py

The fixtures pin the oracle against future changes (tests/test_golden.py) and
give the GPU tests fixed expected values independent of the oracle build."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "pdesolver.jl_amd"), os.path.join(ROOT, "oracle")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

import oracle_py  # noqa: E402
from pdes_amd import abi, problems  # noqa: E402

CASES = {
    # name: (kind, n, p, operator, flux)
    "vortex_4x4_omega_p1_roe": ("vortex", 4, 1, "omega", "roe"),
    "vortex_3x3_omega_p3_roe": ("vortex", 3, 3, "omega", "roe"),
    "vortex_3x3_diage_p2_ir": ("vortex", 3, 2, "diage", "ir"),
    "vortex_2x2_diage_p4_ir": ("vortex", 2, 4, "diage", "ir"),
    "cube_1_omega_p2_roe": ("cube", 1, 2, "omega", "roe"),
    "vortex_3x3_omega_p2_es": ("vortex", 3, 2, "omega", "es"),
}


def state(kind, m, ph):
    if kind == "vortex":
        return problems.vortex_fourier_state(m, ph, seed=7, amp=1e-2)
    q = problems.freestream_bump_state(m, ph, seed=7)
    rng = np.random.default_rng(107)
    q[..., 1:m.dim + 1] += 1e-3 * rng.uniform(0.5, 1.0, size=q[..., 1:m.dim + 1].shape)
    return q


def main():
    oracle_py.build()
    for name, (kind, n, p, opk, flux) in CASES.items():
        m, op, ph = problems.make_problem(kind, n, p, opk, flux, Ma=0.5, aoa_deg=2.0)
        b = abi.make_descriptors(m, op, ph)
        q = state(kind, m, ph)
        r = oracle_py.residual(b, q)
        sc = oracle_py.residual_scale(b, q)
        v = np.random.default_rng(8).standard_normal(q.shape)
        jv = oracle_py.jv_complex_step(b, q, v)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), q=q, res=r, scale=sc, v=v, jv=jv)
        print(name, q.shape)


if __name__ == "__main__":
    main()
