"""Initial conditions and the synthetic benchmark problems.

`vortex_exact` restates calcIsentropicVortex
(src/solver/euler/common_funcs.jl:25-76, used by ICIsentropicVortex ic.jl:326
and isentropicVortexBC bc.jl:554-567) vectorised over nodes, with the inflow
Mach number as a parameter (the reference hard-codes 0.95).  `freestream`
restates calcFreeStream (common_funcs.jl:312-347).
"""
from __future__ import annotations

import math

import numpy as np

from . import mesh as meshmod
from . import sbp as sbpmod
from .abi import Physics, PDES_FLUX_ROE, PDES_FLUX_IRSLF, PDES_FLUX_ESLF, PDES_BC_FREESTREAM, PDES_BC_ISENTROPIC_VORTEX


def vortex_exact(coords: np.ndarray, phys: Physics) -> np.ndarray:
    x = coords[..., 0]
    y = coords[..., 1]
    gamma = phys.gamma
    R = phys.R_gas
    cv = R / (gamma - 1)
    r_in, rho_in, M_in, p_in = 1.0, 2.0, phys.vortex_Min, 1.0 / gamma
    r = np.sqrt(x * x + y * y)
    theta = np.arctan2(y, x)
    tmp1 = ((gamma - 1) / 2) * M_in * M_in
    rho_r = rho_in * (1 + tmp1 * (1 - (r_in * r_in) / (r * r))) ** (1 / (gamma - 1))
    p_r = p_in * (rho_r / rho_in) ** gamma
    a_r = np.sqrt(gamma * p_r / rho_r)
    M_r = np.sqrt((2 / (gamma - 1)) * ((rho_in / rho_r) ** (gamma - 1)) * (1 + tmp1) - 2 / (gamma - 1))
    U_r = M_r * a_r
    u_r = U_r * np.sin(theta)
    v_r = -U_r * np.cos(theta)
    e_r = cv * p_r / (rho_r * R)
    E_r = rho_r * e_r + 0.5 * rho_r * U_r * U_r
    return np.stack([rho_r, rho_r * u_r, rho_r * v_r, E_r], axis=-1)


def freestream(shape, dim: int, phys: Physics) -> np.ndarray:
    aoa = phys.aoa_deg * math.pi / 180.0
    rho = phys.rho_free
    if dim == 2:
        val = [rho, rho * phys.Ma * math.cos(aoa), rho * phys.Ma * math.sin(aoa), phys.E_free]
    else:
        val = [rho, rho * phys.Ma * math.cos(aoa), 0.0, -rho * phys.Ma * math.sin(aoa), phys.E_free]
    return np.broadcast_to(np.asarray(val), tuple(shape) + (dim + 2,)).copy()


def vortex_noise_state(m: meshmod.DGMesh, phys: Physics, seed: int = 0, amp: float = 1e-3):
    """Exact vortex times (1 + U(-amp, amp)) (config 0 / config 3 initial state)."""
    rng = np.random.default_rng(seed)
    q = vortex_exact(m.coords, phys)
    return q * (1.0 + rng.uniform(-amp, amp, size=q.shape))


def vortex_fourier_state(m: meshmod.DGMesh, phys: Physics, seed: int = 1, amp: float = 1e-2,
                         nmodes: int = 8):
    """Exact vortex plus a seeded sum of Fourier modes (config 1 state):
    q += amp * mean|q_k| * sum_m sin(2 pi (a_m x + b_m y) + phi_m)."""
    rng = np.random.default_rng(seed)
    q = vortex_exact(m.coords, phys)
    a = rng.integers(1, 9, size=nmodes)
    b = rng.integers(1, 9, size=nmodes)
    ph = rng.uniform(0, 2 * math.pi, size=nmodes)
    x = m.coords[..., 0]
    y = m.coords[..., 1]
    pert = np.zeros_like(x)
    for k in range(nmodes):
        pert += np.sin(2 * math.pi * (a[k] * x + b[k] * y) + ph[k])
    scale = np.abs(q).reshape(-1, q.shape[-1]).mean(axis=0)
    return q + amp * pert[..., None] * scale


def freestream_bump_state(m: meshmod.DGMesh, phys: Physics, seed: int = 2, amp: float = 1e-2):
    """Freestream plus a smooth Gaussian perturbation of density and velocity
    (config 2 initial state), lobe centre jittered by the seed.  Total energy
    is kept at the freestream value so the pressure stays positive."""
    rng = np.random.default_rng(seed)
    c = 0.5 + rng.uniform(-0.05, 0.05, size=m.dim)
    q = freestream(m.coords.shape[:-1], m.dim, phys)
    r2 = ((m.coords - c) ** 2).sum(axis=-1)
    g = amp * phys.rho_free * np.exp(-r2 / 0.05)
    rho = q[..., 0] * (1 + g)
    for d in range(m.dim):
        u = q[..., 1 + d] / q[..., 0]
        q[..., 1 + d] = rho * (u + g)
    q[..., 0] = rho
    return q


def make_problem(kind: str, n: int, p: int, op_kind: str = "omega", flux: str = "roe",
                 Ma: float = 0.5, aoa_deg: float = 0.0, vortex_Min: float = 0.95):
    """Returns (mesh, op, phys).  kind: "vortex" (2D quarter annulus, vortex BC
    on all boundaries) or "cube" (3D unit cube, freestream BC).  flux: "roe"
    (weak form, Roe faces), "ir" (diagonal-E: IR split form, IRSLF faces) or
    "es" (Omega: IR split form, entropy-stable element face integral)."""
    if kind == "vortex":
        op = sbpmod.get_operator(2, p, op_kind)
        m = meshmod.quarter_annulus(n, op)
        bc = [PDES_BC_ISENTROPIC_VORTEX] * m.numBC
    elif kind == "cube":
        op = sbpmod.get_operator(3, p, op_kind)
        m = meshmod.unit_cube(n, op)
        bc = [PDES_BC_FREESTREAM] * m.numBC
    else:
        raise ValueError(kind)
    return m, op, _physics(op, flux, bc, Ma, aoa_deg, vortex_Min)


def _physics(op, flux, bc, Ma, aoa_deg, vortex_Min):
    if flux == "roe":
        return Physics(Ma=Ma, aoa_deg=aoa_deg, vortex_Min=vortex_Min, volume_integral_type=1,
                       face_flux=PDES_FLUX_ROE, bc_type=bc)
    if flux == "es":
        if op.kind != "omega":
            raise ValueError("the entropy-stable element face integral is for dense-face (Omega) operators")
        return Physics(Ma=Ma, aoa_deg=aoa_deg, vortex_Min=vortex_Min, volume_integral_type=2,
                       face_flux=PDES_FLUX_ESLF, bc_type=bc)
    if flux == "ir":
        if op.kind != "diage":
            raise ValueError("the IR split form with IRSLF face flux needs a diagonal-E operator")
        return Physics(Ma=Ma, aoa_deg=aoa_deg, vortex_Min=vortex_Min, volume_integral_type=2,
                       face_flux=PDES_FLUX_IRSLF, bc_type=bc)
    raise ValueError(flux)


def smb_problem(path: str, p: int, op_kind: str = "omega", flux: str = "roe", bc: str = "vortex",
                bc_tags=None, Ma: float = 0.5, aoa_deg: float = 0.0, vortex_Min: float = 0.95):
    """A problem on a serial PUMI mesh file (the reference's `smb_name`): the
    dimension comes from the file, bc "vortex" (isentropicVortexBC, 2D) or
    "freestream" on every boundary group.  Returns (mesh, op, phys)."""
    from . import pumi
    dim = pumi.read_smb(path).dim
    op = sbpmod.get_operator(dim, p, op_kind)
    m = pumi.load_smb(path, op, bc_tags=bc_tags)
    code = {"vortex": PDES_BC_ISENTROPIC_VORTEX, "freestream": PDES_BC_FREESTREAM}[bc]
    if code == PDES_BC_ISENTROPIC_VORTEX and dim != 2:
        raise ValueError("the isentropic vortex is two-dimensional")
    return m, op, _physics(op, flux, [code] * m.numBC, Ma, aoa_deg, vortex_Min)


def smb_vortex_problem(path: str, p: int, op_kind: str = "omega", flux: str = "roe",
                       bc_tags=None, vortex_Min: float = 0.95):
    """Steady isentropic vortex on a serial PUMI mesh file, isentropicVortexBC
    on every boundary group (the convergence cases of
    test/euler/convergence/p*/conservative_dg, BC1 => [0, 1, 2, 3]).
    Returns (mesh, op, phys)."""
    return smb_problem(path, p, op_kind, flux, "vortex", bc_tags, vortex_Min=vortex_Min)


def config3_instances(count: int = 96, seed: int = 3, nmin: int = 8, nmax: int = 128, amp: float = 1e-3):
    """Batched uneven solves: `count` 2D vortex problems with the grid edge
    count uniform in nmin..nmax, operator Omega or diagonal-E and inflow Mach
    uniform in [0.3, 0.95]; state = exact * (1 + U(-amp, amp)).

    Omega instances use the weak form with the Roe face flux and p in {1,2,3};
    diagonal-E instances use the entropy-stable pairing the reference runs
    steady diagonal-E cases with (volume_integral_type 2, IRSLF faces; e.g.
    test/euler/test_jacp.jl:66-95) and p in {2,3}: at p = 1 that
    discretisation has no Newton-reachable steady state on these grids (the
    oracle's direct Newton stalls near 1e-3 with or without the implicit-Euler
    globalization).  Returns a list of (mesh, op, phys, q0)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        kind = "omega" if rng.integers(0, 2) == 0 else "diage"
        p = int(rng.integers(1, 4)) if kind == "omega" else int(rng.integers(2, 4))
        Min = float(rng.uniform(0.3, 0.95))
        m, op, ph = make_problem("vortex", n, p, kind, "roe" if kind == "omega" else "ir", vortex_Min=Min)
        q0 = vortex_noise_state(m, ph, seed=seed * 1000 + i, amp=amp)
        out.append((m, op, ph, q0))
    return out
