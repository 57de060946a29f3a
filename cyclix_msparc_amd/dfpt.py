"""First-order (DFPT) right-hand sides for atomic displacements.

The reference builds, for each perturbation and each occupied state n of a
channel (``src/phonon.m:1358-1444``, "Method-2", q = 0),

    rhs_n = sum_m psi_m (psi_m^H W dH psi_n) coeff_mn - dH psi_n,
    dH = dVeff + sum_J gamma_J (chi_J dchi_J^H + dchi_J chi_J^H) W,

and solves (H - lambda_n + psi_v psi_v^H W) x_n = rhs_n with bicgstab
(``:1516-1521``).  dVeff at the first DFPT iteration is the potential of the
pseudocharge change alone (S.Drho_dR = 0, ``:188,219``).

Set-up here (host, once per perturbation): the displacement patterns, the
pseudocharge change db and the projector changes dchi (closed-form
derivatives of the set-up tables, ``CyclixSystem.pseudocharge_derivative``
and ``projector_derivatives``), the coefficient matrix coeff
(``:112-158``).  On the device: dphi from the Poisson solve of db, the
right-hand sides (``cyx_sternheimer_rhs``) and the batched solves
(``SternheimerPlan``).
"""
from __future__ import annotations

import math

import numpy as np

from .system import CyclixSystem, ProjAtom, ProjImage, to_cart

OCC_TOL = 1e-5  # S.Ns_nzocc = sum(S.occ > 1e-5), src/phonon.m:38


def random_displacements(n_atm, npert, seed=1):
    """``npert`` random Cartesian displacement patterns (n_atm x 3 each,
    N(0,1) entries, unit Frobenius norm) from numpy's default generator."""
    rng = np.random.default_rng(seed)
    U = rng.standard_normal((npert, n_atm, 3))
    return U / np.linalg.norm(U.reshape(npert, -1), axis=1)[:, None, None]


def displaced_atoms(system: CyclixSystem, u, eps):
    """Atom positions (r, theta, z) after the Cartesian displacement eps * u,
    theta kept next to its old value."""
    g = system.grid
    X = to_cart(g, system.atoms) + eps * np.asarray(u, dtype=float).reshape(-1, 3)
    r = np.hypot(X[:, 0], X[:, 1])
    z = X[:, 2]
    th = np.arctan2(X[:, 1], X[:, 0]) - g.twist * z
    th = th + 2 * math.pi * np.round((system.atoms[:, 1] - th) / (2 * math.pi))
    return np.stack([r, th, z], axis=1)


def occupied_count(occ):
    """S.Ns_nzocc for one channel."""
    return int(np.sum(np.asarray(occ) > OCC_TOL))


def coeff_pv_psi(eigval, occ, bet, nv=None, deg_tol=1e-4):
    """S.coeff_Pv_psi_kq(1:nv,1:nv,kpt) for q = 0 (EigVal_kq = EigVal,
    occ_kq = occ; ``src/phonon.m:112-158``), after the transpose at ``:158``:
    entry (m, n) = 1 / (lambda_n - lambda_m) for a non-degenerate pair,
    1 - (1 - occ_n) bet / 2 for a degenerate one (|d lambda| and |d occ|
    below 1e-4), 1 on the diagonal."""
    lam = np.asarray(eigval, dtype=float)
    f = np.asarray(occ, dtype=float)
    nv = occupied_count(f) if nv is None else int(nv)
    lam, f = lam[:nv], f[:nv]
    dl = lam[:, None] - lam[None, :]  # (i, j) = lambda_i - lambda_j
    do = f[:, None] - f[None, :]
    deg = (np.abs(dl) < deg_tol) & (np.abs(do) < deg_tol)
    with np.errstate(divide="ignore"):
        c = np.where(deg, 1 - (1 - f[:, None]) * bet / 2, 1.0 / np.where(deg, 1.0, dl))
    np.fill_diagonal(c, 1.0)
    return c.T.copy()


def perturbation_projectors(proj, dchi):
    """Projector tables whose KB term is the first-order nonlocal operator
    sum_J gamma_J (chi dchi^H + dchi chi^H) W: per image the columns
    (chi + dchi)/sqrt(2) and (chi - dchi)/sqrt(2) with gammas +gamma, -gamma
    (their difference of squares is the symmetric product)."""
    out = []
    s = 1 / math.sqrt(2)
    for A, dA in zip(proj, dchi):
        imgs = [ProjImage(rc_pos=im.rc_pos, chi=np.hstack([(im.chi + d) * s, (im.chi - d) * s]), coords=im.coords)
                for im, d in zip(A.images, dA)]
        out.append(ProjAtom(coords=A.coords, gamma=np.concatenate([A.gamma, -A.gamma]), images=imgs))
    return out


def sternheimer_rhs_batch(dev, system: CyclixSystem, kpt, psi, eigval, occ, bet, poisson, U, proj=None,
                          poisson_tol=1e-7, poisson_maxit=3000):
    """Right-hand sides of every displacement pattern in U (npert x n_atm x 3)
    for the occupied states of one channel, concatenated: pattern p owns
    columns [p nv, (p + 1) nv).  psi: (ns, N) device block of the channel's
    states (ascending eigenvalues), eigval and occ: host arrays.  Returns
    (B, lam, dlambda, dveff): the (npert nv, N) block, the per-column shifts
    lambda_n, the first-order eigenvalues (npert x nv) and the dVeff of each
    pattern (device, FP64).  dVeff is solved to poisson_tol: Jacobi-AAR
    stalls near 1e-8 on these dipole-like charges at 120 x 24 x 96 (the
    oracle's AAR reaches 1.2e-8 after 3000 iterations)."""
    import torch

    from .api import Hamiltonian

    nv = occupied_count(occ)
    coeff = coeff_pv_psi(eigval, occ, bet, nv)
    proj = system.projectors() if proj is None else proj
    N = system.grid.N
    npert = len(U)
    B = torch.empty((npert * nv, N), dtype=psi.dtype, device=psi.device)
    dl = np.zeros((npert, nv))
    dveffs = []
    zero = torch.zeros(N, dtype=torch.float64, device=psi.device)
    for p, u in enumerate(U):
        db = torch.from_numpy(system.pseudocharge_derivative(u)).to(psi.device)
        dphi, info = poisson.solve(zero, db, tol=poisson_tol, max_iter=poisson_maxit)
        if not info.converged:
            raise RuntimeError(f"Poisson solve of the pseudocharge change did not reach {poisson_tol} "
                               f"({info.iterations} iterations, relres {info.relres:.3e})")
        dveffs.append(dphi)
        dham = Hamiltonian(dev, system, kpt, dphi, perturbation_projectors(proj, system.projector_derivatives(u, proj)))
        try:
            _, dl[p] = dham.sternheimer_rhs(psi, nv, coeff, out=B[p * nv:(p + 1) * nv])
        finally:
            dham.close()
    lam = np.tile(np.asarray(eigval, dtype=float)[:nv], npert)
    return B, lam, dl, dveffs
