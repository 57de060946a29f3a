"""First-order (DFPT) right-hand sides for atomic displacements.

The reference builds, for each perturbation and each occupied state n of a
channel (``src/phonon.m:1358-1444``, "Method-2", q = 0),

    rhs_n = sum_m psi_m (psi_m^H W dH psi_n) coeff_mn - dH psi_n,
    dH = dVeff + sum_J gamma_J (chi_J dchi_J^H + dchi_J chi_J^H) W,

and solves (H - lambda_n + psi_v psi_v^H W) x_n = rhs_n with bicgstab
(``:1516-1521``), inside the self-consistency loop over dVeff
(``:195-366``, ``response_scf``).

Set-up here (host, once per perturbation): the displacement patterns, the
pseudocharge change db, the atomic-density change (the loop's starting
Drho) and the projector changes dchi (closed-form derivatives of the set-up
tables, ``CyclixSystem.pseudocharge_derivative``,
``atomic_density_derivative`` and ``projector_derivatives``), the
coefficient matrix coeff (``:112-158``).  On the device: the right-hand
sides (``cyx_sternheimer_rhs``), the batched solves (``SternheimerPlan``),
the density response (``cyx_density_response``), Dphi (AAR), the
exchange-correlation kernel (``cyx_lda_pw_fxc``) and the mixing.
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


def sternheimer_rhs_batch(dev, system: CyclixSystem, kpt, psi, eigval, occ, bet, rho, U, proj=None,
                          rhotol=1e-14):
    """Right-hand sides of every displacement pattern in U (npert x n_atm x 3)
    for the occupied states of one channel at the first iteration of the
    DFPT loop, concatenated: pattern p owns columns [p nv, (p + 1) nv).
    The first iteration's potential is the reference's starting point
    (``src/phonon.m:195-219``): Drho = the atomic-density guess
    (``CyclixSystem.atomic_density_derivative``), Dphi = 0, so
    DVeff = fxc(rho) Drho with the LDA-PW kernel at the ground-state density
    ``rho`` (device FP64).  psi: (ns, N) device block of the channel's
    states (ascending eigenvalues), eigval and occ: host arrays.  Returns
    (B, lam, dlambda, dveff): the (npert nv, N) block, the per-column shifts
    lambda_n, the first-order eigenvalues (npert x nv) and the DVeff of each
    pattern (device, FP64)."""
    import torch

    from .api import Hamiltonian, lda_pw_fxc

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
        drho0 = torch.from_numpy(system.atomic_density_derivative(u)).to(psi.device)
        _, dveff = lda_pw_fxc(dev, rho, drho0, zero, rhotol, fxc=False)
        dveffs.append(dveff)
        dham = Hamiltonian(dev, system, kpt, dveff, perturbation_projectors(proj, system.projector_derivatives(u, proj)))
        try:
            _, dl[p] = dham.sternheimer_rhs(psi, nv, coeff, out=B[p * nv:(p + 1) * nv])
        finally:
            dham.close()
    lam = np.tile(np.asarray(eigval, dtype=float)[:nv], npert)
    return B, lam, dl, dveffs


def cartesian_patterns(n_atm):
    """The reference's perturbations: one unit Cartesian displacement per
    atom and direction, scf_natmc = 3 (J - 1) + dir (``src/phonon.m:235,1283-1286``)."""
    U = np.zeros((3 * n_atm, n_atm, 3))
    for k in range(3 * n_atm):
        U[k, k // 3, k % 3] = 1.0
    return U


class ResponseResult:
    """Self-consistent first-order response to one displacement pattern:
    Drho, Dphi and the last output DVeff (device FP64), the first-order
    eigenvalues (per channel, host) and Fermi level, the Sternheimer
    solutions Dpsi (per channel, device blocks of the occupied states), the
    iteration count, the SCF errors and the solver iteration counts."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def response_scf(gs, u, SCF_tol=1e-5, poisson_tol=None, lin_tol=1e-8, max_scf=100, min_scf=3, lin_maxit=3000,
                 poisson_maxit=1000, dpsi0=None, drho0=None, proj=None, poll=8, verbose=False, precond="jacobi"):
    """The DFPT self-consistency loop of ``src/phonon.m:195-366`` (q = 0,
    potential mixing) for one displacement pattern u (n_atm x 3) on a
    converged FP64 ground state ``gs`` (``scf.GroundState`` after ``run``).

    Start (``:195-256``): Drho = the atomic-density guess
    (``CyclixSystem.atomic_density_derivative``), Dphi = 0,
    DVeff = Dphi + fxc Drho, a cleared Periodic-Pulay history.  Each
    iteration (``:259-359``), all on the device except the Ns-sized scalars:
    per channel the Sternheimer right-hand sides of DVeff and the projector
    changes (``cyx_sternheimer_rhs``) and their BiCGSTAB solves warm-started
    from the previous solutions (a ``SternheimerPlan``; all ones at the
    start, ``:224``); the first-order occupations (``occupation_pert``,
    host); the density response (``cyx_density_response``); Dphi by
    AAR on -4 pi (Drho + Db) from the previous Dphi, with the ground
    state's Poisson preconditioner; the output
    DVeff = Dphi + fxc Drho (``cyx_lda_pw_fxc``); the error
    ||DVeff_out - DVeff_in|| / ||DVeff_out||; Periodic-Pulay mixing unless
    converged.  SCF_tol 1e-5, Poisson tolerance SCF_tol 1e-3 and the linear
    solver tolerance 1e-8 are the reference's (``src/phonon.m:11-13``).
    ``precond`` is the BiCGSTAB preconditioner where the reference passes the
    ILU factors of Lap_std (``:1519``): "jacobi" (the diagonal of -1/2 Lap,
    the default: at 120 x 24 x 96 the exact inverse below leaves the
    BiCGSTAB counts about where they are, the solves being bound by the
    small gaps around each shift rather than the Laplacian's spread, and
    costs four transforms per column and iteration), "spectral" (the exact
    Lap_std^-1) or "none"."""
    import torch

    from .api import Hamiltonian, Mixer, SternheimerPlan, lda_pw_fxc
    if gs.precision != "fp64":
        raise ValueError("the response runs on an FP64 ground state")
    if gs.result is None:
        raise RuntimeError("the ground state has not been solved")
    system = gs.system
    dev = gs.dev
    tdev = dev.torch_device
    o = gs.opt
    poisson_tol = SCF_tol * 1e-3 if poisson_tol is None else poisson_tol
    u = np.asarray(u, dtype=float).reshape(system.n_atm, 3)
    proj = system.projectors() if proj is None else proj
    EigVal, occ = gs.result.EigVal, gs.result.occ
    nch = len(gs.hams)
    nvs = [occupied_count(occ[:, k]) for k in range(nch)]
    coeffs = [coeff_pv_psi(EigVal[:, k], occ[:, k], o.bet, nvs[k]) for k in range(nch)]
    pairs = perturbation_projectors(proj, system.projector_derivatives(u, proj))
    db = torch.from_numpy(system.pseudocharge_derivative(u)).to(tdev)
    if drho0 is None:
        drho0 = system.atomic_density_derivative(u)
    drho = torch.from_numpy(np.ascontiguousarray(drho0, dtype=float)).to(tdev)
    dphi = torch.zeros(gs.N, dtype=torch.float64, device=tdev)
    _, dveff_in = lda_pw_fxc(dev, gs.rho, drho, dphi, o.rhotol, fxc=False)
    dhams = [Hamiltonian(dev, system, gs.kpts[k], dveff_in, pairs) for k in range(nch)]
    mixer = Mixer(dev, dveff_in, o.MixingHistory, o.MixingParameter, o.MixingParameterSimple, o.PulayFrequency)
    if dpsi0 is None:
        dpsi = [torch.ones((nvs[k], gs.N), dtype=gs.psi[k].dtype, device=tdev) for k in range(nch)]
    else:
        dpsi = [d.clone().contiguous() for d in dpsi0]
    B = [torch.empty((nvs[k], gs.N), dtype=gs.psi[k].dtype, device=tdev) for k in range(nch)]
    lam = [np.asarray(EigVal[:nvs[k], k], dtype=float) for k in range(nch)]
    err, count = 100.0, 1
    errors, lin_its, aar_its = [], [], []
    dlams = [None] * nch
    dlf = 0.0
    dveff_out = dveff_in
    try:
        while (err > SCF_tol and count <= max_scf) or count <= min_scf:
            its_k = []
            for k in range(nch):
                dhams[k].set_veff(dveff_in)
                _, dlams[k] = dhams[k].sternheimer_rhs(gs.psi[k], nvs[k], coeffs[k], out=B[k])
                sp = SternheimerPlan(gs.hams[k], B[k], lam[k], gs.psi[k][:nvs[k]], X0=dpsi[k], tol=lin_tol,
                                     precond=precond)
                try:
                    act, info, its, rel = sp.solve(max_iter=lin_maxit, poll=poll)
                    dpsi[k] = sp.X
                finally:
                    sp.close()
                its_k.append(int(info.max_iterations))
            lin_its.append(its_k)
            dlf, dgs = _occupation_pert([occ[:nvs[k], k] for k in range(nch)], gs.wkpt, dlams, o.bet)
            drho.zero_()
            for k in range(nch):
                gs.hams[k].density_response(drho, gs.psi[k], dpsi[k], occ[:nvs[k], k], dgs[k], gs.wkpt[k])
            dphi, pinfo = gs.poisson.solve(drho, db, dphi, poisson_tol, poisson_maxit)
            aar_its.append(int(pinfo.iterations))
            _, dveff_out = lda_pw_fxc(dev, gs.rho, drho, dphi, o.rhotol, fxc=False)
            _, dv2, v2 = gs.hams[0].scf_terms(dveff_out, dveff_in, drho)
            err = math.sqrt(dv2 / v2)
            errors.append(err)
            if verbose:
                print(f" DFPT SCF {count:3d}  err = {err:.4e}  bicgstab {its_k}  AAR {pinfo.iterations}", flush=True)
            if err > SCF_tol or count < o.MINIT_SCF:
                dveff_in = mixer.step(dveff_out, dveff_in, count)
            else:
                dveff_in = dveff_out
            count += 1
    finally:
        mixer.close()
        for h in dhams:
            h.close()
    return ResponseResult(drho=drho, dphi=dphi, dveff=dveff_out, dlambda=dlams, dlambda_f=dlf, dpsi=dpsi,
                          count=count - 1, errors=errors, converged=err <= SCF_tol,
                          sternheimer_iterations=lin_its, aar_iterations=aar_its)


def _occupation_pert(occs, wkpt, dlams, bet):
    """First-order Fermi level and occupations (``src/phonon.m:1254-1276``):
    G = occ (1 - occ); DlambdaF = sum w G.dlambda / sum w G; Dg = -bet G
    (dlambda - DlambdaF); both 0 when sum w G < 1e-4."""
    num = den = 0.0
    for oc, dl, wk in zip(occs, dlams, wkpt):
        G = oc * (1 - oc)
        num += wk * float(np.dot(G, dl))
        den += wk * float(np.sum(G))
    if den < 1e-4:
        return 0.0, [np.zeros_like(oc) for oc in occs]
    dlf = num / den
    return dlf, [-bet * oc * (1 - oc) * (dl - dlf) for oc, dl in zip(occs, dlams)]


def response_all(gs, patterns=None, verbose=False, **kw):
    """The reference's loop over perturbations (``src/phonon.m:235``):
    ``patterns`` (default: the 3 n_atm Cartesian unit displacements), each
    solved by ``response_scf`` with the Sternheimer solutions of the previous
    pattern as the starting guess, as S.Dpsi_dR_kq carries them there."""
    patterns = cartesian_patterns(gs.system.n_atm) if patterns is None else patterns
    out, dpsi = [], None
    for p, u in enumerate(patterns):
        r = response_scf(gs, u, dpsi0=dpsi, verbose=verbose, **kw)
        dpsi = r.dpsi
        out.append(r)
    return out
