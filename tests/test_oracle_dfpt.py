"""First-order (DFPT) right-hand sides of atomic displacements
(src/phonon.m:1358-1444, coefficients :112-158, solves :1516-1521): the
set-up derivatives and the oracle, pinned on the CPU.

* the pseudocharge change db and the projector changes dchi against central
  finite differences of the set-up tables at displaced atoms (generic
  positions, so that no box or node set moves for the small steps);
* the (chi +- dchi)/sqrt(2) projector pairs against the oracle's
  first-order nonlocal term;
* the oracle's right-hand sides and their dense solves against first-order
  perturbation theory of the displaced system itself: the first-order
  eigenvalues equal the finite-difference derivatives of the eigenvalues of
  H(R +- eps u) (Hellmann-Feynman), the solutions equal the derivatives of
  the eigenvectors (parallel-transport gauge), and the density response
  equals the derivative of the occupied density;
* the coefficient table's branches against hand-computed entries.

The dense checks use the metric weights W' (in which H is self-adjoint) as
the quadrature, so that perturbation theory holds exactly on the grid."""
import math
from dataclasses import replace

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse.linalg as spla

from cyclix_msparc_amd import dfpt
from cyclix_msparc_amd.system import small_system
from oracle import cyclix_oracle as O

SHIFT = np.array([0.0137, 0.0071, 0.0213])  # off the grid lines
BET = 27.211386245988 / 0.2
CASES = {"gamma": (0.0, (0.0, 0.0, 0.0)), "helical_bloch": (0.05, (0.0, 1.0, 0.6))}


def _system(name, nr=12, nt=12, nz=10):
    tw, kpt = CASES[name]
    s = small_system(nr=nr, nt=nt, nz=nz, twist=tw, kpt=kpt)
    return replace(s, atoms=s.atoms + SHIFT)


def _grid(s):
    g = s.grid
    return O.Grid(g.nr, g.nt, g.nz, g.xin, g.dr, g.dt, g.dz, g.twist, g.fdn, g.cell_typ)


def _oatoms(proj):
    return [O.ProjAtom(a.coords, a.gamma, [O.ProjImage(im.rc_pos, im.chi, im.coords) for im in a.images])
            for a in proj]


def _odatoms(proj, dchi):
    return [O.ProjAtom(a.coords, a.gamma, [O.ProjImage(im.rc_pos, d, im.coords) for im, d in zip(a.images, dA)])
            for a, dA in zip(proj, dchi)]


def test_displacements_are_seeded_unit_patterns():
    U = dfpt.random_displacements(8, 3, seed=1)
    assert U.shape == (3, 8, 3)
    assert np.allclose(np.linalg.norm(U.reshape(3, -1), axis=1), 1.0)
    assert np.array_equal(U, dfpt.random_displacements(8, 3, seed=1))


def test_displaced_atoms_move_by_the_cartesian_step():
    s = _system("helical_bloch")
    u = dfpt.random_displacements(s.n_atm, 1, seed=3)[0]
    from cyclix_msparc_amd.system import to_cart
    moved = dfpt.displaced_atoms(s, u, 0.01)
    assert np.allclose(to_cart(s.grid, moved) - to_cart(s.grid, s.atoms), 0.01 * u, atol=1e-14)
    assert np.all(np.abs(moved[:, 1] - s.atoms[:, 1]) < 0.1)


@pytest.mark.parametrize("name", list(CASES))
def test_pseudocharge_derivative_matches_finite_differences(name):
    s = _system(name)
    for u in dfpt.random_displacements(s.n_atm, 2, seed=1):
        db = s.pseudocharge_derivative(u)
        eps = 1e-4
        bp = replace(s, atoms=dfpt.displaced_atoms(s, u, eps)).pseudocharge()["b"]
        bm = replace(s, atoms=dfpt.displaced_atoms(s, u, -eps)).pseudocharge()["b"]
        fd = (bp - bm) / (2 * eps)
        # central-difference error ~ eps^2 (observed 6e-9)
        assert np.max(np.abs(db - fd)) <= 5e-8 * np.max(np.abs(db))


@pytest.mark.parametrize("name", list(CASES))
def test_projector_derivative_matches_finite_differences(name):
    s = _system(name)
    P = s.projectors()
    eps = 1e-4
    for u in dfpt.random_displacements(s.n_atm, 2, seed=1):
        dchi = s.projector_derivatives(u, P)
        Pp = replace(s, atoms=dfpt.displaced_atoms(s, u, eps)).projectors()
        Pm = replace(s, atoms=dfpt.displaced_atoms(s, u, -eps)).projectors()
        worst, scale = 0.0, 0.0
        for J in range(s.n_atm):
            assert len(P[J].images) == len(Pp[J].images) == len(Pm[J].images)
            for k, im in enumerate(P[J].images):
                a, b, c = im, Pp[J].images[k], Pm[J].images[k]
                assert np.array_equal(a.rc_pos, b.rc_pos) and np.array_equal(a.rc_pos, c.rc_pos)
                fd = (b.chi - c.chi) / (2 * eps)
                worst = max(worst, np.max(np.abs(dchi[J][k] - fd)))
                scale = max(scale, np.max(np.abs(dchi[J][k])))
        assert worst <= 5e-8 * scale


@pytest.mark.parametrize("name", list(CASES))
def test_projector_pairs_give_the_first_order_nonlocal_term(name):
    """KB term of the (chi +- dchi)/sqrt(2), +-gamma pairs = the nonlocal part
    of dH psi in the oracle's restatement of src/phonon.m:1365-1444."""
    s = _system(name)
    g = s.grid
    P = s.projectors()
    u = dfpt.random_displacements(s.n_atm, 1, seed=2)[0]
    dchi = s.projector_derivatives(u, P)
    rng = np.random.default_rng(0)
    psi = rng.standard_normal((g.N, 3)) + 1j * rng.standard_normal((g.N, 3))
    pairs = dfpt.perturbation_projectors(P, dchi)
    via_pairs = O.nonlocal_apply(_oatoms(pairs), s.W, psi, s.kpt)
    # the oracle rhs with coeff = 0 and dVeff = 0 is minus the nonlocal dH psi
    rhs, _ = O.sternheimer_rhs(s.W, s.kpt, np.zeros(g.N), _oatoms(P), _odatoms(P, dchi), psi, np.zeros((3, 3)))
    assert np.max(np.abs(via_pairs + rhs)) <= 1e-13 * np.max(np.abs(rhs))


def _dense_case(name):
    s = _system(name)
    g = s.grid
    og = _grid(s)
    Wm = O.metric_weights(og)
    r = np.tile(g.r, g.nt * g.nz)
    veff0 = -1.5 * np.exp(-((r - s.atoms[:, 0].max()) / 1.2) ** 2) + 0.1 * np.cos(np.arange(g.N) * 0.37)
    lu = spla.splu(O.lap_matrix(og).tocsc())
    b0 = s.pseudocharge()["b"]
    sq = np.sqrt(Wm)

    def eig_at(sys_):
        # Veff(R') = Veff0 + phi[b(R')] - phi[b(R)], the Poisson part of the
        # potential following the atoms (Vxc held, S.Drho_dR = 0)
        veff = veff0 + lu.solve(-4 * np.pi * (sys_.pseudocharge()["b"] - b0))
        H = O.Hamiltonian(og, s.kpt, veff, Wm, _oatoms(sys_.projectors())).dense()
        Hs = H * sq[:, None] / sq[None, :]
        ev, V = sla.eigh(0.5 * (Hs + Hs.conj().T))
        return ev, V / sq[:, None], H

    return s, Wm, lu, eig_at


@pytest.mark.parametrize("name", list(CASES))
def test_sternheimer_solution_is_the_eigenvector_derivative(name):
    s, Wm, lu, eig_at = _dense_case(name)
    N = s.grid.N
    nv, eps = 6, 1e-4
    u = dfpt.random_displacements(s.n_atm, 1, seed=1)[0]
    ev, psi, H = eig_at(s)
    assert np.min(np.diff(ev[:nv + 1])) > 1e-3  # non-degenerate occupied states and a gap
    evp, psip, _ = eig_at(replace(s, atoms=dfpt.displaced_atoms(s, u, eps)))
    evm, psim, _ = eig_at(replace(s, atoms=dfpt.displaced_atoms(s, u, -eps)))
    P = psi[:, :nv]

    def gauge(Q):  # phase with <psi_n | psi_n(R')> real and positive
        ov = np.sum(np.conj(P) * Wm[:, None] * Q[:, :nv], 0)
        return Q[:, :nv] * (np.conj(ov) / np.abs(ov))[None, :]

    dlam_fd = (evp[:nv] - evm[:nv]) / (2 * eps)
    dpsi_fd = (gauge(psip) - gauge(psim)) / (2 * eps)
    proj = s.projectors()
    dchi = s.projector_derivatives(u, proj)
    dveff = lu.solve(-4 * np.pi * s.pseudocharge_derivative(u))
    coeff = O.coeff_pv_psi(ev, np.ones(nv), BET, nv)
    rhs, dlam = O.sternheimer_rhs(Wm, s.kpt, dveff, _oatoms(proj), _odatoms(proj, dchi), P.astype(complex), coeff)
    assert np.max(np.abs(dlam - dlam_fd)) <= 1e-7 * np.max(np.abs(dlam_fd))  # observed 7e-8
    X = np.empty_like(rhs)
    for n in range(nv):
        A = H - ev[n] * np.eye(N) + P @ (P.conj().T * Wm[None, :])
        X[:, n] = np.linalg.solve(A, rhs[:, n])
    assert np.max(np.abs(X - dpsi_fd)) <= 1e-6 * np.max(np.abs(dpsi_fd))  # observed 8e-8
    drho = 2 * np.real(np.sum(np.conj(P) * X, 1))
    drho_fd = (np.sum(np.abs(psip[:, :nv]) ** 2, 1) - np.sum(np.abs(psim[:, :nv]) ** 2, 1)) / (2 * eps)
    assert np.max(np.abs(drho - drho_fd)) <= 1e-6 * np.max(np.abs(drho_fd))
    # the batched bicgstab of the oracle reaches the same solutions
    ham = O.Hamiltonian(_grid(s), s.kpt, np.zeros(N), Wm, [])
    ham.apply = lambda Z: H @ Z
    Xb, rel, its, flags = O.sternheimer_solve(ham, ev[:nv], P, rhs, tol=1e-10, maxit=3000, precond=True)
    assert np.all(flags == 0)
    assert np.max(np.abs(Xb - X)) <= 1e-7 * np.max(np.abs(X))


def test_coefficient_table_branches():
    """Hand-computed entries: states 0 and 1 non-degenerate, 1 and 2
    degenerate (|d lambda|, |d occ| < 1e-4) with occ_2 < 1."""
    lam = np.array([-1.0, -0.5, -0.5 + 5e-5])
    occ = np.array([1.0, 1.0, 1.0 - 2e-5])
    bet = 40.0
    c = O.coeff_pv_psi(lam, occ, bet, 3)
    # entry (m, n) = original (n, m): 1 / (lambda_n - lambda_m) when non-degenerate
    assert c[1, 0] == pytest.approx(1 / (-1.0 + 0.5))
    assert c[0, 1] == pytest.approx(1 / (-0.5 + 1.0))
    assert c[0, 2] == pytest.approx(1 / (-0.5 + 5e-5 + 1.0))
    # degenerate: 1 - (1 - occ_n) bet / 2 with n the column
    assert c[2, 1] == pytest.approx(1 - (1 - 1.0) * bet / 2)
    assert c[1, 2] == pytest.approx(1 - (2e-5) * bet / 2)
    assert np.all(np.diag(c) == 1.0)
    assert np.allclose(dfpt.coeff_pv_psi(lam, occ, bet, 3), c, rtol=0, atol=1e-15)
    rng = np.random.default_rng(5)
    lam = np.sort(rng.standard_normal(12))
    lam[4] = lam[3] + 2e-5
    occ = np.clip(1 - np.exp(lam), 0, 1)
    assert np.allclose(dfpt.coeff_pv_psi(lam, occ, bet, 12), O.coeff_pv_psi(lam, occ, bet, 12), rtol=1e-15, atol=0)
    assert dfpt.occupied_count(np.array([1.0, 0.5, 2e-5, 5e-6])) == 3
