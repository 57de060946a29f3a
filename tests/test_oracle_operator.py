"""Pins of the oracle's grid, FD weights, Laplacian and Hamiltonian.

Each check is independent of the formula it pins: closed forms, exactness on
polynomials, brute force against the explicit Kronecker matrix, the
reference's second (direct-stencil) formulation, a manufactured analytic
solution, Bloch (anti)periodicity on a doubled domain, and W-symmetry.
"""
import math

import numpy as np
import pytest

from oracle import cyclix_oracle as O


def small_grid(helical=False, nr=12, nt=14, nz=13):
    return O.Grid(nr=nr, nt=nt, nz=nz, xin=2.0, dr=0.3, dt=2 * math.pi / 6 / nt, dz=0.35,
                  twist=0.07 if helical else 0.0, fdn=6, cell_typ=5 if helical else 3)


# ----------------------------------------------------------------- FD weights
def test_fd_weights_fdn1_closed_form():
    w2, w1 = O.fd_weights(1)
    assert np.allclose(w2, [-2.0, 1.0]) and np.allclose(w1, [0.0, 0.5])


@pytest.mark.parametrize("fdn", [2, 4, 6])
def test_fd_weights_exact_on_polynomials(fdn):
    w2, w1 = O.fd_weights(fdn)
    k = np.arange(-fdn, fdn + 1)
    W2 = w2[np.abs(k)]
    W1 = np.sign(k) * w1[np.abs(k)]
    for n in range(0, 2 * fdn + 1):
        d2 = np.sum(W2 * k.astype(float) ** n)
        d1 = np.sum(W1 * k.astype(float) ** n)
        assert d2 == pytest.approx(2.0 if n == 2 else 0.0, abs=1e-9)
        assert d1 == pytest.approx(1.0 if n == 1 else 0.0, abs=1e-9)


def test_fd_weight_identity_k_w2_eq_2_w1():
    w2, w1 = O.fd_weights(6)
    k = np.arange(1, 7)
    assert np.allclose(k * w2[1:], 2 * w1[1:], rtol=1e-14)


def test_product_setup_uses_same_weights():
    from cyclix_msparc_amd.system import fd_weights
    for fdn in (1, 3, 6):
        a, b = O.fd_weights(fdn), fd_weights(fdn)
        assert np.allclose(a[0], b[0], rtol=1e-15) and np.allclose(a[1], b[1], rtol=1e-15)


def test_weights_integrate_annulus_exactly():
    g = small_grid()
    W = O.intg_weights(g)
    xout = g.xin + (g.nr - 1) * g.dr
    assert W.sum() == pytest.approx(0.5 * (xout ** 2 - g.xin ** 2) * g.L2 * g.L3, rel=1e-13)


# ----------------------------------------------------------------- Laplacian
@pytest.mark.parametrize("helical", [False, True])
@pytest.mark.parametrize("kpt", [(0, 0, 0), (0, 1.0, 0.0), (0, 2.0, 0.9)])
def test_lap_vec_matches_kronecker_matrix(helical, kpt):
    g = small_grid(helical)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((g.N, 3)) + (1j * rng.standard_normal((g.N, 3)) if any(kpt) else 0)
    F = O.bloch_factors(g, kpt)
    L = O.lap_matrix(g, kpt)
    assert np.allclose(O.lap_vec(F, X, g), L @ X, rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("helical", [False, True])
def test_lap_vec_matches_direct_stencil(helical):
    """Interior agreement with Cychel/pseudochargeDensity_atom_cychel.m's form."""
    g = small_grid(helical, nr=20)
    n0 = g.fdn
    rng = np.random.default_rng(5)
    V = rng.standard_normal((g.nr, g.nt, g.nz))
    V[:n0] = 0.0
    V[-n0:] = 0.0  # Dirichlet zero beyond the box in the Kronecker form
    Xpad = np.zeros((g.nr + 2 * n0, g.nt + 2 * n0, g.nz + 2 * n0))
    Xpad[n0:-n0] = np.pad(V, ((0, 0), (n0, n0), (n0, n0)), mode="wrap")
    direct = O.lap_direct_box(Xpad, g, g.xin - n0 * g.dr)
    F = O.bloch_factors(g)
    kron = O.lap_vec(F, V.reshape(-1, order="F"), g).reshape(g.nr, g.nt, g.nz, order="F")
    assert np.allclose(direct, kron, rtol=1e-11, atol=1e-9)


@pytest.mark.parametrize("helical", [False, True])
def test_lap_manufactured_solution(helical):
    """Order-12 operator vs the analytic helical-coordinate Laplacian."""
    tau = 0.08 if helical else 0.0
    nr, nt, nz = 90, 32, 32
    L2, L3 = 2 * math.pi / 4, 6.0
    g = O.Grid(nr=nr, nt=nt, nz=nz, xin=3.0, dr=0.12, dt=L2 / nt, dz=L3 / nz, twist=tau,
               cell_typ=5 if helical else 3)
    r = g.r[:, None, None]
    th = (np.arange(nt) * g.dt)[None, :, None]
    z = (np.arange(nz) * g.dz)[None, None, :]
    rc, s = g.xin + 0.5 * (nr - 1) * g.dr, 0.7
    a, b = 2 * math.pi / L2, 2 * math.pi / L3
    G = np.exp(-(r - rc) ** 2 / (2 * s * s))
    Gp = -(r - rc) / s ** 2 * G
    Gpp = ((r - rc) ** 2 / s ** 4 - 1 / s ** 2) * G
    f = G * np.cos(a * th) * np.cos(b * z)
    lap = (Gpp + Gp / r) * np.cos(a * th) * np.cos(b * z) \
        - (1 / r ** 2 + tau ** 2) * a * a * f - b * b * f \
        - 2 * tau * G * (a * b) * np.sin(a * th) * np.sin(b * z)
    F = O.bloch_factors(g)
    num = O.lap_vec(F, f.reshape(-1, order="F"), g).reshape(nr, nt, nz, order="F")
    err = np.max(np.abs(num - lap)) / np.max(np.abs(lap))
    assert err < 1e-6, err


@pytest.mark.parametrize("axis", ["theta", "z"])
def test_bloch_phase_equals_antiperiodic_extension(axis):
    """Channel with e^{ikL} = -1 equals the k = 0 operator on the doubled domain."""
    g = small_grid(True)
    if axis == "theta":
        kpt = (0, math.pi / g.L2, 0.0)
        g2 = O.Grid(g.nr, 2 * g.nt, g.nz, g.xin, g.dr, g.dt, g.dz, g.twist, g.fdn, g.cell_typ)
    else:
        kpt = (0, 0.0, math.pi / g.L3)
        g2 = O.Grid(g.nr, g.nt, 2 * g.nz, g.xin, g.dr, g.dt, g.dz, g.twist, g.fdn, g.cell_typ)
    rng = np.random.default_rng(9)
    X = rng.standard_normal((g.nr, g.nt, g.nz)) + 1j * rng.standard_normal((g.nr, g.nt, g.nz))
    X2 = np.concatenate([X, -X], axis=1 if axis == "theta" else 2)
    out = O.lap_vec(O.bloch_factors(g, kpt), X.reshape(-1, order="F"), g).reshape(X.shape, order="F")
    out2 = O.lap_vec(O.bloch_factors(g2), X2.reshape(-1, order="F"), g2).reshape(X2.shape, order="F")
    half = out2[:, :g.nt] if axis == "theta" else out2[:, :, :g.nz]
    assert np.allclose(out, half, rtol=1e-12, atol=1e-10)


def test_w_symmetry_interior():
    """W*Lap is symmetric except for the two end-corrected radial planes
    (the identity k w2_k = 2 w1_k makes r(D_rr + r^-1 D_r) symmetric)."""
    g = small_grid(True, nr=16)
    W = O.intg_weights(g)
    WL = (O.lap_matrix(g).toarray()) * W[:, None]
    i = np.arange(g.N) % g.nr
    inner = (i > 0) & (i < g.nr - 1)
    D = (WL - WL.T)[np.ix_(inner, inner)]
    assert np.max(np.abs(D)) < 1e-10 * np.max(np.abs(WL))


# ----------------------------------------------------------------- KB term
def _dense_kb(atoms, W, kpt, N):
    kpt = np.asarray(kpt, float)
    fac = 1j if np.any(kpt) else 1.0
    V = np.zeros((N, N), dtype=complex)
    for A in atoms:
        P = np.zeros((N, A.gamma.size), dtype=complex)
        for im in A.images:
            ph = np.exp(-np.dot(kpt, (A.coords - im.coords) * fac))
            np.add.at(P, im.rc_pos, im.chi * ph)
        V += P @ np.diag(A.gamma) @ P.conj().T @ np.diag(W)
    return V


@pytest.mark.parametrize("kpt", [(0, 0, 0), (0, 1.0, 0.7)])
def test_nonlocal_matches_dense_projector(kpt):
    from cyclix_msparc_amd.system import small_system
    sysm = small_system(twist=0.05, kpt=kpt)
    atoms = [O.ProjAtom(a.coords, a.gamma, [O.ProjImage(im.rc_pos, im.chi, im.coords) for im in a.images])
             for a in sysm.projectors()]
    N = sysm.grid.N
    rng = np.random.default_rng(1)
    X = rng.standard_normal((N, 2)) + 1j * rng.standard_normal((N, 2))
    dense = _dense_kb(atoms, sysm.W, kpt, N)
    assert np.allclose(O.nonlocal_apply(atoms, sysm.W, X, kpt), dense @ X, rtol=1e-11, atol=1e-12)
