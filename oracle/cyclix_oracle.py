"""FP64 CPU restatement of the Cyclix-M-SPARC CheFSI + AAR-Poisson hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP path is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package
(``cyclix_msparc_amd``) never does.

Every function restates one reference routine and cites it (paths relative
to the reference root ``Cyclix_MS_200220/``).  Conventions follow MATLAB:

* a grid function is a vector of length N = Nr*Nt*Nz, radial index fastest,
  then theta, then z (``src/calculate_nloc_projector.m:145``);
* a block of states is an (N, Ns) array in column-major (Fortran) order, i.e.
  the same bytes as MATLAB's ``psi(:,:,ks)`` (``src/scf.m:82``);
* ``kpt`` is the reference's ``kptvec = [k_r, nu, eta]`` with k_r = 0
  (Dirichlet radial direction), ``nu`` the integer cyclic channel and
  ``eta`` the helical/axial Bloch wavenumber (``src/initialization.m:1624-1641``).

Parity pins (what fixes each function independently of its own formula) are
listed in ``tests/test_oracle_*.py``.  One branch is "parity unpinned": the
degenerate-pair entry 1 - (1 - occ_n) bet / 2 of ``coeff_pv_psi`` for
partially occupied states (checked only by hand-computed entries of the same
formula; with full occupation it is 1 and the perturbation-theory pin of the
right-hand sides covers it).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla


# --------------------------------------------------------------------------
# finite-difference weights and grid
# --------------------------------------------------------------------------
def fd_weights(fdn: int):
    """Central FD weights of order 2*fdn.  ``src/initialization.m:403-419``.

    Returns (w2, w1): w2[0..fdn] second-derivative weights (w2[0] centre),
    w1[0..fdn] first-derivative weights (w1[0] = 0).
    """
    w2 = np.zeros(fdn + 1)
    w1 = np.zeros(fdn + 1)
    f = math.factorial
    for k in range(1, fdn + 1):
        w2[k] = (2 * (-1) ** (k + 1)) * (f(fdn) ** 2) / (k * k * f(fdn - k) * f(fdn + k))
        w2[0] -= 2.0 / (k * k)
        w1[k] = ((-1) ** (k + 1)) * (f(fdn) ** 2) / (k * f(fdn - k) * f(fdn + k))
    return w2, w1


@dataclass
class Grid:
    """One cyclic/helical sector discretised in (r, theta, z).

    ``nr, nt, nz`` are node counts (S.Nx, S.Ny, S.Nz after the Dirichlet
    node is added, ``src/initialization.m:341-344``); ``dr, dt, dz`` the
    spacings (S.dx, S.dy, S.dz); ``xin`` the inner radius (S.xin,
    ``src/initialization.m:189-193``); ``twist`` the helical twist tau
    (S.twist).  ``cell_typ`` 3 = cyclic, 5 = helical
    (``src/initialization.m:287-295``).
    """

    nr: int
    nt: int
    nz: int
    xin: float
    dr: float
    dt: float
    dz: float
    twist: float = 0.0
    fdn: int = 6
    cell_typ: int = 3

    @property
    def N(self) -> int:
        return self.nr * self.nt * self.nz

    @property
    def L2(self) -> float:
        return self.nt * self.dt

    @property
    def L3(self) -> float:
        return self.nz * self.dz

    @property
    def r(self) -> np.ndarray:
        return self.xin + np.arange(self.nr) * self.dr

    @property
    def helical(self) -> bool:
        return self.cell_typ in (4, 5)


def intg_weights(g: Grid) -> np.ndarray:
    """Quadrature weights W.  ``Cychel/intgWts_cychel.m:2-14`` (BCy = BCz = 0)."""
    wx = (g.xin + np.arange(g.nr) * g.dr) * g.dr
    wx[0] = 0.5 * g.xin * g.dr + 0.125 * g.dr * g.dr
    wx[-1] = 0.5 * (g.xin + (g.nr - 1) * g.dr) * g.dr - 0.125 * g.dr * g.dr
    wy = np.full(g.nt, g.dt)
    wz = np.full(g.nz, g.dz)
    return np.kron(wz, np.kron(wy, wx))


def bloch_channels(nkpt, L1, L2, L3, kptshift=(0.0, 0.0, 0.0), time_rev=True):
    """Cyclix branch of ``Generate_kpts`` (``src/initialization.m:1591-1682``):
    channels [0, nu, eta] with nu = round(N_c j / n2) (N_c = 2 pi / L2, MATLAB
    round-half-away), eta = 2 pi/L3 * (Monkhorst-Pack point + shift / n3) in
    (-1/2, 1/2], weights 1/n; under time reversal a channel whose partner
    (nu' = N_c - nu or -nu, eta' = -eta) appears later in the ndgrid order is
    dropped and the partner's weight doubled.  Returns (list of 3-tuples,
    weights)."""
    n1, n2, n3 = nkpt
    assert n1 == 1
    Nc = 2 * math.pi / L2
    nus = [math.floor(Nc * j / n2 + 0.5) for j in range(n2)]
    start = -((n3 - 1) // 2)
    etas = []
    for k in range(n3):
        f = (start + k) / n3 + kptshift[2] / n3
        f = ((f + 0.5 - np.finfo(float).eps) % 1.0) - 0.5 + np.finfo(float).eps
        etas.append(2 * math.pi / L3 * f)
    chans = [(0.0, float(nu), eta) for eta in etas for nu in nus]
    w = [1.0 / len(chans)] * len(chans)
    if not time_rev:
        return chans, np.array(w)
    keep, wk = [], []
    partnered = set()
    for i, (_, nu_i, eta_i) in enumerate(chans):
        later = [j for j in range(i + 1, len(chans))
                 if (abs(nu_i + chans[j][1] - Nc) < 1e-8 or abs(nu_i + chans[j][1]) < 1e-8)
                 and abs(eta_i + chans[j][2]) < 1e-8]
        partnered.update(later)
        if not later:
            keep.append(i)
    out = [chans[i] for i in keep]
    weights = np.array([w[i] * (2.0 if i in partnered else 1.0) for i in keep])
    return out, weights


def metric_weights(g: Grid) -> np.ndarray:
    """W' = r dr dtheta dz without the end corrections of ``intg_weights``.
    The identity k w2_k = 2 w1_k (``src/initialization.m:407-417``) makes
    W' * Lap exactly symmetric, rows of the two radial end planes included,
    so H is self-adjoint in this inner product; Rayleigh-Ritz and the
    Lanczos bounds use it.  (W' differs from W only on those two planes.)"""
    wx = (g.xin + np.arange(g.nr) * g.dr) * g.dr
    return np.kron(np.full(g.nz, g.dz), np.kron(np.full(g.nt, g.dt), wx))


# --------------------------------------------------------------------------
# 1-D operator factors
# --------------------------------------------------------------------------
@dataclass
class Factors:
    """The 1-D factors returned by ``blochLaplacian_1d`` plus S.R2inv."""

    DL11: np.ndarray
    DL22: np.ndarray
    DL33: np.ndarray
    DG1: np.ndarray
    DG2: np.ndarray
    DG3: np.ndarray
    R2inv: np.ndarray  # diagonal, as a vector
    helical: bool


def _band_1d(n, coef, sym, periodic, phase_l=1.0, phase_r=1.0, include_centre=True):
    """Assemble one 1-D banded factor exactly as the triplet loops of
    ``Cychel/lapIndicesValues_1d_cychel.m`` build it and ``sparse()`` sums it.

    coef[q] for q = 0..fdn; sym=True -> value coef[q] on both sides (second
    derivative), sym=False -> +coef[q] on the right, -coef[q] on the left
    (first derivative).  Dirichlet (periodic=False) drops out-of-range
    entries (``:53-54``); periodic wraps them and multiplies the wrapped
    entries by the Bloch phase (``src/blochLaplacian_1d.m:46-77``).
    coef may be per-row (2-D array [n, fdn+1]) for the radial first derivative.
    """
    fdn = coef.shape[-1] - 1
    dtype = np.result_type(coef, np.asarray(phase_l), np.asarray(phase_r))
    M = np.zeros((n, n), dtype=dtype)
    for i in range(n):
        c = coef[i] if coef.ndim == 2 else coef
        if include_centre:
            M[i, i] += c[0]
        for q in range(1, fdn + 1):
            for jj, s in ((i + q, 1.0), (i - q, -1.0)):
                v = c[q] if sym else s * c[q]
                if 0 <= jj < n:
                    M[i, jj] += v
                elif periodic:
                    ph = phase_r if jj >= n else phase_l
                    M[i, jj % n] += v * ph
    return M


def bloch_factors(g: Grid, kpt=(0.0, 0.0, 0.0)) -> Factors:
    """``src/blochLaplacian_1d.m:14-123`` over ``Cychel/lapIndicesValues_1d_cychel.m:9-261``."""
    w2, w1 = fd_weights(g.fdn)
    tau = g.twist
    k2, k3 = kpt[1], kpt[2]
    ph2l = 1.0 if k2 == 0 else np.exp(-1j * k2 * g.L2)
    ph2r = 1.0 if k2 == 0 else np.exp(1j * k2 * g.L2)
    ph3l = 1.0 if k3 == 0 else np.exp(-1j * k3 * g.L3)
    ph3r = 1.0 if k3 == 0 else np.exp(1j * k3 * g.L3)
    r = g.r
    # D_xx, Dirichlet truncation (:17-54)
    DL11 = _band_1d(g.nr, w2 / g.dr ** 2, True, False)
    # (1/r) d/dr with coefficient 1/(r dx) per row (:154-180)
    c1 = np.outer(1.0 / (r * g.dr), w1)
    DG1 = _band_1d(g.nr, c1, False, False, include_centre=False)
    # D_yy periodic with phases (:56-103), times tau^2 for helical (:101-103)
    c22 = w2 / g.dt ** 2 * (tau * tau if g.helical else 1.0)
    DL22 = _band_1d(g.nt, c22, True, True, ph2l, ph2r)
    # -2 tau / dy first derivative (:182-214)
    DG2 = _band_1d(g.nt, w1 * (-2.0 * tau / g.dt), False, True, ph2l, ph2r, include_centre=False)
    # D_zz (:105-148) and d/dz (:217-252)
    DL33 = _band_1d(g.nz, w2 / g.dz ** 2, True, True, ph3l, ph3r)
    DG3 = _band_1d(g.nz, w1 / g.dz, False, True, ph3l, ph3r, include_centre=False)
    # R2inv (:254-261)
    if g.helical:
        R2inv = 1.0 / (tau * tau) / (r * r)
    else:
        R2inv = 1.0 / (r * r)
    return Factors(DL11, DL22, DL33, DG1, DG2, DG3, R2inv, g.helical)


def lap_vec(F: Factors, X: np.ndarray, g: Grid) -> np.ndarray:
    """Matrix-free Laplacian of an (N, Ns) block.  ``Cychel/lapVec_cychel.m:19-53``."""
    X = np.asarray(X)
    ns = X.shape[1] if X.ndim == 2 else 1
    X4 = X.reshape(g.nr, g.nt, g.nz, ns, order="F")
    dtype = np.result_type(X4, F.DL22, F.DL33)
    A = F.DL11 + F.DG1
    # (DL11+DG1)*X1 (:24 / :40)
    out = np.einsum("ia,ajks->ijks", A, X4).astype(dtype)
    # R2inv*X1*DL22.' (:24 / :40)
    xt = np.einsum("jb,ibks->ijks", F.DL22, X4)
    out += F.R2inv[:, None, None, None] * xt
    if F.helical:
        out += xt  # X1*DL22.' (:40)
        # Hlx2 = (X1*DG2.')*DG3.' (:41, :49)
        t = np.einsum("jb,ibks->ijks", F.DG2, X4)
        out += np.einsum("kc,ijcs->ijks", F.DG3, t)
    # X2*DL33.' (:30 / :49)
    out += np.einsum("kc,ijcs->ijks", F.DL33, X4)
    out = out.reshape(g.N, ns, order="F")
    return out if X.ndim == 2 else out[:, 0]


def lap_matrix(g: Grid, kpt=(0.0, 0.0, 0.0)) -> sp.csr_matrix:
    """Explicit sparse Kronecker Laplacian.  ``src/initialization.m:734-741``
    (the reference assembles it at k = 0; the same Kronecker form with the
    Bloch factors of ``blochLaplacian_1d`` gives the channel operator)."""
    F = bloch_factors(g, kpt)
    Iz, It, Ir = sp.identity(g.nz), sp.identity(g.nt), sp.identity(g.nr)
    L = (sp.kron(Iz, sp.kron(It, sp.csr_matrix(F.DL11 + F.DG1)))
         + sp.kron(sp.csr_matrix(F.DL33), sp.kron(It, Ir))
         + sp.kron(Iz, sp.kron(sp.csr_matrix(F.DL22), sp.diags(F.R2inv))))
    if F.helical:
        L = L + sp.kron(Iz, sp.kron(sp.csr_matrix(F.DL22), Ir)) \
              + sp.kron(sp.csr_matrix(F.DG3), sp.kron(sp.csr_matrix(F.DG2), Ir))
    return sp.csr_matrix(L)


def lap_direct_box(V: np.ndarray, g: Grid, xin_box: float) -> np.ndarray:
    """The reference's second, direct-stencil form of the same operator.
    ``Cychel/pseudochargeDensity_atom_cychel.m:1-33``.

    V is a 3-D box (nx, ny, nz) padded by fdn on every side; returns the
    Laplacian on the interior (nx-2fdn, ny-2fdn, nz-2fdn).  xin_box is the
    radius of the box's first (padding) node.
    """
    w2, w1 = fd_weights(g.fdn)
    n0 = g.fdn
    c1, c2, c3 = 1 / g.dr ** 2, 1 / g.dt ** 2, 1 / g.dz ** 2
    c4 = g.twist ** 2 / g.dt ** 2
    c5 = -(2 * g.twist) / (g.dt * g.dz)
    c6 = 1 / g.dr
    nx, ny, nz = V.shape
    I = slice(n0, nx - n0)
    J = slice(n0, ny - n0)
    K = slice(n0, nz - n0)
    rr = (xin_box + np.arange(n0, nx - n0) * g.dr)[:, None, None]

    def sh(a, b, c):
        return V[n0 + a:nx - n0 + a, n0 + b:ny - n0 + b, n0 + c:nz - n0 + c]

    b = w2[0] * (c1 + c2 / (rr * rr) + c3 + c4) * V[I, J, K]
    for p in range(1, n0 + 1):
        b = b + (c1 * w2[p] + c6 * w1[p] / rr) * sh(p, 0, 0) \
              + (c1 * w2[p] - c6 * w1[p] / rr) * sh(-p, 0, 0) \
              + (c2 / (rr * rr) + c4) * w2[p] * (sh(0, p, 0) + sh(0, -p, 0)) \
              + c3 * w2[p] * (sh(0, 0, p) + sh(0, 0, -p))
        for q in range(1, n0 + 1):
            coeff = c5 * w1[p] * w1[q]
            b = b + coeff * (sh(0, p, q) - sh(0, p, -q) - sh(0, -p, q) + sh(0, -p, -q))
    return b


# --------------------------------------------------------------------------
# nonlocal (Kleinman-Bylander) projectors and the Hamiltonian
# --------------------------------------------------------------------------
@dataclass
class ProjImage:
    """``S.Atom(J).rcImage(img)``: 0-based node indices, Chi_mat, coordinates."""

    rc_pos: np.ndarray  # int, 0-based
    chi: np.ndarray  # (n_nodes, angnum) real
    coords: np.ndarray  # (3,) image position in (r, theta, z)


@dataclass
class ProjAtom:
    coords: np.ndarray  # (3,) S.Atoms(J,:)
    gamma: np.ndarray  # (angnum,) S.Atom(J).gamma_Jl
    images: list = field(default_factory=list)


def nonlocal_apply(atoms, W, X, kpt):
    """KB term  sum_J chi_J gamma_J chi_J^H W X  with image Bloch phases.
    ``src/h_nonlocal_vector_mult.m:24-43``."""
    kpt = np.asarray(kpt, dtype=float)
    fac = 1.0 if not np.any(kpt) else 1.0j
    X2 = X if X.ndim == 2 else X[:, None]
    dtype = np.result_type(X2, complex if np.any(kpt) else float)
    out = np.zeros(X2.shape, dtype=dtype)
    for A in atoms:
        C = np.zeros((A.gamma.size, X2.shape[1]), dtype=dtype)
        for im in A.images:
            ph = np.exp(np.dot(kpt, (A.coords - im.coords) * fac))
            C += (im.chi * W[im.rc_pos][:, None]).conj().T @ X2[im.rc_pos, :] * ph
        C *= A.gamma[:, None]
        for im in A.images:
            ph = np.exp(-np.dot(kpt, (A.coords - im.coords) * fac))
            np.add.at(out, im.rc_pos, (im.chi * ph) @ C)
    return out if X.ndim == 2 else out[:, 0]


Y00 = 0.5 / np.sqrt(np.pi)  # src/sphericalHarmonics.m:64-67


def to_cartesian(g: Grid, rtz):
    """(r, theta, z) -> (x, y, z) with the helical angle theta + twist z.
    ``Cychel/coordinateTransformation_cychel.m:3-7``."""
    rtz = np.atleast_2d(np.asarray(rtz, dtype=float))
    ang = rtz[:, 1] + g.twist * rtz[:, 2]
    return np.stack([rtz[:, 0] * np.cos(ang), rtz[:, 0] * np.sin(ang), rtz[:, 2]], axis=1)


def projector_tables_bruteforce(g: Grid, atoms_rtz, radial, rc, max_image=3):
    """s-channel KB tables by brute force: for every atom and every periodic
    image (theta shifts of L2 and z shifts of L3 up to +-max_image), every
    grid node within Cartesian distance rc of the image, with
    chi = radial(d) * Y00.  ``src/calculate_nloc_projector.m:23-287`` finds
    the same sets through rb boxes around each image; this scans every node.
    Returns {(atom, ky, kz): (sorted 0-based nodes, chi)}."""
    ii, jj, kk = np.meshgrid(np.arange(g.nr), np.arange(g.nt), np.arange(g.nz), indexing="ij")
    ii, jj, kk = (a.ravel(order="F") for a in (ii, jj, kk))
    node = kk * g.nr * g.nt + jj * g.nr + ii
    P = to_cartesian(g, np.stack([g.xin + ii * g.dr, jj * g.dt, kk * g.dz], axis=1))
    out = {}
    for J, x0 in enumerate(np.atleast_2d(atoms_rtz)):
        for kz in range(-max_image, max_image + 1):
            for ky in range(-max_image, max_image + 1):
                c = to_cartesian(g, [x0[0], x0[1] + ky * g.L2, x0[2] + kz * g.L3])[0]
                d = np.linalg.norm(P - c, axis=1)
                sel = d <= rc
                if np.any(sel):
                    order = np.argsort(node[sel], kind="stable")
                    out[(J, ky, kz)] = (node[sel][order], radial(d[sel][order]) * Y00)
    return out


@dataclass
class Hamiltonian:
    """Everything ``h_nonlocal_vector_mult`` reads for one channel."""

    g: Grid
    kpt: tuple
    veff: np.ndarray
    W: np.ndarray
    atoms: list
    F: Factors = None

    def __post_init__(self):
        if self.F is None:
            self.F = bloch_factors(self.g, self.kpt)

    def apply(self, X):
        """``src/h_nonlocal_vector_mult.m:21-43``."""
        Xa = np.asarray(X)
        X2 = Xa if Xa.ndim == 2 else Xa[:, None]
        H = -0.5 * lap_vec(self.F, X2, self.g) + self.veff[:, None] * X2
        if self.atoms:
            H = H + nonlocal_apply(self.atoms, self.W, X2, self.kpt)
        return H if Xa.ndim == 2 else H[:, 0]

    def dense(self):
        """Dense N x N matrix (brute force; small grids only)."""
        E = np.eye(self.g.N)
        return self.apply(E)


# --------------------------------------------------------------------------
# CheFSI
# --------------------------------------------------------------------------
def chebyshev_filter(hfun, X, m, a, b, a0):
    """Zhou-Saad scaled three-term recurrence.  ``src/chebyshev_filter.m:24-42``."""
    e = (b - a) / 2
    c = (b + a) / 2
    sigma = e / (a0 - c)
    sigma1 = sigma
    gamma = 2 / sigma1
    HX = hfun(X)
    Y = (sigma1 / e) * (HX - c * X)
    ii = 2
    while ii <= m:
        sigma2 = 1 / (gamma - sigma)
        HX = hfun(Y)
        Ynew = (2 * sigma2 / e) * (HX - c * Y) - ((sigma * sigma2) * X)
        X = Y
        Y = Ynew
        sigma = sigma2
        ii += 1
    return Y


def chebyshev_scalars(m, a, b, a0):
    """The per-step coefficients of the recurrence above, as the GPU path
    consumes them: step 1 Y = s1*(HX - c X); step i Y = al*(HY - c Y) - be*X."""
    e = (b - a) / 2
    c = (b + a) / 2
    sigma = e / (a0 - c)
    gamma = 2 / sigma
    steps = [(sigma / e, 0.0)]
    for _ in range(2, m + 1):
        sigma2 = 1 / (gamma - sigma)
        steps.append((2 * sigma2 / e, sigma * sigma2))
        sigma = sigma2
    return c, steps


def rayleigh_ritz_euclid(ham: Hamiltonian, psi):
    """The reference's post-filter block: orth, Hs = psi'*H*psi, general eig,
    sort, rotate, W-normalise.  ``src/eigSolver.m:99-124`` (cell_typ >= 3:
    no symmetrisation, ``:107-109``)."""
    q = sla.orth(psi)
    assert q.shape[1] == psi.shape[1], "Number of states have changed within SCF"
    Hs = q.conj().T @ ham.apply(q)
    lam, Q = sla.eig(Hs)
    ev = np.real(lam)
    ind = np.argsort(ev, kind="stable")
    ev = ev[ind]
    Q = Q[:, ind]
    p = q @ Q
    scfac = 1.0 / np.sqrt(np.sum(ham.W[:, None] * (p * p.conj()).real, axis=0))
    return ev, p * scfac[None, :]


def rayleigh_ritz_w(ham: Hamiltonian, psi):
    """Metric Rayleigh-Ritz in W' (``metric_weights``): M = psi^H W' psi,
    Hs = psi^H W' H psi, solve the Hermitian-definite pencil (Hs, M), rotate,
    then W-normalise columns as the reference does (``src/eigSolver.m:122-124``).
    Same span as the reference's orth + eig (``src/eigSolver.m:99-119``) and,
    at convergence, the same eigenpairs; this is the form the GPU computes."""
    Wm = metric_weights(ham.g)
    M = psi.conj().T @ (Wm[:, None] * psi)
    Hp = ham.apply(psi)
    Hs = psi.conj().T @ (Wm[:, None] * Hp)
    M = 0.5 * (M + M.conj().T)
    Hs = 0.5 * (Hs + Hs.conj().T)
    ev, Q = sla.eigh(Hs, M)
    p = psi @ Q
    scfac = 1.0 / np.sqrt(np.sum(ham.W[:, None] * (p * p.conj()).real, axis=0))
    return ev, p * scfac[None, :]


def lanczos_bounds(hfun, v0, kmax=300, tol=1e-3, W=None):
    """Extreme Ritz values of H by Lanczos in the inner product with weights
    W (pass ``metric_weights``, in which H is self-adjoint).  Stands in for
    ``eigs(Hfun,N,1,'lr'/'sr')`` at ``src/eigSolver.m:47-63``; stops when both
    extreme Ritz values change by less than tol*max(1,|lambda|) in one step
    (at least 5 steps).  Returns (lambda_min, lambda_max).  The device
    estimator (cyx_lanczos_bounds) runs the same iteration."""
    Wv = np.ones(v0.size) if W is None else W

    def dot(a, b):
        return np.real(np.vdot(a, Wv * b))

    v = v0 / np.sqrt(dot(v0, v0))
    vprev = np.zeros_like(v)
    alpha, beta = [], []
    b = 0.0
    lo = hi = lo_p = hi_p = 0.0
    for j in range(kmax):
        w = hfun(v)
        a = dot(v, w)
        w = w - a * v - b * vprev
        alpha.append(a)
        b = np.sqrt(dot(w, w))
        beta.append(b)
        th = sla.eigvalsh_tridiagonal(np.array(alpha), np.array(beta[:-1])) if j > 0 else np.array([a])
        lo, hi = th[0], th[-1]
        if j >= 4 and abs(lo - lo_p) <= tol * max(1.0, abs(lo)) and abs(hi - hi_p) <= tol * max(1.0, abs(hi)):
            break
        lo_p, hi_p = lo, hi
        if b == 0:
            break
        vprev, v = v, w / b
    return lo, hi


def eig_solver_pass(ham: Hamiltonian, psi, npl, lambda_cutoff, bup, a0, metric="W"):
    """One CheFSI pass for one channel.  ``src/eigSolver.m:98-124``."""
    psi = chebyshev_filter(ham.apply, psi, npl, lambda_cutoff, bup, a0)
    if metric == "W":
        return rayleigh_ritz_w(ham, psi)
    return rayleigh_ritz_euclid(ham, psi)


def filter_bounds(count, ev_prev, bup, a0, lam_lo=None, lam_hi=None, first=True):
    """Filter bounds of ``src/eigSolver.m:47-94``: on the first pass
    bup = 1.01*lambda_max, a0 = lambda_min - 0.1, cutoff = (bup+a0)/2;
    afterwards a0 = min(EigVal), cutoff = max(EigVal) + 0.1."""
    if count == 1:
        bup = 1.01 * lam_hi
        a0 = lam_lo - 0.1
        cut = 0.5 * (bup + a0) if first else np.max(ev_prev) + 0.10
    else:
        a0 = np.min(ev_prev)
        cut = np.max(ev_prev) + 0.10
    return bup, a0, cut


# --------------------------------------------------------------------------
# AAR Poisson
# --------------------------------------------------------------------------
def aar(A, b, x_guess, tol, max_iter, omega, beta, m, p, precond):
    """Alternating Anderson-Richardson.  ``src/aar.m:33-72``.

    A: callable x -> A x;  precond: callable r -> M^{-1} r (the reference's
    ``U\\(L\\r)``; Jacobi r ./ diag(A) for the AAJ variant of
    ``extras/AAR_old.m:45-46,73-75``).  Returns (x, relres, iterations)."""
    N = b.size
    DX = np.zeros((N, m))
    DF = np.zeros((N, m))
    nb = 1 / np.linalg.norm(b)
    x_prev = np.ones(N) if x_guess is None else np.array(x_guess, dtype=float)
    relres = 1 + tol
    count = 1
    x = x_prev
    Xold = Fold = None
    while count <= max_iter and relres > tol:
        res1 = b - A(x_prev)
        relres = np.linalg.norm(res1) * nb
        res = precond(res1)
        if count > 1:
            k = (count - 2) % m
            DX[:, k] = x_prev - Xold
            DF[:, k] = res - Fold
        Xold = x_prev
        Fold = res
        if count % p != 0:
            x_new = x_prev + omega * res
        else:
            x_new = x_prev + beta * res - (DX + beta * DF) @ (np.linalg.pinv(DF.T @ DF) @ (DF.T @ res))
        x = x_prev
        x_prev = x_new
        count += 1
    return x, relres, count - 1


def circulant_symbol(C):
    """Eigenvalues of a circulant n x n matrix C (C[a, b] = c[(b - a) mod n])
    on the Fourier vectors numpy's FFT uses: (C x)^(m) = lam[m] x^(m) with
    lam[m] = sum_d c[d] e^{2 pi i m d / n}."""
    n = C.shape[0]
    return np.fft.ifft(C[0, :]) * n


def lap_spectral_factors(g: Grid):
    """S.Lap_std (k = 0) split by the theta/z Fourier transform: its theta and
    z factors (DL22, DG2, DL33, DG3 of ``src/blochLaplacian_1d.m`` at k = 0)
    are circulant, so on Fourier mode (m, n) Lap_std is the nr x nr radial
    matrix A_mn = (DL11 + DG1) + lam22(m) diag(R2inv) + lam33(n) I
    [+ (lam22(m) + lamG2(m) lamG3(n)) I for helical cells], real for every
    mode.  Returns the stack A (nt, nz, nr, nr)."""
    F = bloch_factors(g)
    l22 = circulant_symbol(F.DL22)
    l33 = circulant_symbol(F.DL33)
    Ar = F.DL11 + F.DG1
    shift = l22[:, None] * 0 + l33[None, :]
    scale = np.broadcast_to(l22[:, None], (g.nt, g.nz))
    if F.helical:
        shift = shift + l22[:, None] + circulant_symbol(F.DG2)[:, None] * circulant_symbol(F.DG3)[None, :]
    assert np.max(np.abs(np.imag(shift))) <= 1e-9 * np.max(np.abs(shift)) + 1e-300
    shift, scale = np.real(shift), np.real(scale)
    A = (Ar[None, None, :, :] + scale[:, :, None, None] * np.diag(F.R2inv)[None, None, :, :]
         + shift[:, :, None, None] * np.eye(g.nr)[None, None, :, :])
    return A


def lap_spectral_solve(g: Grid, r, A=None, Ainv=None):
    """x = Lap_std^{-1} r exactly (k = 0): FFT over theta and z, one radial
    solve per mode (or the product with the per-mode inverses Ainv),
    inverse FFT.  r: (N,) or (N, ns), real or complex."""
    r = np.asarray(r)
    ns = r.shape[1] if r.ndim == 2 else 1
    R4 = r.reshape(g.nr, g.nt, g.nz, ns, order="F")
    Rh = np.fft.fft2(R4, axes=(1, 2))                        # (nr, nt, nz, ns)
    Rh = np.moveaxis(Rh, 0, 2)                                # (nt, nz, nr, ns)
    if Ainv is not None:
        Xh = Ainv @ Rh
    else:
        Xh = np.linalg.solve(lap_spectral_factors(g) if A is None else A, Rh)
    X4 = np.fft.ifft2(np.moveaxis(Xh, 2, 0), axes=(1, 2))
    if not np.iscomplexobj(r):
        X4 = X4.real
    out = X4.reshape(g.N, ns, order="F")
    return out if r.ndim == 2 else out[:, 0]


_SPECTRAL_CACHE = {}


def spectral_inverses(g: Grid):
    """The per-mode inverses of lap_spectral_factors(g), kept per grid."""
    key = (g.nr, g.nt, g.nz, g.xin, g.dr, g.dt, g.dz, g.twist, g.fdn, g.cell_typ)
    if key not in _SPECTRAL_CACHE:
        if len(_SPECTRAL_CACHE) > 8:
            _SPECTRAL_CACHE.clear()
        _SPECTRAL_CACHE[key] = np.linalg.inv(lap_spectral_factors(g))
    return _SPECTRAL_CACHE[key]


def poisson_rhs(rho, b):
    """Cyclix RHS f = -4 pi (rho + b).  ``src/poissonSolve.m:19-23``."""
    return -4 * np.pi * (rho + b)


def poisson_solve(g: Grid, rho, bcharge, tol, phi_guess=None, maxit=1000, precond="jacobi"):
    """``src/poissonSolve.m:17-51``: aar(Lap_std, f, guess, tol, MAXIT, 0.6, 0.6, 7, 6, L, U)."""
    L = lap_matrix(g)
    f = poisson_rhs(rho, bcharge)
    if precond == "jacobi":
        d = L.diagonal()
        pc = lambda r: r / d
    elif precond == "spectral":
        Ainv = spectral_inverses(g)
        pc = lambda r: lap_spectral_solve(g, r, Ainv=Ainv)
    else:
        ilu = spla.spilu(sp.csc_matrix(L), drop_tol=1e-5)
        pc = ilu.solve
    return aar(lambda x: L @ x, f, phi_guess, tol, maxit, 0.6, 0.6, 7, 6, pc)


# --------------------------------------------------------------------------
# SCF glue
# --------------------------------------------------------------------------
_PW = dict(p=1.0, A=0.031091, alpha1=0.21370, beta1=7.5957, beta2=3.5876, beta3=1.6382, beta4=0.49294)


def lda_pw_vxc(rho, rhotol=1e-14):
    """LDA (Perdew-Wang) potential.  ``src/exchangeCorrelationPotential.m:82-106``."""
    rho = np.maximum(rho, rhotol)
    p, A, a1 = _PW["p"], _PW["A"], _PW["alpha1"]
    b1, b2, b3, b4 = _PW["beta1"], _PW["beta2"], _PW["beta3"], _PW["beta4"]
    C3 = 0.9847450218427
    rs = (0.75 / (np.pi * rho)) ** (1 / 3)
    G = 2 * A * (b1 * rs ** 0.5 + b2 * rs + b3 * rs ** 1.5 + b4 * rs ** (p + 1))
    dG = A * (b1 * rs ** -0.5 + 2 * b2 + 3 * b3 * rs ** 0.5 + 2 * (p + 1) * b4 * rs ** p)
    ec = -2 * A * (1 + a1 * rs) * np.log(1 + 1 / G)
    vc = ec - (rs / 3) * (-2 * A * a1 * np.log(1 + 1 / G)
                          - (-2 * A * (1 + a1 * rs)) * dG / (G * G + G))
    return -C3 * rho ** (1 / 3) + vc


def lda_pw_fxc(rho, rhotol=1e-14):
    """dVxc/drho of the LDA (Perdew-Wang) potential ``lda_pw_vxc``: the kernel
    S.Dvxc_drho that the DFPT loop multiplies the density response by
    (``src/phonon.m:215,303``).  The reference sets it only for LDA_PZ through
    libxc (``src/exchangeCorrelationPotential.m:125``); this is the closed-form
    derivative of the PW92 expressions of ``src/exchangeCorrelationPotential.m:82-106``
    (J. P. Perdew and Y. Wang, Phys. Rev. B 45, 13244 (1992)) with respect to
    rho: with rs = (3/(4 pi rho))^(1/3), drs/drho = -rs/(3 rho),
    vc = ec - rs ec'/3 and dvc/drho = (2 ec'/3 - rs ec''/3) drs/drho; the
    exchange part -C3 rho^(1/3) gives -C3 rho^(-2/3)/3.  Below rhotol the
    potential is constant (rho is clamped there), so the kernel is 0."""
    rho = np.asarray(rho, dtype=float)
    clamp = rho < rhotol
    r = np.maximum(rho, rhotol)
    p, A, a1 = _PW["p"], _PW["A"], _PW["alpha1"]
    b1, b2, b3, b4 = _PW["beta1"], _PW["beta2"], _PW["beta3"], _PW["beta4"]
    C3 = 0.9847450218427
    rs = (0.75 / (np.pi * r)) ** (1 / 3)
    G = 2 * A * (b1 * rs ** 0.5 + b2 * rs + b3 * rs ** 1.5 + b4 * rs ** (p + 1))
    G1 = A * (b1 * rs ** -0.5 + 2 * b2 + 3 * b3 * rs ** 0.5 + 2 * (p + 1) * b4 * rs ** p)
    G2 = A * (-0.5 * b1 * rs ** -1.5 + 1.5 * b3 * rs ** -0.5 + 2 * p * (p + 1) * b4 * rs ** (p - 1))
    D = G * G + G
    L = np.log(1 + 1 / G)
    ec1 = -2 * A * a1 * L + 2 * A * (1 + a1 * rs) * G1 / D
    ec2 = 4 * A * a1 * G1 / D + 2 * A * (1 + a1 * rs) * (G2 * D - G1 * G1 * (2 * G + 1)) / (D * D)
    fc = (2 * ec1 / 3 - rs * ec2 / 3) * (-rs / (3 * r))
    fx = -C3 / 3 * r ** (-2 / 3)
    return np.where(clamp, 0.0, fx + fc)


def lda_pw_exc(rho, W, rhotol=1e-14):
    """Exchange-correlation energy.  ``src/evaluateTotalEnergy.m:27-47``."""
    rho = np.maximum(rho, rhotol)
    p, A, a1 = _PW["p"], _PW["A"], _PW["alpha1"]
    b1, b2, b3, b4 = _PW["beta1"], _PW["beta2"], _PW["beta3"], _PW["beta4"]
    C2 = 0.73855876638202
    rs = (0.75 / (np.pi * rho)) ** (1 / 3)
    ec = -2 * A * (1 + a1 * rs) * np.log(1 + 1 / (2 * A * (b1 * rs ** 0.5 + b2 * rs + b3 * rs ** 1.5 + b4 * rs ** (p + 1))))
    return np.sum(ec * rho * W) - C2 * np.sum(rho ** (4 / 3) * W)


def smearing_occ(ev, lam_f, bet, kind):
    """``src/occupations.m:11-15``: 0 = Fermi-Dirac, 1 = Gaussian."""
    if kind == 0:
        return 1.0 / (1.0 + np.exp(bet * (ev - lam_f)))
    from scipy.special import erf
    return 0.5 * (1.0 - erf(bet * (ev - lam_f)))


def fermi_level(eigvals, wkpt, bet, nelec, occfac=2.0, kind=1):
    """Root of  occfac * sum_k w_k sum_n occ(e_nk) = Nelectron.
    ``src/occupations.m:3-5,21-36`` (fzero there; bracketed Brent here)."""
    from scipy.optimize import brentq

    def f(lf):
        return sum(occfac * wk * np.sum(smearing_occ(ev, lf, bet, kind)) for ev, wk in zip(eigvals, wkpt)) - nelec

    lo = min(np.min(e) for e in eigvals) - 10.0 / bet - 1.0
    hi = max(np.max(e) for e in eigvals) + 10.0 / bet + 1.0
    return brentq(f, lo, hi, xtol=1e-15, rtol=4 * np.finfo(float).eps, maxiter=500)


def electron_density(psis, occs, wkpt):
    """rho = 2 sum_k w_k sum_n occ |psi|^2.  ``src/electronDensity.m:13-17``."""
    rho = 0
    for psi, occ, wk in zip(psis, occs, wkpt):
        rho = rho + 2 * wk * np.sum((psi * psi.conj()).real * occ[None, :], axis=1)
    return np.real(rho)


def anderson_wtd_avg(x_k, f_k, X, F):
    """``src/andersonWtdAvg.m:19-21``."""
    Gamma = np.linalg.pinv(F.T @ F) @ (F.T @ f_k)
    return x_k - X @ Gamma, f_k - F @ Gamma


class PeriodicPulay:
    """Potential mixing state.  ``src/mixing.m:20-116`` (MixingPrecond = 0,
    PulayFrequency = 1, PulayRestartFreq = 0, MixingVariable = 1)."""

    def __init__(self, N, hist=7, beta=0.3, omega=None, pfreq=1):
        self.m, self.beta = hist, beta
        self.omega = beta if omega is None else omega
        self.p = pfreq
        self.X = np.zeros((N, hist))
        self.F = np.zeros((N, hist))
        self.xkm1 = None
        self.fkm1 = np.zeros(N)

    def __call__(self, g_k, x_k, it):
        f_k = g_k - x_k
        if it > 1:
            i = (it - 2) % self.m
            self.X[:, i] = x_k - self.xkm1
            self.F[:, i] = f_k - self.fkm1
        if it % self.p == 0 and it > 1:
            x_w, f_w = anderson_wtd_avg(x_k, f_k, self.X, self.F)
        else:
            x_w, f_w = x_k, f_k
        x_kp1 = x_w + (self.beta if it % self.p == 0 else self.omega) * f_w
        self.fkm1 = f_k
        self.xkm1 = x_k
        return x_kp1


def band_energy(eigvals, occs, wkpt, occfac=2.0):
    """``src/evaluateTotalEnergy.m:12-20``."""
    return sum(occfac * wk * np.sum(ev * oc) for ev, oc, wk in zip(eigvals, occs, wkpt))


def entropy_energy(eigvals, occs, wkpt, lam_f, bet, kind=1, occfac=2.0):
    """``src/evaluateTotalEnergy.m:58-74``."""
    E = 0.0
    for ev, oc, wk in zip(eigvals, occs, wkpt):
        if kind == 0:
            with np.errstate(divide="ignore", invalid="ignore"):
                v = occfac * (1 / bet) * (oc * np.log(oc) + (1 - oc) * np.log(1 - oc))
            v[np.isnan(v)] = 0.0
        else:
            v = -occfac * (1 / bet) / (2 * np.sqrt(np.pi)) * np.exp(-(bet * (ev - lam_f)) ** 2)
        E += wk * np.sum(v)
    return E


def total_energy(eigvals, occs, wkpt, lam_f, bet, rho, vxc, phi, bcharge, W, Eself, Ecorr, kind=1, rhotol=1e-14):
    """``src/evaluateTotalEnergy.m:12-93`` for nspin = 1, LDA_PW.
    Returns (Etot, Eband, Exc, Exc_dc, Eelec_dc, Eent)."""
    Eband = band_energy(eigvals, occs, wkpt)
    rho_c = np.maximum(rho, rhotol)
    Exc = lda_pw_exc(rho, W, rhotol)
    Exc_dc = np.sum(vxc * rho_c * W)
    Eelec_dc = 0.5 * np.sum((bcharge - rho) * phi * W)
    Eent = entropy_energy(eigvals, occs, wkpt, lam_f, bet, kind)
    Etot = Eband + Exc - Exc_dc + Eelec_dc - Eself + Ecorr + Eent
    return Etot, Eband, Exc, Exc_dc, Eelec_dc, Eent


# --------------------------------------------------------------------------
# Sternheimer operator (DFPT right-hand sides)
# --------------------------------------------------------------------------
def sternheimer_apply(ham: Hamiltonian, lam, psi_v, X):
    """(H - lambda_n + psi_v psi_v^H W) x_n, one shift per column.
    ``src/phonon.m:2139-2157``."""
    HX = ham.apply(X)
    HX = HX - X * np.asarray(lam)[None, :]
    HX = HX + psi_v @ (psi_v.conj().T @ (ham.W[:, None] * X))
    return HX


def coeff_pv_psi(eigval, occ, bet, nv):
    """S.coeff_Pv_psi_kq(1:nv,1:nv,kpt) for q = 0 (EigVal_kq = EigVal, occ_kq =
    occ, S.Kdeltaq0 = 1), after the transpose.  ``src/phonon.m:116-158``."""
    c = np.zeros((nv, nv))
    for i in range(nv):
        for j in range(nv):
            temp = eigval[i] - eigval[j]
            temp_occ = occ[i] - occ[j]
            if j != i:
                if abs(temp) < 1e-4 and abs(temp_occ) < 1e-4:
                    c[i, j] = 1 - (1 - occ[i]) * bet / 2
                else:
                    c[i, j] = 1 / temp
            else:
                c[i, j] = 1.0
    return c.T.copy()


def sternheimer_rhs(W, kpt, dveff, atoms, datoms, psi, coeff):
    """Right-hand sides of one perturbation, ``src/phonon.m:1358-1444``
    ("Method-2"): PsiW, DVeff_psi, Chi_full and DChi per atom (DChi is the
    gradient of Chi_full with respect to the field point, i.e. minus the
    derivative with respect to the atom position held in ``datoms``),
    integral_1, integral_2, psi_DH_psi and rhs.  psi: N x nv.  Returns
    (rhs N x nv, dlambda = diag(psi_DH_psi))."""
    kpt = np.asarray(kpt, dtype=float)
    fac = 1.0j
    N, nv = psi.shape
    PsiW = (np.conj(psi) * W[:, None]).T
    DVeff_psi = dveff[:, None] * psi
    psi_DVeff_psi = PsiW @ DVeff_psi
    psi_DVnl_psi = np.zeros((nv, nv), dtype=complex)
    rhs_nl = np.zeros((N, nv), dtype=complex)
    for A, dA in zip(atoms, datoms):
        ang = A.gamma.size
        Chi_full = np.zeros((N, ang), dtype=complex)
        DChi = np.zeros((N, ang), dtype=complex)
        for im, dim in zip(A.images, dA.images):
            phase_fac = np.exp(-np.dot(kpt, (A.coords - im.coords) * fac))
            Chi_full[im.rc_pos, :] += im.chi * phase_fac
            DChi[im.rc_pos, :] += -dim.chi * phase_fac
        integral_1 = np.conj(PsiW) @ np.conj(Chi_full)
        integral_2 = -PsiW @ DChi
        temp = integral_2 @ (integral_1.T * A.gamma[:, None])
        psi_DVnl_psi += temp + temp.conj().T
        rhs_nl += -Chi_full @ (integral_2.conj().T * A.gamma[:, None]) + DChi @ (integral_1.T * A.gamma[:, None])
    psi_DH_psi = psi_DVeff_psi + psi_DVnl_psi
    rhs = psi @ (psi_DH_psi * coeff) + rhs_nl - DVeff_psi
    return rhs, np.real(np.diag(psi_DH_psi))


def occupation_pert(occs, wkpt, dlams, bet):
    """First-order Fermi level and occupations at q = 0,
    ``src/phonon.m:1254-1276``: with G = occ (1 - occ) over the occupied
    states of each channel, DlambdaF = sum_k w_k G.dlambda / sum_k w_k sum G
    and Dg = -bet G (dlambda - DlambdaF); both 0 when the denominator is
    below 1e-4 (an insulator).  occs, dlams: one array per channel (the
    occupied states).  Returns (DlambdaF, [Dg per channel])."""
    num = den = 0.0
    for oc, dl, wk in zip(occs, dlams, wkpt):
        G = oc * (1 - oc)
        num += wk * np.dot(G, dl)
        den += wk * np.sum(G)
    if den < 1e-4:
        return 0.0, [np.zeros_like(oc) for oc in occs]
    dlf = num / den
    return dlf, [-bet * oc * (1 - oc) * (dl - dlf) for oc, dl in zip(occs, dlams)]


def electron_density_pert(psis, dpsis, occs, dgs, wkpt):
    """First-order density at q = 0, ``src/phonon.m:1232-1240``:
    Drho = sum_k 2 w_k (sum_n Dg_n |psi_n|^2 + 2 Re sum_n occ_n Dpsi_n conj(psi_n))
    over the occupied states (psis[k], dpsis[k]: N x nv_k)."""
    drho = 0.0
    for psi, dpsi, oc, dg, wk in zip(psis, dpsis, occs, dgs, wkpt):
        drho = drho + 2 * wk * (np.sum((psi * psi.conj()).real * dg[None, :], axis=1)
                                + 2 * np.real(np.sum(dpsi * np.conj(psi) * oc[None, :], axis=1)))
    return np.real(drho)


def dfpt_scf(g: Grid, W, hams, atoms, datoms, psis, eigvals, occs, wkpt, rho, db, drho0, bet,
             SCF_tol=1e-5, poisson_tol=1e-8, lin_tol=1e-8, max_scf=100, min_scf=3, MINIT_SCF=3,
             hist=7, beta=0.3, omega=None, pfreq=1, lin_maxit=3000, poisson_maxit=1000, rhotol=1e-14,
             poisson=None, dpsi0=None, linsolve=None, poisson_precond="spectral", stern_precond="jacobi"):
    """Self-consistent first-order response to one atomic-displacement
    pattern at q = 0 with potential mixing, ``src/phonon.m:195-366``:

    * start: Drho = drho0 (the atomic-density guess of ``electrostatics_local``,
      ``:2405-2408``), Dphi = 0 (``:207``), DVxc = fxc Drho, DVeff = Dphi + DVxc
      (``:212-219``); the mixing history cleared, its x_{k-1} = DVeff (``:236-256``);
    * each iteration (``:259-359``): per channel the Sternheimer right-hand
      sides of DVeff and the projector changes and their bicgstab solves,
      warm-started from the previous solutions (``sternheimer`` ``:1279-1521``;
      ones at the very start, ``:224``); ``occupation_pert``;
      ``electronDensity_pert``; Dphi from AAR on f = -4 pi (Drho + Db) from the
      previous Dphi (``aar_phonon`` with the Laplacian, ``:286-293``);
      DVeff_out = Dphi + fxc Drho; err = ||DVeff_out - DVeff_in|| / ||DVeff_out||;
      Periodic-Pulay mixing unless converged (``:326-334``).

    hams: per channel the oracle Hamiltonian at the ground state (its veff);
    atoms, datoms: projector tables and their changes per channel (image
    phases come from each channel's kpt); psis: occupied states per channel
    (N x nv); rho: ground-state density (for fxc).  ``poisson(drho, db, dphi)``
    overrides the AAR solve (preconditioned by ``poisson_precond``: "spectral",
    the exact Lap_std^-1 that stands in for the reference's ILU factors, or
    "jacobi") and ``linsolve(k, lam, psi, rhs, x0)`` the
    Sternheimer bicgstab of channel k.  Returns a dict with Drho, Dphi, Dveff
    (the last output potential), Dlambda and Dpsi per channel, DlambdaF,
    count and the error history."""
    fxc = lda_pw_fxc(rho, rhotol)
    nch = len(hams)
    coeffs = [coeff_pv_psi(ev, oc, bet, psi.shape[1]) for ev, oc, psi in zip(eigvals, occs, psis)]
    if poisson is None:
        def poisson(drho, dbb, guess):
            x, _, _ = poisson_solve(g, drho, dbb, poisson_tol, guess, poisson_maxit, poisson_precond)
            return x
    drho = np.array(drho0, dtype=float, copy=True)
    dphi = np.zeros(g.N)
    dveff_in = dphi + fxc * drho
    mix = PeriodicPulay(g.N, hist, beta, omega, pfreq)
    dpsis = ([np.ones(p.shape, dtype=complex) for p in psis] if dpsi0 is None
             else [np.array(d, copy=True) for d in dpsi0])
    err, count = 100.0, 1
    errs = []
    dlams = [None] * nch
    dlf = 0.0
    dveff_out = dveff_in
    while (err > SCF_tol and count <= max_scf) or count <= min_scf:
        for k in range(nch):
            psi = psis[k].astype(complex)
            rhs, dlams[k] = sternheimer_rhs(W, hams[k].kpt, dveff_in, atoms[k], datoms[k], psi, coeffs[k])
            lam = eigvals[k][:psi.shape[1]]
            if linsolve is None:
                dpsis[k], _, _, _ = sternheimer_solve(hams[k], lam, psi, rhs, dpsis[k], lin_tol, lin_maxit,
                                                      stern_precond)
            else:
                dpsis[k] = linsolve(k, lam, psi, rhs, dpsis[k])
        dlf, dgs = occupation_pert([oc[:p.shape[1]] for oc, p in zip(occs, psis)], wkpt, dlams, bet)
        drho = electron_density_pert(psis, dpsis, [oc[:p.shape[1]] for oc, p in zip(occs, psis)], dgs, wkpt)
        dphi = poisson(drho, db, dphi)
        dveff_out = dphi + fxc * drho
        err = np.linalg.norm(dveff_out - dveff_in) / np.linalg.norm(dveff_out)
        errs.append(err)
        if err > SCF_tol or count < MINIT_SCF:
            dveff_in = mix(dveff_out, dveff_in, count)
        else:
            dveff_in = dveff_out
        count += 1
    return dict(Drho=drho, Dphi=dphi, Dveff=dveff_out, Dlambda=dlams, Dpsi=dpsis, DlambdaF=dlf,
                count=count - 1, errors=errs, converged=err <= SCF_tol)


# --------------------------------------------------------------------------
# SCF loop
# --------------------------------------------------------------------------
def scf(g: Grid, W, atoms, channels, wkpt, bcharge, rho0, Eself, Ecorr, nelec, n_atom, psi0, npl, SCF_tol,
        poisson_tol, maxit_scf=100, minit_scf=3, rho_trigger=4, hist=7, beta=0.3, omega=None, pfreq=1,
        bet=27.211386245988 / 0.2, kind=1, tol_lanczos=1e-2, lanczos_maxit=300, maxit_poisson=1000,
        rhotol=1e-14, poisson_precond="spectral"):
    """``src/scf.m:32-306`` for nspin = 1, potential mixing, BC = 3 (no mean
    shift of Veff) and the first relaxation step.  psi0: list of (N, Nev)
    blocks, one per channel.  The Poisson solves are AAR preconditioned by
    ``poisson_precond`` ("spectral": the exact Lap_std^-1, in the place of the
    reference's ILU factors; or "jacobi").  Returns a dict with the converged
    energies, eigenvalues and the per-iteration record."""
    phi, _, _ = poisson_solve(g, rho0, bcharge, poisson_tol, None, maxit_poisson, poisson_precond)
    vxc = lda_pw_vxc(rho0, rhotol)
    veff = phi + vxc                                        # src/scf.m:39
    mix = PeriodicPulay(g.N, hist, beta, omega, pfreq)
    veff_in = veff.copy()
    psis = [np.array(p, copy=True) for p in psi0]
    nch = len(channels)
    Nev = psis[0].shape[1]
    EigVal = np.zeros((Nev, nch))
    bup, a0, cut = np.zeros(nch), np.zeros(nch), np.zeros(nch)
    Wm = metric_weights(g)
    count, count_scf, err = 1, 1, 100.0
    history = []
    Etot, lam_f, occs = 0.0, 0.0, None
    min_scf = min(minit_scf, maxit_scf)
    while (err > SCF_tol and count_scf <= maxit_scf) or count_scf <= min_scf:
        for ks, kpt in enumerate(channels):                 # src/eigSolver.m:26-124
            ham = Hamiltonian(g, tuple(kpt), veff, W, atoms)
            if count == 1:
                v0 = np.random.default_rng(ks + 1).random(g.N)
                lo, hi = lanczos_bounds(ham.apply, v0, lanczos_maxit, tol_lanczos, Wm)
                bup[ks] = 1.01 * hi
                a0[ks] = lo - 0.1
                cut[ks] = 0.5 * (bup[ks] + a0[ks])
            else:
                if count > rho_trigger:
                    a0[ks] = np.min(EigVal[:, ks])
                cut[ks] = np.max(EigVal[:, ks]) + 0.10
            ev, psis[ks] = eig_solver_pass(ham, psis[ks], npl, cut[ks], bup[ks], a0[ks])
            EigVal[:, ks] = ev
        evs = [EigVal[:, ks] for ks in range(nch)]
        lam_f = fermi_level(evs, wkpt, bet, nelec, 2.0, kind)  # src/occupations.m
        occs = [smearing_occ(e, lam_f, bet, kind) for e in evs]
        if count >= rho_trigger:
            rho = electron_density(psis, occs, wkpt)
            phi, _, it_p = poisson_solve(g, rho, bcharge, poisson_tol, phi, maxit_poisson, poisson_precond)
            vxc = lda_pw_vxc(rho, rhotol)
            veff_out = phi + vxc
            Etot, Eband, Exc, Exc_dc, Eelec_dc, Eent = total_energy(evs, occs, wkpt, lam_f, bet, rho, vxc, phi,
                                                                   bcharge, W, Eself, Ecorr, kind, rhotol)
            Escc = float(np.sum((veff_out - veff_in) * rho * W))     # src/scf.m:160-162
            Etot += Escc
            err = np.linalg.norm(veff_out - veff_in) / np.linalg.norm(veff_out)   # src/scf.m:184
            history.append(dict(count_SCF=count_scf, Etotal=Etot, Eband=Eband, Exc=Exc, Exc_dc=Exc_dc,
                                Eelec_dc=Eelec_dc, Eent=Eent, Escc=Escc, lambda_f=lam_f, err=err,
                                aar_iterations=it_p))
            if err > SCF_tol or count_scf < minit_scf:
                veff = mix(veff_out, veff_in, count_scf)    # src/scf.m:193-209
                veff_in = veff.copy()
            else:
                veff = veff_out
            count_scf += 1
        count += 1
    converged = not (count_scf == maxit_scf + 1 and err > SCF_tol)
    # veff_in: the potential the last eigenpairs belong to; vxc: Vxc of rho
    return dict(converged=converged, count_SCF=count_scf - 1, Etotal=Etot, Eatom=Etot / n_atom, lambda_f=lam_f,
                EigVal=EigVal, occ=np.stack(occs, 1), history=history, psi=psis, rho=rho, phi=phi, veff=veff,
                veff_in=veff_in, vxc=vxc)


def bicgstab(A, b, x0, tol, maxit, Minv=None):
    """BiCGSTAB (van der Vorst, SIAM J. Sci. Stat. Comput. 13 (1992) 631),
    right-preconditioned, with the half-step exit of MATLAB's bicgstab
    (src/phonon.m:1519 calls it).  Stops when ||b - A x|| <= tol ||b|| by the
    recursive residual.  Returns (x, relres, iterations, flag): flag 0
    converged, 1 iteration limit, 4 breakdown."""
    M = (lambda z: z) if Minv is None else Minv
    x = np.zeros_like(b) if x0 is None else np.array(x0, copy=True)
    bn = np.linalg.norm(b)
    if bn == 0:
        return np.zeros_like(b), 0.0, 0, 0
    r = b - A(x)
    rt = r.copy()
    if np.linalg.norm(r) <= tol * bn:
        return x, np.linalg.norm(r) / bn, 0, 0
    rho_prev = alpha = omega = 1.0
    p = np.zeros_like(b)
    v = np.zeros_like(b)
    for it in range(1, maxit + 1):
        rho = np.vdot(rt, r)
        if rho == 0:
            return x, np.linalg.norm(r) / bn, it, 4
        if it == 1:
            p = r.copy()
        else:
            beta = (rho / rho_prev) * (alpha / omega)
            p = r + beta * (p - omega * v)
        ph = M(p)
        v = A(ph)
        rtv = np.vdot(rt, v)
        if rtv == 0:
            return x, np.linalg.norm(r) / bn, it, 4
        alpha = rho / rtv
        s = r - alpha * v
        if np.linalg.norm(s) <= tol * bn:
            return x + alpha * ph, np.linalg.norm(s) / bn, it, 0
        sh = M(s)
        t = A(sh)
        tt = np.vdot(t, t)
        if tt == 0:
            return x + alpha * ph, np.linalg.norm(s) / bn, it, 4
        omega = np.vdot(t, s) / tt
        x = x + alpha * ph + omega * sh
        r = s - omega * t
        if np.linalg.norm(r) <= tol * bn:
            return x, np.linalg.norm(r) / bn, it, 0
        if omega == 0:
            return x, np.linalg.norm(r) / bn, it, 4
        rho_prev = rho
    return x, np.linalg.norm(r) / bn, maxit, 1


def jacobi_lap_half(g: Grid):
    """Per-point inverse of the diagonal of -1/2 Lap (the Jacobi stand-in for
    the ILU factors of S.Lap_std that src/phonon.m:1519 passes)."""
    d = -0.5 * lap_matrix(g).diagonal()
    return 1.0 / d


def sternheimer_solve(ham: Hamiltonian, lam, psi_v, B, X0=None, tol=1e-8, maxit=3000, precond=True):
    """Column by column bicgstab on (H - lam_n + psi_v psi_v^H W) x_n = b_n
    (src/phonon.m:1516-1521).  precond: "spectral" (the exact Lap_std^-1, in
    the place of the reference's ILU factors of Lap_std), True / "jacobi"
    (the diagonal of -1/2 Lap), False / "none".  Returns (X, relres[],
    iterations[], flags[])."""
    if precond == "spectral":
        Ainv = spectral_inverses(ham.g)
        Minv = lambda z: lap_spectral_solve(ham.g, z, Ainv=Ainv)
    elif precond is True or precond == "jacobi":
        dinv = jacobi_lap_half(ham.g)
        Minv = lambda z: dinv * z
    else:
        Minv = None
    ncol = B.shape[1]
    X = np.zeros_like(B)
    rel, its, flags = np.zeros(ncol), np.zeros(ncol, dtype=int), np.zeros(ncol, dtype=int)
    for j in range(ncol):
        A = lambda z, j=j: sternheimer_apply(ham, [lam[j]], psi_v, z[:, None])[:, 0]
        x0 = None if X0 is None else X0[:, j]
        X[:, j], rel[j], its[j], flags[j] = bicgstab(A, B[:, j], x0, tol, maxit, Minv)
    return X, rel, its, flags
