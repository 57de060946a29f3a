"""Host-side set-up of one Cyclix sector: grid, quadrature weights, atoms,
synthetic Gaussian pseudopotential, KB projector tables and pseudocharge.

This is the one-off geometry work that the reference does in MATLAB before
the SCF loop (``src/initialization.m``, ``src/calculate_nloc_projector.m``,
``src/calculate_b_guessRho_Eself.m``).  It produces exactly the arrays the
reference keeps in ``S`` (W, Atom(J).rcImage(img).{rc_pos,Chi_mat},
gamma_Jl, b, b_ref, Eself, E_corr, rho_at) so that the device library can be
handed the same data a MATLAB host would hand it.

Synthetic pseudopotential (one species):

* local part: potential of a Gaussian ion charge of width ``sigma_loc``,
  V_loc(d) = -Z erf(d / (sqrt(2) sigma_loc)) / d;
* one s-channel KB projector with Gaussian radial form,
  chi(d) = exp(-d^2 / (2 sigma_nl^2)) * Y00, truncated at ``rc_nl``, energy
  gamma = ``gamma_nl`` Ha;
* isolated-atom density guess: Gaussian of width ``sigma_rho`` holding Z
  electrons.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy.special import erf

Y00 = 0.282094791773878  # 0.5*sqrt(1/pi), src/sphericalHarmonics.m:64-67
BOHR_PER_ANG = 1.0 / 0.529177210903


def fd_weights(fdn: int):
    """Order-2*fdn central weights (w2 second, w1 first derivative).
    ``src/initialization.m:403-419``."""
    w2 = np.zeros(fdn + 1)
    w1 = np.zeros(fdn + 1)
    f = math.factorial
    for k in range(1, fdn + 1):
        c = f(fdn) ** 2 / (f(fdn - k) * f(fdn + k))
        w2[k] = 2 * (-1) ** (k + 1) * c / (k * k)
        w1[k] = (-1) ** (k + 1) * c / k
        w2[0] -= 2.0 / (k * k)
    return w2, w1


@dataclass
class GridSpec:
    """Node counts and spacings of one sector (S.Nx/Ny/Nz, S.dx/dy/dz,
    S.xin, S.twist, S.FDn, S.cell_typ)."""

    nr: int
    nt: int
    nz: int
    xin: float
    dr: float
    dt: float
    dz: float
    twist: float = 0.0
    fdn: int = 6
    cell_typ: int = 3  # 3 cyclic, 5 helical

    @property
    def N(self):
        return self.nr * self.nt * self.nz

    @property
    def L2(self):
        return self.nt * self.dt

    @property
    def L3(self):
        return self.nz * self.dz

    @property
    def r(self):
        return self.xin + np.arange(self.nr) * self.dr

    @property
    def helical(self):
        return self.cell_typ in (4, 5)

    def weights(self):
        """W = kron(W_z, kron(W_t, W_r)).  ``Cychel/intgWts_cychel.m:2-14``."""
        wx = self.r * self.dr
        wx[0] = 0.5 * self.xin * self.dr + 0.125 * self.dr ** 2
        wx[-1] = 0.5 * (self.xin + (self.nr - 1) * self.dr) * self.dr - 0.125 * self.dr ** 2
        return np.kron(np.full(self.nz, self.dz), np.kron(np.full(self.nt, self.dt), wx))


def _matlab_round(x):
    """MATLAB round: halves away from zero (numpy rounds halves to even)."""
    x = np.asarray(x, dtype=float)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def generate_kpts(nkpt, L1, L2, L3, kptshift=(0.0, 0.0, 0.0), cell_typ=3, time_rev=True):
    """Bloch channels of a cyclic / helical sector and their weights:
    ``Generate_kpts`` (``src/initialization.m:1591-1682``, the cell_typ 3-5
    branch).  The cyclic channel is the integer nu = round(2 pi/L2 * j/n2),
    j = 0..n2-1 (MATLAB rounding); the axial one eta = 2 pi/L3 * k with k on
    the Monkhorst-Pack grid of n3 points (shifted, mapped into (-1/2, 1/2]).
    Time-reversal symmetry drops channel i when a later channel j has
    nu_i + nu_j in {0, 2 pi/L2} and eta_i + eta_j = 0, and doubles j's
    weight.  Returns (kptgrid (n, 3) rows [0, nu, eta], weights (n,))."""
    n1, n2, n3 = (int(v) for v in nkpt)
    if n1 != 1:
        raise ValueError("nkpt cannot be greater than 1 in the Dirichlet radial direction")
    eps = np.finfo(float).eps
    mp1 = lambda n: np.arange(-((n - 1) // 2), -((n - 1) // 2) + n)   # noqa: E731  (MP grid, infinite order)
    kx = mp1(n1) / n1
    ky = np.arange(n2) / n2                                              # finite group order
    kz = mp1(n3) / n3 + kptshift[2] / n3
    kz = np.mod(kz + 0.5 - eps, 1.0) - 0.5 + eps
    kx = (2 * math.pi / L1) * kx
    ky = _matlab_round((2 * math.pi / L2) * ky)
    kz = (2 * math.pi / L3) * kz
    # ndgrid order: x fastest, then y, then z
    grid = np.array([[x, y, z] for z in kz for y in ky for x in kx], dtype=float)
    n = grid.shape[0]
    w = np.full(n, 1.0 / n)
    if time_rev:
        sumy = 2 * math.pi / L2
        tol = 1e-8
        drop = np.zeros(n, dtype=bool)
        dbl = np.zeros(n, dtype=bool)
        for i in range(n):
            for j in range(i + 1, n):
                if (abs(grid[i, 0] + grid[j, 0]) < tol
                        and (abs(grid[i, 1] + grid[j, 1] - sumy) < tol or abs(grid[i, 1] + grid[j, 1]) < tol)
                        and abs(grid[i, 2] + grid[j, 2]) < tol):
                    drop[i] = True
                    dbl[j] = True
        w[dbl] *= 2
        grid, w = grid[~drop], w[~drop]
    return grid, w


def to_cart(g: GridSpec, pts):
    """(r, theta, z) -> Cartesian.  ``Cychel/coordinateTransformation_cychel.m:3-7``."""
    pts = np.atleast_2d(pts)
    ang = pts[:, 1] + g.twist * pts[:, 2]
    return np.stack([pts[:, 0] * np.cos(ang), pts[:, 0] * np.sin(ang), pts[:, 2]], axis=1)


@dataclass
class Species:
    Z: float = 4.0
    sigma_loc: float = 0.6
    sigma_nl: float = 0.8
    gamma_nl: float = 2.0
    rc_nl: float = 3.0
    sigma_rho: float = 1.0
    rb: float = 4.2  # pseudocharge support radius (about 7 sigma_loc)

    def vloc(self, d):
        d = np.asarray(d, dtype=float)
        out = np.empty_like(d)
        small = d < 1e-12
        out[small] = -self.Z * math.sqrt(2 / math.pi) / self.sigma_loc
        dd = d[~small]
        out[~small] = -self.Z * erf(dd / (math.sqrt(2) * self.sigma_loc)) / dd
        return out

    def dvloc_over_d(self, d):
        """V_loc'(d) / d, finite at d = 0 (the radial derivative of the local
        potential divided by the distance, for the displacement derivative
        dV/dX = -V'(d)/d (x - X))."""
        d = np.asarray(d, dtype=float)
        s = math.sqrt(2) * self.sigma_loc
        x = d / s
        out = np.empty_like(d)
        small = x < 0.05
        # series of erf(x)/x: V'/d = (Z/s^3)(2/sqrt(pi))(2/3 - 2x^2/5 + x^4/7 - x^6/27)
        xs = x[small] ** 2
        out[small] = (self.Z / s ** 3) * (2 / math.sqrt(math.pi)) * (2 / 3 - xs * (2 / 5 - xs * (1 / 7 - xs / 27)))
        dd = d[~small]
        g = math.sqrt(2 / math.pi) / self.sigma_loc * np.exp(-dd ** 2 / (2 * self.sigma_loc ** 2))
        out[~small] = -self.Z * (g / dd ** 2 - erf(dd / s) / dd ** 3)
        return out

    def udv(self, d):
        return np.exp(-np.asarray(d) ** 2 / (2 * self.sigma_nl ** 2))

    def dudv_over_d(self, d):
        """UdV'(d) / d of the Gaussian projector radial form."""
        return -self.udv(d) / self.sigma_nl ** 2

    def rho_iso(self, d):
        s = self.sigma_rho
        return self.Z / (2 * math.pi * s * s) ** 1.5 * np.exp(-np.asarray(d) ** 2 / (2 * s * s))

    def drho_iso_over_d(self, d):
        """rho_iso'(d) / d of the Gaussian atomic density."""
        return -self.rho_iso(d) / self.sigma_rho ** 2

    @property
    def rc(self):
        """Projector ball radius (S.Atm(ityp).rc)."""
        return self.rc_nl

    def projector_channels(self):
        """(l, UdV radial form, gamma) per projector, in the reference's
        (l, pp) order (``src/calculate_nloc_projector.m:211-236``)."""
        return [(0, self.udv, self.gamma_nl)]


def real_ylm(l, m, X, Y, Z):
    """Real spherical harmonics of ``src/sphericalHarmonics.m`` (l <= 2) on
    the vectors (X, Y, Z).  At the origin the l >= 1 functions are set to 0,
    their limit along any ray times the vanishing radial factor of an l >= 1
    projector (the reference's formula divides 0 by 0 there)."""
    if l == 0:
        return np.full(np.shape(X), Y00)
    r = np.sqrt(X * X + Y * Y + Z * Z)
    rs = np.where(r > 0, r, 1.0)
    if l == 1:
        c = 0.488602511902920  # sqrt(3/(4 pi))
        v = {-1: Y, 0: Z, 1: X}[m] / rs
        return np.where(r > 0, c * v, 0.0)
    if l == 2:
        r2 = rs * rs
        v = {-2: 1.092548430592079 * X * Y, -1: 1.092548430592079 * Y * Z,
             0: 0.315391565252520 * (-X * X - Y * Y + 2 * Z * Z), 1: 1.092548430592079 * Z * X,
             2: 0.546274215296040 * (X * X - Y * Y)}[m] / r2
        return np.where(r > 0, v, 0.0)
    raise ValueError("l > 2 is not supported")


@dataclass
class ProjImage:
    rc_pos: np.ndarray  # int32, 0-based node index
    chi: np.ndarray  # (n, angnum)
    coords: np.ndarray  # (3,)


@dataclass
class ProjAtom:
    coords: np.ndarray
    gamma: np.ndarray
    images: list = field(default_factory=list)


@dataclass
class CyclixSystem:
    grid: GridSpec
    atoms: np.ndarray  # (n_atm, 3) in (r, theta, z)
    species: Species
    nstates: int
    npl: int
    kpt: tuple = (0.0, 0.0, 0.0)  # (0, nu, eta)
    name: str = ""
    W: np.ndarray = None

    def __post_init__(self):
        if self.W is None:
            self.W = self.grid.weights()

    @property
    def n_atm(self):
        return self.atoms.shape[0]

    @property
    def nelectron(self):
        return self.species.Z * self.n_atm

    # ------------------------------------------------------------------
    def rb_xyz(self):
        """Per-direction support of the pseudocharge / projector boxes.
        ``src/initialization.m:55-61,104-124`` (rb_y is an angle).  A species
        that carries its own radii (found by the reference's search,
        ``deck.find_rb``) supplies them as ``rb_xyz``."""
        own = getattr(self.species, "rb_xyz", None)
        if own is not None:
            return tuple(own)
        g = self.grid
        R = self.atoms[:, 0].max()
        rb = self.species.rb
        rbx = math.ceil(rb / g.dr - 1e-12) * g.dr
        rbz = math.ceil(rb / g.dz - 1e-12) * g.dz
        rby = math.acos(max(-1.0, 1 - rbz ** 2 / (2 * R * R))) + g.twist * rbz * 0.5
        rby = math.ceil(rby / g.dt - 1e-12) * g.dt
        return rbx, rby, rbz

    def image_rotation(self, J, im):
        """Cartesian rotation taking atom J's displacement to that of its image
        at ``im`` (r, theta, z): the image angle offset (theta_img - theta_J) +
        twist (z_img - z_J) about the axis.  The reference's
        ``S.RotM1^fac1 * S.RotM2^fac2`` (``src/phonon.m:1381-1386``)."""
        a = (im[1] - self.atoms[J, 1]) + self.grid.twist * (im[2] - self.atoms[J, 2])
        c, s_ = math.cos(a), math.sin(a)
        return np.array([[c, -s_, 0.0], [s_, c, 0.0], [0.0, 0.0, 1.0]])

    def _images(self, x0):
        """Periodic images in theta and z.  ``src/calculate_nloc_projector.m:50-81``."""
        g = self.grid
        rbx, rby, rbz = self.rb_xyz()
        ny_l = math.floor((x0[1] + rby) / g.L2)
        ny_r = math.floor((g.L2 - x0[1] + rby - g.dt) / g.L2)
        nz_l = math.floor((x0[2] + rbz) / g.L3)
        nz_r = math.floor((g.L3 - x0[2] + rbz - g.dz) / g.L3)
        out = []
        for kz in range(-nz_l, nz_r + 1):
            for ky in range(-ny_l, ny_r + 1):
                out.append(np.array([x0[0], x0[1] + ky * g.L2, x0[2] + kz * g.L3]))
        return out

    def projectors(self):
        """KB projector tables.  ``src/calculate_nloc_projector.m:23-287``
        (psptyp 1: every (l, pp) projector of the species times the real
        Y_lm, m = -l..l, evaluated on the node-minus-image vectors rotated
        back by the image's angle, as ``ROT = S.RotM1^fac1 * S.RotM2^fac2``
        does at ``:178-184``)."""
        g = self.grid
        sp = self.species
        rbx, rby, rbz = self.rb_xyz()
        chans = sp.projector_channels()
        gam = np.array([gm for (l, _, gm) in chans for _m in range(2 * l + 1)], dtype=float)
        s_only = all(l == 0 for (l, _, _) in chans)
        out = []
        for J in range(self.n_atm):
            x0 = self.atoms[J]
            A = ProjAtom(coords=x0.copy(), gamma=gam.copy())
            for im in self._images(x0):
                x0i, y0i, z0i = im
                inx = ((x0i - g.xin) - (g.nr - 1) * g.dr + g.dr <= rbx) or abs(x0i - g.xin) <= rbx
                iny = (y0i - g.L2 + g.dt <= rby) or abs(y0i) <= rby
                inz = (z0i - g.L3 + g.dz <= rbz) or abs(z0i) <= rbz
                if not (inx and iny and inz):
                    continue
                i_s = max(math.ceil(((x0i - g.xin) - rbx) / g.dr) + 1, 1)
                i_e = min(math.floor(((x0i - g.xin) + rbx) / g.dr) + 1, g.nr)
                j_s = max(math.ceil((y0i - rby) / g.dt) + 1, 1)
                j_e = min(math.floor((y0i + rby) / g.dt) + 1, g.nt)
                k_s = max(math.ceil((z0i - rbz) / g.dz) + 1, 1)
                k_e = min(math.floor((z0i + rbz) / g.dz) + 1, g.nz)
                if i_e < i_s or j_e < j_s or k_e < k_s:
                    continue
                ii, jj, kk = np.meshgrid(np.arange(i_s, i_e + 1), np.arange(j_s, j_e + 1),
                                         np.arange(k_s, k_e + 1), indexing="ij")
                ii, jj, kk = (a.ravel(order="F") for a in (ii, jj, kk))
                pts = np.stack([g.xin + (ii - 1) * g.dr, (jj - 1) * g.dt, (kk - 1) * g.dz], axis=1)
                rel = to_cart(g, pts) - to_cart(g, im)
                d = np.linalg.norm(rel, axis=1)
                ball = d <= sp.rc
                if not np.any(ball):
                    continue
                rc_pos = ((kk[ball] - 1) * g.nr * g.nt + (jj[ball] - 1) * g.nr + ii[ball] - 1).astype(np.int32)
                if s_only:
                    chi = np.stack([radial(d[ball]) * Y00 for (_, radial, _) in chans], axis=1)
                else:
                    P = rel[ball] @ self.image_rotation(J, im)  # rows R^T v: the reference's ROT v
                    cols = []
                    for (l, radial, _) in chans:
                        u = radial(d[ball])
                        for m in range(-l, l + 1):
                            cols.append(u * real_ylm(l, m, P[:, 0], P[:, 1], P[:, 2]))
                    chi = np.stack(cols, axis=1)
                A.images.append(ProjImage(rc_pos=rc_pos, chi=chi, coords=np.array(im, dtype=float)))
            out.append(A)
        return out

    def projector_derivatives(self, u, proj=None):
        """First-order change of every projector image under the Cartesian
        atom displacements u (n_atm x 3): dchi = grad_X chi . (R_img u_J), on
        the node set of ``projectors()`` (exact derivative of the discrete
        tables, whose node sets do not move for small displacements).  The
        reference forms the same term from the finite-difference gradient of
        Chi_full (``src/phonon.m:1365-1399``, DChi = -dchi); here the radial
        form is differentiated in closed form.  Returns a list per atom of
        arrays per image (n_nodes, angnum)."""
        g = self.grid
        sp = self.species
        proj = self.projectors() if proj is None else proj
        u = np.asarray(u, dtype=float).reshape(self.n_atm, 3)
        out = []
        for J, A in enumerate(proj):
            per = []
            for im in A.images:
                q = im.rc_pos.astype(np.int64)
                ii, jj, kk = q % g.nr, (q // g.nr) % g.nt, q // (g.nr * g.nt)
                pts = np.stack([g.xin + ii * g.dr, jj * g.dt, kk * g.dz], axis=1)
                diff = to_cart(g, pts) - to_cart(g, im.coords)
                d = np.linalg.norm(diff, axis=1)
                v = self.image_rotation(J, im.coords) @ u[J]
                # d chi / d X_img = -UdV'(d)/d (x - X_img) Y00
                dchi = -(sp.dudv_over_d(d) * Y00) * (diff @ v)
                per.append(dchi[:, None] * np.ones((1, im.chi.shape[1])))
            out.append(per)
        return out

    # ------------------------------------------------------------------
    def _box_laplacian(self, V, xin_box):
        """Direct stencil of ``Cychel/pseudochargeDensity_atom_cychel.m:1-33``."""
        g = self.grid
        w2, w1 = fd_weights(g.fdn)
        n0 = g.fdn
        c1, c2, c3 = 1 / g.dr ** 2, 1 / g.dt ** 2, 1 / g.dz ** 2
        c4 = g.twist ** 2 / g.dt ** 2
        c5 = -(2 * g.twist) / (g.dt * g.dz)
        c6 = 1 / g.dr
        nx, ny, nz = V.shape
        rr = (xin_box + np.arange(n0, nx - n0) * g.dr)[:, None, None]

        def sh(a, b, c):
            return V[n0 + a:nx - n0 + a, n0 + b:ny - n0 + b, n0 + c:nz - n0 + c]

        out = w2[0] * (c1 + c2 / rr ** 2 + c3 + c4) * sh(0, 0, 0)
        for p in range(1, n0 + 1):
            out = out + (c1 * w2[p] + c6 * w1[p] / rr) * sh(p, 0, 0) \
                + (c1 * w2[p] - c6 * w1[p] / rr) * sh(-p, 0, 0) \
                + (c2 / rr ** 2 + c4) * w2[p] * (sh(0, p, 0) + sh(0, -p, 0)) \
                + c3 * w2[p] * (sh(0, 0, p) + sh(0, 0, -p))
            for q in range(1, n0 + 1):
                cf = c5 * w1[p] * w1[q]
                out = out + cf * (sh(0, p, q) - sh(0, p, -q) - sh(0, -p, q) + sh(0, -p, -q))
        return out

    def pseudocharge(self, rc_ref=0.5):
        """b, b_ref, Eself, E_corr, rho_at.  ``src/calculate_b_guessRho_Eself.m:29-216``."""
        g = self.grid
        sp = self.species
        W = self.W
        N = g.N
        b = np.zeros(N)
        b_ref = np.zeros(N)
        rho_at = np.zeros(N)
        V_c = np.zeros(N)
        Eself = 0.0
        Eself_ref = 0.0
        rbx, rby, rbz = self.rb_xyz()
        n0 = g.fdn
        for J in range(self.n_atm):
            for im in self._images(self.atoms[J]):
                x0i, y0i, z0i = im
                i_s = max(math.ceil(((x0i - g.xin) - rbx) / g.dr) + 1, 1)
                i_e = min(math.floor(((x0i - g.xin) + rbx) / g.dr) + 1, g.nr)
                j_s = max(math.ceil((y0i - rby) / g.dt) + 1, 1)
                j_e = min(math.floor((y0i + rby) / g.dt) + 1, g.nt)
                k_s = max(math.ceil((z0i - rbz) / g.dz) + 1, 1)
                k_e = min(math.floor((z0i + rbz) / g.dz) + 1, g.nz)
                if i_e < i_s or j_e < j_s or k_e < k_s:
                    continue
                xx = g.xin + np.arange(i_s - n0 - 1, i_e + n0) * g.dr
                yy = np.arange(j_s - n0 - 1, j_e + n0) * g.dt
                zz = np.arange(k_s - n0 - 1, k_e + n0) * g.dz
                XX, YY, ZZ = np.meshgrid(xx, yy, zz, indexing="ij")
                pts = np.stack([XX.ravel(), YY.ravel(), ZZ.ravel()], axis=1)
                dd = np.linalg.norm(to_cart(g, pts) - to_cart(g, im), axis=1).reshape(XX.shape)
                V_PS = sp.vloc(dd)
                V_ref = np.where(dd < rc_ref,
                                 -sp.Z * (9 * dd ** 7 - 30 * rc_ref * dd ** 6 + 28 * rc_ref ** 2 * dd ** 5
                                          - 14 * rc_ref ** 5 * dd ** 2 + 12 * rc_ref ** 7) / (5 * rc_ref ** 8),
                                 -sp.Z / np.maximum(dd, 1e-300))
                rho_iso = sp.rho_iso(dd)
                bJ = self._box_laplacian(V_PS, xx[0])
                bJr = self._box_laplacian(V_ref, xx[0])
                core = (slice(n0, -n0),) * 3
                IIb, JJb, KKb = np.meshgrid(np.arange(i_s, i_e + 1), np.arange(j_s, j_e + 1),
                                            np.arange(k_s, k_e + 1), indexing="ij")
                rows = ((KKb - 1) * g.nr * g.nt + (JJb - 1) * g.nr + IIb - 1).ravel()
                np.add.at(b, rows, bJ.ravel())
                np.add.at(b_ref, rows, bJr.ravel())
                np.add.at(rho_at, rows, rho_iso[core].ravel())
                Eself += 0.5 * np.sum(bJ.ravel() * V_PS[core].ravel() * W[rows])
                Eself_ref += 0.5 * np.sum(bJr.ravel() * V_ref[core].ravel() * W[rows])
                np.add.at(V_c, rows, (V_ref[core] - V_PS[core]).ravel())
        b *= -1 / (4 * math.pi)
        b_ref *= -1 / (4 * math.pi)
        Eself *= -1 / (4 * math.pi)
        Eself_ref *= -1 / (4 * math.pi)
        pos = abs(np.dot(W, b))
        neg = -pos
        rho_scale = abs(neg / np.dot(W, rho_at))
        rho_at *= rho_scale
        E_corr = 0.5 * np.sum((b_ref + b) * V_c * W) + Eself - Eself_ref
        return dict(b=b, b_ref=b_ref, Eself=Eself, E_corr=E_corr, rho_at=rho_at, V_c=V_c,
                    PosCharge=pos, NegCharge=neg, rho_scale=rho_scale)

    def pseudocharge_derivative(self, u):
        """First-order change of the pseudocharge b under the Cartesian atom
        displacements u (n_atm x 3): the box Laplacian of
        dV_J = grad_X V_J . (R_img u_J) on the boxes of ``pseudocharge()``
        (the reference's S.Db_dR of ``electrostatics_local``,
        ``src/phonon.m:2164-2330``, for one displacement pattern)."""
        g = self.grid
        sp = self.species
        u = np.asarray(u, dtype=float).reshape(self.n_atm, 3)
        db = np.zeros(g.N)
        rbx, rby, rbz = self.rb_xyz()
        n0 = g.fdn
        for J in range(self.n_atm):
            for im in self._images(self.atoms[J]):
                x0i, y0i, z0i = im
                i_s = max(math.ceil(((x0i - g.xin) - rbx) / g.dr) + 1, 1)
                i_e = min(math.floor(((x0i - g.xin) + rbx) / g.dr) + 1, g.nr)
                j_s = max(math.ceil((y0i - rby) / g.dt) + 1, 1)
                j_e = min(math.floor((y0i + rby) / g.dt) + 1, g.nt)
                k_s = max(math.ceil((z0i - rbz) / g.dz) + 1, 1)
                k_e = min(math.floor((z0i + rbz) / g.dz) + 1, g.nz)
                if i_e < i_s or j_e < j_s or k_e < k_s:
                    continue
                xx = g.xin + np.arange(i_s - n0 - 1, i_e + n0) * g.dr
                yy = np.arange(j_s - n0 - 1, j_e + n0) * g.dt
                zz = np.arange(k_s - n0 - 1, k_e + n0) * g.dz
                XX, YY, ZZ = np.meshgrid(xx, yy, zz, indexing="ij")
                pts = np.stack([XX.ravel(), YY.ravel(), ZZ.ravel()], axis=1)
                diff = to_cart(g, pts) - to_cart(g, im)
                dd = np.linalg.norm(diff, axis=1)
                v = self.image_rotation(J, im) @ u[J]
                # dV/dX_img = -V'(d)/d (x - X_img)
                dV = (-sp.dvloc_over_d(dd) * (diff @ v)).reshape(XX.shape)
                dbJ = self._box_laplacian(dV, xx[0])
                IIb, JJb, KKb = np.meshgrid(np.arange(i_s, i_e + 1), np.arange(j_s, j_e + 1),
                                            np.arange(k_s, k_e + 1), indexing="ij")
                rows = ((KKb - 1) * g.nr * g.nt + (JJb - 1) * g.nr + IIb - 1).ravel()
                np.add.at(db, rows, dbJ.ravel())
        return db * (-1 / (4 * math.pi))

    def atomic_density_derivative(self, u, scale=None):
        """First-order change of the superposed atomic density rho_at under
        the Cartesian displacements u (n_atm x 3): the initial density
        response S.Drho_dR of the DFPT loop (``electrostatics_local``,
        ``src/phonon.m:2402-2408``: minus the field-point gradient of each
        image's density on its pseudocharge box, rotated into the atom's
        frame).  ``scale`` is the normalisation that ``pseudocharge()``
        applies to rho_at (default: the one at the current geometry)."""
        g = self.grid
        sp = self.species
        u = np.asarray(u, dtype=float).reshape(self.n_atm, 3)
        if scale is None:
            scale = self.pseudocharge()["rho_scale"]
        drho = np.zeros(g.N)
        rbx, rby, rbz = self.rb_xyz()
        for J in range(self.n_atm):
            for im in self._images(self.atoms[J]):
                x0i, y0i, z0i = im
                i_s = max(math.ceil(((x0i - g.xin) - rbx) / g.dr) + 1, 1)
                i_e = min(math.floor(((x0i - g.xin) + rbx) / g.dr) + 1, g.nr)
                j_s = max(math.ceil((y0i - rby) / g.dt) + 1, 1)
                j_e = min(math.floor((y0i + rby) / g.dt) + 1, g.nt)
                k_s = max(math.ceil((z0i - rbz) / g.dz) + 1, 1)
                k_e = min(math.floor((z0i + rbz) / g.dz) + 1, g.nz)
                if i_e < i_s or j_e < j_s or k_e < k_s:
                    continue
                II, JJ, KK = np.meshgrid(np.arange(i_s, i_e + 1), np.arange(j_s, j_e + 1), np.arange(k_s, k_e + 1),
                                         indexing="ij")
                pts = np.stack([g.xin + (II.ravel() - 1) * g.dr, (JJ.ravel() - 1) * g.dt, (KK.ravel() - 1) * g.dz],
                               axis=1)
                diff = to_cart(g, pts) - to_cart(g, im)
                dd = np.linalg.norm(diff, axis=1)
                v = self.image_rotation(J, im) @ u[J]
                # d rho(|x - X|)/dX . v = -rho'(d)/d (x - X) . v
                rows = ((KK - 1) * g.nr * g.nt + (JJ - 1) * g.nr + II - 1).ravel()
                np.add.at(drho, rows, -sp.drho_iso_over_d(dd) * (diff @ v))
        return drho * scale


# --------------------------------------------------------------------------
# synthetic nanotube sectors
# --------------------------------------------------------------------------
def zigzag_sector_atoms(n, atoms_per_sector, twist=0.0, d_cc=1.42):
    """Atoms of an (n,0) zigzag tube in (r, theta, z) over one sector holding
    ``atoms_per_sector`` atoms (4 per circumferential period a = sqrt(3) d_cc).

    Returns (atoms, R, L2, L3).  Positions are given in helical coordinates
    (theta = phi - twist z is the coordinate the grid uses), so with a
    non-zero twist the Cartesian tube is twisted by ``twist`` rad/Bohr.
    """
    d = d_cc * BOHR_PER_ANG
    a = math.sqrt(3) * d
    R = n * a / (2 * math.pi)
    periods = atoms_per_sector // 4
    L2 = 2 * math.pi * periods / n
    L3 = 3 * d
    base = [(0.0, 0.0), (0.0, d), (a / 2, 1.5 * d), (a / 2, 2.5 * d)]
    th0, z0 = 0.25 * (2 * math.pi / n), 0.125 * L3
    atoms = []
    for pnum in range(periods):
        for x, z in base:
            atoms.append([R, th0 + (x + pnum * a) / R, (z0 + z) % L3])
    return np.array(atoms), R, L2, L3


def make_system(cfg: str) -> CyclixSystem:
    """The two single-channel workloads.

    ``cnt6_cyclic``: (6,0) tube, 1/6 sector, 4 C, no twist, 40 x 16 x 24
    nodes, 24 real states, Chebyshev degree 16 (cyclic, Gamma channel).

    ``cnt10_helical``: (10,0) tube, 1/5 sector, 8 C, twist 0.05 rad/Bohr,
    120 x 24 x 96 nodes, 256 complex states, degree 24, channel
    (nu, eta) = (1, 0.25 * 2 pi / L3).

    Radial extent: outer vacuum 8 Bohr; inner radius R - 8 Bohr clamped to
    1.5 Bohr so that r stays positive (the reference's symmetric-vacuum rule
    ``src/initialization.m:186-193`` would put the inner radius below 0).
    """
    if cfg in ("cnt6_cyclic", "cfg1"):
        atoms, R, L2, L3 = zigzag_sector_atoms(6, 4)
        nr, nt, nz = 40, 16, 24
        xin = max(R - 8.0, 1.5)
        xout = R + 8.0
        g = GridSpec(nr, nt, nz, xin, (xout - xin) / (nr - 1), L2 / nt, L3 / nz, 0.0, 6, 3)
        return CyclixSystem(g, atoms, Species(), nstates=24, npl=16, kpt=(0.0, 0.0, 0.0), name="cnt6_cyclic")
    if cfg in ("cnt10_helical", "cfg2", "cnt10_helical_80x64"):
        tw = 0.05
        atoms, R, L2, L3 = zigzag_sector_atoms(10, 8, twist=tw)
        # the second grid of the batched workload: 80 x 24 x 64 nodes over the same sector
        nr, nt, nz = (80, 24, 64) if cfg == "cnt10_helical_80x64" else (120, 24, 96)
        xin = max(R - 8.0, 1.5)
        xout = R + 8.0
        g = GridSpec(nr, nt, nz, xin, (xout - xin) / (nr - 1), L2 / nt, L3 / nz, tw, 6, 5)
        return CyclixSystem(g, atoms, Species(), nstates=256, npl=24,
                            kpt=(0.0, 1.0, 0.25 * 2 * math.pi / L3), name=cfg if cfg != "cfg2" else "cnt10_helical")
    raise ValueError(f"unknown workload {cfg!r}")


def small_system(nr=14, nt=16, nz=14, twist=0.0, kpt=(0.0, 0.0, 0.0), nstates=8, npl=8, natoms=4):
    """A grid small enough for dense brute force (N of a few thousand)."""
    n = 6
    atoms, R, L2, L3 = zigzag_sector_atoms(n, natoms)
    xin = R - 2.0
    dr = 4.0 / (nr - 1)
    cell = 5 if twist != 0.0 else 3
    g = GridSpec(nr, nt, nz, xin, dr, L2 / nt, L3 / nz, twist, 6, cell)
    sp = Species(rc_nl=1.6, rb=2.4, sigma_nl=0.5)
    return CyclixSystem(g, atoms, sp, nstates=nstates, npl=npl, kpt=kpt, name="small")


def scf_system(nr=28, nt=12, nz=16, vacuum=6.0, twist=0.0, nstates=12, npl=12):
    """(6,0) tube, 1/6 sector, 4 C, on a coarse grid (default 28 x 12 x 16,
    N = 5376) with 6 Bohr of radial vacuum on each side: a whole SCF runs
    in seconds on the CPU.  The pseudocharge integrates to -16 within 1e-5."""
    atoms, R, L2, L3 = zigzag_sector_atoms(6, 4, twist=twist)
    xin = max(R - vacuum, 1.5)
    xout = R + vacuum
    cell = 5 if twist != 0.0 else 3
    g = GridSpec(nr, nt, nz, xin, (xout - xin) / (nr - 1), L2 / nt, L3 / nz, twist, 6, cell)
    return CyclixSystem(g, atoms, Species(), nstates=nstates, npl=npl, kpt=(0.0, 0.0, 0.0), name="cnt6_scf")
