"""Reference run decks: ``.inpt`` / ``.ion`` / ``.psp8`` files to a
``CyclixSystem`` and the SCF options of the run.

Host set-up only, in the order ``src/initialization.m:13-41`` runs it:
the input keys (``src/read_inpt.m``), the atoms (``src/read_ion.m``), the
ONCV psp8 pseudopotential (``src/readPseudopot.m:15-96``), the Cyclix grid
and inner radius (``src/initialization.m:166-202,318-349``) and the
pseudocharge radii (``Calculate_rb``, ``src/initialization.m:47-134``).
Only what the Cyclix decks of the reference use is read: one atom type,
LDA, cyclic or helical sectors (``BC: d c p`` / ``d c h``).
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
from scipy.interpolate import CubicSpline

from .system import CyclixSystem, GridSpec, _matlab_round, fd_weights, generate_kpts, to_cart


# ---------------------------------------------------------------- psp8
@dataclass
class Psp8:
    """``src/readPseudopot.m:15-96``: Z (valence), lmax, lloc (forced to 4),
    nproj per l, gamma_Jl per l, the radial grid, the projectors chi(r)
    (the file's r chi divided by r, the r = 0 value copied from the next
    point), Vloc, the isolated-atom valence density / (4 pi), and rc = the
    largest r_core of the title line."""

    Z: float
    lmax: int
    lloc: int
    r: np.ndarray
    nproj: list
    gamma: list   # per l: array (nproj,)
    proj: list    # per l: array (mmax, nproj), None for lloc
    vloc: np.ndarray
    rho_isolated: np.ndarray
    rc: float

    @staticmethod
    def read(path):
        text = Path(path).read_text()
        lines = text.splitlines()
        # title: "<sym> ONCVPSP-x r_core= rc_0 rc_1 ..." (lmax + 1 values)
        tit = lines[0].split()
        zline = lines[1].split()
        Z = float(zline[1])
        hdr = lines[2].split()
        lmax, mmax = int(hdr[2]), int(hdr[4])
        lloc = 4  # src/readPseudopot.m:29
        rcs = [float(v) for v in tit[3:3 + lmax + 1]]
        nproj = [int(v) for v in lines[4].split()[:lmax + 1]]
        # numbers from line 7 on (after the extension_switch line), in order
        toks = []
        for ln in lines[6:]:
            if ln.strip().startswith("<") or ln.strip() == "END_PSP":
                break
            toks.extend(ln.split())
        vals = np.array([float(t) for t in toks])
        pos = 0
        gamma, proj = [], []
        r = None
        for l in range(lmax + 1):
            pos += 1  # the l index
            if l != lloc:
                g = vals[pos:pos + nproj[l]]
                pos += nproj[l]
                blk = vals[pos:pos + (2 + nproj[l]) * mmax].reshape(mmax, 2 + nproj[l])
                pos += (2 + nproj[l]) * mmax
                r = blk[:, 1].copy()
                p = blk[:, 2:].copy()
                p[1:] = p[1:] / r[1:, None]
                p[0] = p[1]
                gamma.append(g.copy())
                proj.append(p)
            else:
                gamma.append(np.zeros(0))
                proj.append(None)
        if lloc > lmax:
            pos += 1
            blk = vals[pos:pos + 3 * mmax].reshape(mmax, 3)
            pos += 3 * mmax
            r = blk[:, 1].copy()
            vloc = blk[:, 2].copy()
        blk = vals[pos:pos + 5 * mmax].reshape(mmax, 5)
        rho = blk[:, 2] / (4 * math.pi)
        return Psp8(Z=Z, lmax=lmax, lloc=lloc, r=r, nproj=nproj, gamma=gamma, proj=proj, vloc=vloc,
                    rho_isolated=rho, rc=max(rcs))


class PspSpecies:
    """A psp8 species in the interface ``CyclixSystem`` reads: vloc(d),
    rho_iso(d), the projector channels and the radii.  Radial functions are
    interpolated with not-a-knot cubic splines, the interpolant of MATLAB's
    ``interp1(..., 'spline')``."""

    def __init__(self, psp: Psp8, name="X"):
        self.psp = psp
        self.name = name
        self.Z = psp.Z
        r = psp.r
        self._rV = CubicSpline(r, r * psp.vloc, bc_type="not-a-knot")
        self._rho = CubicSpline(r, psp.rho_isolated, bc_type="not-a-knot")
        self._proj = []
        for l in range(psp.lmax + 1):
            if l == psp.lloc or psp.proj[l] is None:
                continue
            for pp in range(psp.nproj[l]):
                self._proj.append((l, CubicSpline(r, psp.proj[l][:, pp], bc_type="not-a-knot"),
                                   float(psp.gamma[l][pp])))
        self.rb_xyz = None

    @property
    def rc(self):
        return self.psp.rc

    def vloc(self, d):
        """V_PS of ``src/calculate_b_guessRho_Eself.m:109-115``: -Z/d beyond
        the grid, spline(r Vloc)(d)/d inside, Vloc(0) below the second grid
        point."""
        d = np.asarray(d, dtype=float)
        r = self.psp.r
        out = np.empty_like(d)
        far = d > r[-1]
        out[far] = -self.Z
        out[~far] = self._rV(d[~far])
        with np.errstate(divide="ignore", invalid="ignore"):
            out = out / d
        out[d < r[1]] = self.psp.vloc[0]
        return out

    def rho_iso(self, d):
        d = np.asarray(d, dtype=float)
        out = self._rho(d)
        out[d > self.psp.r[-1]] = 0.0
        return out

    def drho_iso_over_d(self, d):
        """rho_iso'(d) / d from the spline; its limit rho_iso''(0) at d = 0."""
        d = np.asarray(d, dtype=float)
        out = np.zeros_like(d)
        inside = d <= self.psp.r[-1]
        small = inside & (d < 1e-8)
        reg = inside & ~small
        out[reg] = self._rho(d[reg], 1) / d[reg]
        out[small] = self._rho(d[small], 2)
        return out

    def projector_channels(self):
        return [(l, sp, g) for (l, sp, g) in self._proj]


# ---------------------------------------------------------------- input files
def read_inpt(path):
    """Key: value pairs of a ``.inpt`` file (``src/read_inpt.m``); '#' starts a comment."""
    keys = {}
    for ln in Path(path).read_text().splitlines():
        ln = ln.split("#", 1)[0].strip()
        if not ln or ":" not in ln:
            continue
        k, v = ln.split(":", 1)
        keys[k.strip().upper()] = v.split()
    return keys


def read_ion(path):
    """Atom types of a ``.ion`` file (``src/read_ion.m``): name, valence,
    psp8 file, coordinates and whether they are fractional."""
    types = []
    lines = [ln.split("#", 1)[0].strip() for ln in Path(path).read_text().splitlines()]
    i = 0
    while i < len(lines):
        ln = lines[i]
        if ln.startswith("ATOM_TYPE:"):
            f = ln.split()
            types.append(dict(name=f[1], n=0, psp=None, coords=None, frac=False))
        elif ln.startswith("N_TYPE_ATOM:"):
            types[-1]["n"] = int(ln.split()[1])
        elif ln.startswith("PSEUDO_POT:"):
            types[-1]["psp"] = ln.split()[1]
        elif ln.startswith("COORD_FRAC:") or ln.startswith("COORD:"):
            n = types[-1]["n"]
            rows = []
            j = i + 1
            while len(rows) < n:
                if lines[j]:
                    rows.append([float(v) for v in lines[j].split()[:3]])
                j += 1
            types[-1]["coords"] = np.array(rows)
            types[-1]["frac"] = ln.startswith("COORD_FRAC:")
            i = j - 1
        i += 1
    return types


# ---------------------------------------------------------------- pseudocharge radii
def _intg_wts_cychel(nx, ny, nz, xin, dx, dy, dz):
    """``Cychel/intgWts_cychel.m`` with Dirichlet ends in all three directions."""
    wx = (xin + np.arange(nx) * dx) * dx
    wx[0] = 0.5 * xin * dx + 0.125 * dx * dx
    wx[-1] = 0.5 * (xin + (nx - 1) * dx) * dx - 0.125 * dx * dx
    wy = np.full(ny, dy)
    wy[0] *= 0.5
    wy[-1] *= 0.5
    wz = np.full(nz, dz)
    wz[0] *= 0.5
    wz[-1] *= 0.5
    return wx[:, None, None] * wy[None, :, None] * wz[None, None, :]


def find_rb(species: PspSpecies, grid: GridSpec, xmax_at, xvac, tol):
    """``Calculate_rb`` (``src/initialization.m:47-134``) for a Cyclix
    sector: b of one atom at (xmax_at, 0, 0) on a box reaching the radial
    vacuum and 12 Bohr along theta (as an angle) and z; rb grows by
    min(dx, dz) from rc until |2 int_{z >= 0} b W + Z| <= tol."""
    g = grid
    n0 = g.fdn
    dx, dy, dz = g.dr, g.dt, g.dz

    def f_rby(y):
        return math.acos(1 - y * y / (2 * xmax_at * xmax_at)) + g.twist * y * 0.5

    rb_up_x, rb_up_y, rb_up_z = xvac, f_rby(12.0), 12.0
    ii_s, ii_e = -math.ceil(rb_up_x / dx), math.ceil(rb_up_x / dx)
    jj_s, jj_e = -math.ceil(rb_up_y / dy), math.ceil(rb_up_y / dy)
    kk_s, kk_e = 0, math.ceil(rb_up_z / dz)
    xx = xmax_at + np.arange(ii_s - n0, ii_e + n0 + 1) * dx
    yy = np.arange(jj_s - n0, jj_e + n0 + 1) * dy
    zz = np.arange(kk_s - n0, kk_e + n0 + 1) * dz
    XX, YY, ZZ = np.meshgrid(xx, yy, zz, indexing="ij")
    pts = np.stack([XX.ravel(), YY.ravel(), ZZ.ravel()], axis=1)
    dd = np.linalg.norm(to_cart(g, pts) - to_cart(g, [xmax_at, 0.0, 0.0]), axis=1).reshape(XX.shape)
    W = _intg_wts_cychel(ii_e - ii_s + 1, jj_e - jj_s + 1, kk_e - kk_s + 1, xx[n0], dx, dy, dz)
    tmp = CyclixSystem(g, np.zeros((1, 3)), species, nstates=1, npl=1)
    b = -tmp._box_laplacian(species.vloc(dd), xx[0]) / (4 * math.pi)  # interior of the box
    rc = species.rc
    rb_x = math.ceil(rc / dx - 1e-12) * dx
    rb_z = math.ceil(rc / dz - 1e-12) * dz
    rb_y = math.ceil(f_rby(rc) / dy - 1e-12) * dy
    ds = min(dx, dz)
    err, count = 100.0, 1
    while err > tol and count <= 300 and rb_x <= rb_up_x and rb_y <= rb_up_y and rb_z <= rb_up_z:
        rb_x += ds
        rb_z += ds
        rb_y = f_rby(rb_z)
        fx, fy, fz = math.floor(rb_x / dx), math.floor(rb_y / dy), math.floor(rb_z / dz)
        # interior-box indices (0-based) of the reference's ii_rb - FDn etc.
        ia = np.arange(-ii_s - fx, -ii_s + fx + 1)
        ja = np.arange(-jj_s - fy, -jj_s + fy + 1)
        ka = np.arange(0, fz + 1)
        sub = np.ix_(ia, ja, ka)
        err = abs(2 * float(np.sum(W[sub] * b[sub])) + species.Z)
        count += 1
    if not (rb_x <= rb_up_x and rb_y <= rb_up_y and rb_z <= rb_up_z):
        raise ValueError("pseudocharge radii reached their upper bounds")
    return rb_x, rb_y, rb_z


# ---------------------------------------------------------------- the deck
@dataclass
class Deck:
    system: CyclixSystem
    channels: list            # [(kpt, weight)]
    options: dict = field(default_factory=dict)   # SCFOptions fields
    meta: dict = field(default_factory=dict)


def mesh_to_cheb_degree(h):
    """Default Chebyshev degree from the effective mesh size
    (``Mesh2ChebDegree``, ``src/initialization.m:856-869``): a cubic fit in h,
    14 above h = 0.7, rounded."""
    if h > 0.7:
        return 14
    p3, p2, p1, p0 = -700.0 / 3.0, 1240.0 / 3.0, -773.0 / 3.0, 1078.0 / 15.0
    return int(_matlab_round(((p3 * h + p2) * h + p1) * h + p0))


def scf_tol_for_forces(target_force_accuracy, nelectron, n_atom):
    """Default SCF tolerance from a force target (``src/initialization.m:515-545``:
    log10 SCF_tol = (log10 F - b) / a, a = 1.03, b = -0.33145 + 1.1761
    log10(electrons per atom)); with no accuracy key the target is 1e-3 Ha/Bohr."""
    b = -0.33145 + 1.1761 * math.log10(nelectron / n_atom)
    return 10.0 ** ((math.log10(target_force_accuracy) - b) / 1.03)


def load_deck(directory, name, cell=None, fd_grid=None):
    """``<directory>/<name>.inpt`` + ``.ion`` + the psp8 it names, for the
    Cyclix decks (one atom type, LDA_PW).  Returns a Deck whose system holds
    the grid, the atoms in (r, theta, z), the psp8 species with its
    pseudocharge radii, NSTATES and CHEB_DEGREE.  ``cell`` = (L1, L2, L3)
    replaces CELL with FD_GRID and the fractional coordinates kept, as a
    cell relaxation step does (``src/cell_relax.m``); ``fd_grid`` = (Nx, Ny,
    Nz) replaces FD_GRID (intervals, as the key counts them)."""
    d = Path(directory)
    K = read_inpt(d / f"{name}.inpt")
    types = read_ion(d / f"{name}.ion")
    if len(types) != 1:
        raise ValueError("one atom type is supported")
    if K.get("EXCHANGE_CORRELATION", ["LDA_PW"])[0] != "LDA_PW":
        raise ValueError("LDA_PW is the supported exchange-correlation functional")
    L1, L2, L3 = (float(v) for v in K["CELL"][:3]) if cell is None else (float(v) for v in cell)
    nx, ny, nz = (int(v) for v in (K["FD_GRID"][:3] if fd_grid is None else fd_grid))
    if "BC" in K:
        bc = [v.lower() for v in K["BC"][:3]]
    elif "BOUNDARY_CONDITION" in K:
        # the obsolete numeric key (src/read_inpt.m:92-96): 5 is the cyclic
        # sector, 7 the helical one (src/initialization.m:277-298)
        code = int(K["BOUNDARY_CONDITION"][0])
        if code not in (5, 7):
            raise ValueError(f"BOUNDARY_CONDITION {code} is not a cyclic or helical sector")
        bc = ["d", "c", "p" if code == 5 else "h"]
    else:
        raise ValueError("the deck names no boundary condition (BC or BOUNDARY_CONDITION)")
    if bc[0] != "d" or bc[1] != "c" or bc[2] not in ("p", "h"):
        raise ValueError(f"BC {bc} is not a cyclic or helical sector")
    cell_typ = 5 if bc[2] == "h" else 3
    twist = float(K.get("TWIST_ANGLE", ["0"])[0]) if cell_typ == 5 else 0.0
    fdn = int(K.get("FD_ORDER", ["12"])[0]) // 2
    t = types[0]
    coords = np.array(t["coords"], dtype=float)
    if t["frac"]:
        atoms = coords * np.array([L1, L2, L3])
    else:
        X = coords
        atoms = np.stack([np.hypot(X[:, 0], X[:, 1]), np.arctan2(X[:, 1], X[:, 0]) - twist * X[:, 2], X[:, 2]], 1)
    xmin, xmax = atoms[:, 0].min(), atoms[:, 0].max()
    xvac = (L1 - (xmax - xmin)) / 2
    xin = xmin - xvac
    dx, dy, dz = L1 / nx, L2 / ny, L3 / nz
    # Dirichlet radial direction: S.Nx = intervals + 1 nodes (src/initialization.m:342-345)
    g = GridSpec(nx + 1, ny, nz, xin, dx, dy, dz, twist, fdn, cell_typ)
    psp = Psp8.read(d / t["psp"])
    sp = PspSpecies(psp, t["name"])
    nelectron = psp.Z * len(atoms)
    scf_tol = float(K["TOL_SCF"][0]) if "TOL_SCF" in K else scf_tol_for_forces(1e-3, nelectron, len(atoms))
    # the default pseudocharge tolerance is 0.01 SCF_tol (src/initialization.m:556-560)
    tol = float(K["TOL_PSEUDOCHARGE"][0]) if "TOL_PSEUDOCHARGE" in K else 0.01 * scf_tol
    sp.rb_xyz = find_rb(sp, g, xmax, xvac, tol)
    nkpt = tuple(int(v) for v in K.get("KPOINT_GRID", ["1", "1", "1"])[:3])
    shift = tuple(float(v) for v in K.get("KPOINT_SHIFT", ["0", "0", "0"])[:3])
    kgrid, w = generate_kpts(nkpt, L1, L2, L3, kptshift=shift, cell_typ=cell_typ)
    channels = [(tuple(float(v) for v in k), float(wk)) for k, wk in zip(kgrid, w)]
    if "NSTATES" in K:
        nstates = int(K["NSTATES"][0])
    else:   # src/initialization.m:508-513
        nstates = int(_matlab_round(math.floor(nelectron / 2) * 1.2 + 5))
    if "CHEB_DEGREE" in K:
        npl = int(K["CHEB_DEGREE"][0])
    else:   # src/initialization.m:490-506 on the Cyclix effective mesh size
        xout = xmax + xvac
        h_eff = math.sqrt(3.0 / (1 / dx ** 2 + 1 / (((xin + xout) / 2) * dy) ** 2 + 1 / dz ** 2))
        npl = mesh_to_cheb_degree(h_eff)
    sysm = CyclixSystem(g, atoms, sp, nstates=nstates, npl=npl, kpt=channels[0][0], name=name)
    opts = {"SCF_tol": scf_tol}
    if "TOL_POISSON" in K:
        opts["poisson_tol"] = float(K["TOL_POISSON"][0])
    ttype = K.get("ELEC_TEMP_TYPE", ["gaussian"])[0].lower()
    opts["elec_T_type"] = 0 if ttype in ("fd", "fermi-dirac") else 1
    if "SMEARING" in K:
        opts["bet"] = 1.0 / float(K["SMEARING"][0])
    if "TOL_LANCZOS" in K:
        opts["TOL_LANCZOS"] = float(K["TOL_LANCZOS"][0])
    return Deck(system=sysm, channels=channels, options=opts,
                meta=dict(xvac=xvac, xmin_at=xmin, xmax_at=xmax, rb=sp.rb_xyz, nkpt=nkpt))
