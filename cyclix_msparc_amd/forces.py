"""Atomic forces of a converged ground state (``src/atomicForce.m:1-364``).

force = local + correction + nonlocal, in the reference's discretisation:

* Cartesian gradient operators of a Cyclix grid (``gradIndicesValues_cychel``,
  ``blochGradient_cychel``): d/dx = cos(phi)/dr D_r - sin(phi)/(r dtheta) D_theta,
  d/dy = sin(phi)/dr D_r + cos(phi)/(r dtheta) D_theta,
  d/dz = 1/dz D_z - tau/dtheta D_theta, with phi = theta + tau z, order-2 FDn
  first-derivative weights, Dirichlet truncation in r and Bloch phases across
  the theta and z wraps;
* the vector-field pair (M1, M2) of ``div_gradvec_cychel``: the same x and y
  operators, whose entries across the theta wrap also rotate the field
  components by the sector angle;
* local and correction terms per atom image over the pseudocharge boxes
  (``src/atomicForce.m:27-215``; the box gradient of V_J is
  ``Cychel/dpseudopot_cychel.m``), rotated back to the atom by the image's
  ROT;
* the nonlocal term per channel (``:220-350``): Chi_full, its rotated
  copies and their gradients on the host (sparse operators); the integrals
  psi^H W [Chi, DChi_x, DChi_y, DChi_z] as one weighted product on the device
  (``cyx_weighted_gram``).
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as spm

from .system import CyclixSystem, fd_weights, to_cart


def _phases(kpt, L2, L3):
    nu, eta = float(kpt[1]), float(kpt[2])
    pt = (np.exp(-1j * nu * L2), np.exp(1j * nu * L2)) if nu != 0 else (1.0, 1.0)
    pz = (np.exp(-1j * eta * L3), np.exp(1j * eta * L3)) if eta != 0 else (1.0, 1.0)
    return pt, pz


def gradient_operators(sysm: CyclixSystem, kpt=(0.0, 0.0, 0.0)):
    """(G1, G2, G3, M1, M2): N x N sparse operators of the Cartesian x, y, z
    derivatives with the channel's Bloch phases, and the vector-field pair of
    ``Cychel/div_gradvec_cychel.m``."""
    g = sysm.grid
    nx, ny, nz, n0 = g.nr, g.nt, g.nz, g.fdn
    _, w1 = fd_weights(n0)
    N = g.N
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ii, jj, kk = (a.ravel(order="F") for a in (ii, jj, kk))
    row = kk * nx * ny + jj * nx + ii
    x = g.xin + ii * g.dr
    ph = jj * g.dt + g.twist * kk * g.dz
    Xc, Yc = x * np.cos(ph), x * np.sin(ph)
    wt1_d1, wt1_d2 = Xc / (x * g.dr), -Yc / (x * x * g.dt)
    wt2_d1, wt2_d2 = Yc / (x * g.dr), Xc / (x * x * g.dt)
    (ptl, ptr), (pzl, pzr) = _phases(kpt, g.L2, g.L3)
    cx = isinstance(ptl, complex) or isinstance(pzl, complex) or np.iscomplexobj(ptl) or np.iscomplexobj(pzl)
    R, C, V1, V2, OUT = [], [], [], [], []   # OUT: -1 left theta wrap, +1 right, 0 inside
    for p in range(1, n0 + 1):
        for sgn in (1, -1):
            # radial taps (Dirichlet: dropped outside)
            i2 = ii + sgn * p
            ok = (i2 >= 0) & (i2 < nx)
            R.append(row[ok])
            C.append((kk * nx * ny + jj * nx + i2)[ok])
            V1.append(sgn * wt1_d1[ok] * w1[p])
            V2.append(sgn * wt2_d1[ok] * w1[p])
            OUT.append(np.zeros(ok.sum(), dtype=np.int8))
            # theta taps (periodic with the Bloch phase)
            j2 = jj + sgn * p
            out = np.where(j2 < 0, -1, np.where(j2 >= ny, 1, 0)).astype(np.int8)
            R.append(row)
            C.append(kk * nx * ny + (j2 % ny) * nx + ii)
            V1.append(sgn * wt1_d2 * w1[p])
            V2.append(sgn * wt2_d2 * w1[p])
            OUT.append(out)
    R, C, V1, V2, OUT = (np.concatenate(a) for a in (R, C, V1, V2, OUT))
    dt = complex if cx else float
    ph_t = np.where(OUT < 0, ptl, np.where(OUT > 0, ptr, 1.0)).astype(dt)
    G1 = spm.csr_matrix((V1 * ph_t, (R, C)), shape=(N, N))
    G2 = spm.csr_matrix((V2 * ph_t, (R, C)), shape=(N, N))
    # div_gradvec_cychel: across the wrap the neighbouring sector's field is
    # rotated by L2 (left: by -L2 seen from here)
    c2, s2 = math.cos(g.L2), math.sin(g.L2)
    lft, rgt = OUT < 0, OUT > 0
    M1v = (V1 * ph_t).astype(dt)
    M2v = (V2 * ph_t).astype(dt)
    M1v[lft] = (V1[lft] * c2 - V2[lft] * s2) * ph_t[lft]
    M2v[lft] = (V2[lft] * c2 + V1[lft] * s2) * ph_t[lft]
    M1v[rgt] = (V1[rgt] * c2 + V2[rgt] * s2) * ph_t[rgt]
    M2v[rgt] = (V2[rgt] * c2 - V1[rgt] * s2) * ph_t[rgt]
    M1 = spm.csr_matrix((M1v, (R, C)), shape=(N, N))
    M2 = spm.csr_matrix((M2v, (R, C)), shape=(N, N))
    # z derivative: -tau/dtheta D_theta + 1/dz D_z
    R3, C3, V3 = [], [], []
    for p in range(1, n0 + 1):
        for sgn in (1, -1):
            j2 = jj + sgn * p
            pt = np.where(j2 < 0, ptl, np.where(j2 >= ny, ptr, 1.0)).astype(dt)
            R3.append(row)
            C3.append(kk * nx * ny + (j2 % ny) * nx + ii)
            V3.append(sgn * (-g.twist / g.dt) * w1[p] * pt)
            k2 = kk + sgn * p
            pz = np.where(k2 < 0, pzl, np.where(k2 >= nz, pzr, 1.0)).astype(dt)
            R3.append(row)
            C3.append((k2 % nz) * nx * ny + jj * nx + ii)
            V3.append(sgn * (1.0 / g.dz) * w1[p] * pz)
    G3 = spm.csr_matrix((np.concatenate(V3), (np.concatenate(R3), np.concatenate(C3))), shape=(N, N))
    return G1, G2, G3, M1, M2


def _box_gradient(sysm, V, xx, yy, zz):
    """``Cychel/dpseudopot_cychel.m``: Cartesian gradient of a box function
    V (with its FDn halo) at the box interior."""
    g = sysm.grid
    n0 = g.fdn
    _, w1 = fd_weights(n0)
    X, Y, Z = np.meshgrid(xx[n0:-n0], yy[n0:-n0], zz[n0:-n0], indexing="ij")
    ph = Y + g.twist * Z
    Xc, Yc = X * np.cos(ph), X * np.sin(ph)
    c11, c12 = Xc / (X * g.dr), -Yc / (X * X * g.dt)
    c21, c22 = Yc / (X * g.dr), Xc / (X * X * g.dt)
    c31, c32 = -g.twist / g.dt, 1.0 / g.dz
    nxb, nyb, nzb = V.shape

    def sh(a, b, c):
        return V[n0 + a:nxb - n0 + a, n0 + b:nyb - n0 + b, n0 + c:nzb - n0 + c]

    dx = np.zeros(X.shape)
    dy = np.zeros(X.shape)
    dz = np.zeros(X.shape)
    for p in range(1, n0 + 1):
        dr_ = w1[p] * (sh(p, 0, 0) - sh(-p, 0, 0))
        dt_ = w1[p] * (sh(0, p, 0) - sh(0, -p, 0))
        dz_ = w1[p] * (sh(0, 0, p) - sh(0, 0, -p))
        dx += c11 * dr_ + c12 * dt_
        dy += c21 * dr_ + c22 * dt_
        dz += c31 * dt_ + c32 * dz_
    return dx, dy, dz


def _rot(sysm, J, im):
    """ROT = S.RotM1^fac1 * S.RotM2^fac2 with fac = (atom - image) / L."""
    return sysm.image_rotation(J, im).T


def local_forces(sysm: CyclixSystem, phi, pc=None):
    """Local and correction terms (``src/atomicForce.m:19-215``).  phi: the
    converged electrostatic potential (host, length N).  Returns (n_atm x 3,
    n_atm x 3)."""
    g = sysm.grid
    sp = sysm.species
    pc = sysm.pseudocharge() if pc is None else pc
    G1, G2, G3, _, _ = gradient_operators(sysm)
    dphi = [G.dot(phi).real for G in (G1, G2, G3)]
    dvc = [G.dot(pc["V_c"]).real for G in (G1, G2, G3)]
    bsum = pc["b"] + pc["b_ref"]
    W = sysm.W
    rbx, rby, rbz = sysm.rb_xyz()
    n0 = g.fdn
    rc_ref = 0.5
    f_loc = np.zeros((sysm.n_atm, 3))
    f_cor = np.zeros((sysm.n_atm, 3))
    for J in range(sysm.n_atm):
        for im in sysm._images(sysm.atoms[J]):
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
            V = sp.vloc(dd)
            Vr = np.where(dd < rc_ref,
                          -sp.Z * (9 * dd ** 7 - 30 * rc_ref * dd ** 6 + 28 * rc_ref ** 2 * dd ** 5
                                   - 14 * rc_ref ** 5 * dd ** 2 + 12 * rc_ref ** 7) / (5 * rc_ref ** 8),
                          -sp.Z / np.maximum(dd, 1e-300))
            bJ = -sysm._box_laplacian(V, xx[0]) / (4 * math.pi)
            bJr = -sysm._box_laplacian(Vr, xx[0]) / (4 * math.pi)
            dV = _box_gradient(sysm, V, xx, yy, zz)
            dVr = _box_gradient(sysm, Vr, xx, yy, zz)
            IIb, JJb, KKb = np.meshgrid(np.arange(i_s, i_e + 1), np.arange(j_s, j_e + 1), np.arange(k_s, k_e + 1),
                                        indexing="ij")
            rows = ((KKb - 1) * g.nr * g.nt + (JJb - 1) * g.nr + IIb - 1)
            w = W[rows]
            fl = np.array([-np.sum(w * bJ * (dphi[a][rows] - dV[a])) for a in range(3)])
            fc = np.array([0.5 * np.sum(w * (bsum[rows] * (dVr[a] - dV[a]) - (bJ + bJr) * dvc[a][rows]))
                           for a in range(3)])
            R = _rot(sysm, J, im)
            f_loc[J] += R @ fl
            f_cor[J] += R @ fc
    return f_loc, f_cor


def nonlocal_projection_fields(sysm: CyclixSystem, kpt, proj=None, ops=None):
    """Per atom: (Chi_full_k, DChi_x, DChi_y, DChi_z), each N x angnum
    (``src/atomicForce.m:283-301``)."""
    g = sysm.grid
    proj = sysm.projectors() if proj is None else proj
    _, _, G3, M1, M2 = gradient_operators(sysm, kpt) if ops is None else ops
    kv = np.asarray(kpt, dtype=float)
    out = []
    for J, A in enumerate(proj):
        ang = A.gamma.size
        chi = np.zeros((g.N, ang), dtype=complex)
        r1 = np.zeros((g.N, ang), dtype=complex)
        r2 = np.zeros((g.N, ang), dtype=complex)
        for im in A.images:
            ph = np.exp(-1j * np.dot(kv, A.coords - im.coords))
            R = _rot(sysm, J, im.coords)
            chi[im.rc_pos] += im.chi * ph
            r1[im.rc_pos] += R[0, 0] * im.chi * ph
            r2[im.rc_pos] += R[0, 1] * im.chi * ph
        dx = M1 @ r1 + M2 @ r2
        dy = -(M1 @ r2) + M2 @ r1
        dz = G3 @ chi
        out.append((chi, dx, dy, dz))
    return out


def nonlocal_forces(ham, psi, occ, weight, sysm: CyclixSystem, kpt, proj=None, occfac=2.0):
    """Nonlocal force of one channel (``src/atomicForce.m:274-350``): the
    integrals psi^H W [Chi, DChi_x, DChi_y, DChi_z] on the device
    (``Hamiltonian.weighted_gram``), combined with gamma and the occupations.
    psi: the channel's (Ns, N) device block; occ: host (Ns,)."""
    import torch

    proj = sysm.projectors() if proj is None else proj
    fields = nonlocal_projection_fields(sysm, kpt, proj)
    cols = np.concatenate([np.concatenate(f, axis=1) for f in fields], axis=1)  # N x (4 ang natm)
    if psi.is_complex():
        B = torch.from_numpy(np.ascontiguousarray(cols.T.astype(np.complex128))).to(psi.device).to(psi.dtype)
    else:
        B = torch.from_numpy(np.ascontiguousarray(cols.T.real)).to(psi.device).to(psi.dtype)
    G = ham.weighted_gram(psi, B).cpu().numpy()  # Ns x ncol: (psi^H W B)
    occ = np.asarray(occ, dtype=float)
    f = np.zeros((sysm.n_atm, 3))
    c0 = 0
    for J, A in enumerate(proj):
        ang = A.gamma.size
        g_chi = G[:, c0:c0 + ang]
        I1 = np.conj(g_chi).T                      # integral_1 = transpose(conj(PsiW) * conj(Chi))
        for a in range(3):
            g_d = G[:, c0 + (a + 1) * ang:c0 + (a + 2) * ang]
            I2 = -g_d.T                            # integral_2 = -transpose(PsiW * DChi)
            f[J, a] = A.gamma @ np.real(I1 * I2) @ occ
        c0 += 4 * ang
    return -occfac * 2 * weight * f


def drift_free(force, cell_typ=3):
    """Subtract the mean force, for a Cyclix sector only its z component
    (``src/electronicGroundState.m:108-115``)."""
    corr = force.mean(axis=0)
    if cell_typ in (3, 4, 5):
        corr[:2] = 0.0
    return force - corr[None, :]


def atomic_forces(gs, pc=None):
    """Forces on the atoms of a converged ``scf.GroundState`` (n_atm x 3,
    Ha/Bohr, Cartesian, before the drift correction), ``src/atomicForce.m``."""
    sysm = gs.system
    phi = gs.phi.cpu().numpy()
    f_loc, f_cor = local_forces(sysm, phi, pc)
    proj = sysm.projectors()
    occ = gs.result.occ
    f_nl = np.zeros_like(f_loc)
    for ks, (h, k) in enumerate(zip(gs.hams, gs.kpts)):
        f_nl += nonlocal_forces(h, gs.psi[ks], occ[:, ks], gs.wkpt[ks], sysm, k, proj)
    return f_loc + f_cor + f_nl
