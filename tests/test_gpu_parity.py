"""HIP path vs the FP64 CPU oracle, through the C-ABI.

Small grids (N of a few thousand) run the oracle in seconds; every dtype and
both symmetries (cyclic Gamma, helical Bloch channel) are covered, with r
tiles and z blocks split so that tile seams and z wraps are exercised.
Tolerances: FP64 paths 1e-11 relative to the block's max; FP32 paths 2e-5.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import cyclix_oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu

CASES = {
    # name: (nr, nt, nz, twist, kpt, dtypes)
    "cyclic_gamma": (50, 16, 14, 0.0, (0.0, 0.0, 0.0), ("f64", "f32", "c128", "c64")),
    "helical_bloch": (45, 16, 14, 0.05, (0.0, 1.0, 0.6), ("c128", "c64")),
    "cyclic_bloch": (20, 13, 12, 0.0, (0.0, 2.0, 0.0), ("c128",)),
    # theta rings wider than 42 nodes: radial tiles narrower than the 2 FDN
    # halo (7 and 5 radii); 8 and 9 z planes, fewer than the 13-point z stencil
    "helical_wide": (26, 70, 8, 0.05, (0.0, 1.0, 0.6), ("c128", "c64")),
    "cyclic_wide": (22, 100, 9, 0.0, (0.0, 0.0, 0.0), ("f64", "f32")),
}
TOL = {"f64": 1e-11, "c128": 1e-11, "f32": 2e-5, "c64": 2e-5}
NP = {"f64": np.float64, "c128": np.complex128, "f32": np.float32, "c64": np.complex64}


def make_case(name):
    from cyclix_msparc_amd.system import small_system
    nr, nt, nz, tw, kpt, dts = CASES[name]
    sysm = small_system(nr=nr, nt=nt, nz=nz, twist=tw, kpt=kpt, nstates=8, npl=10)
    g = sysm.grid
    r = np.tile(g.r, g.nt * g.nz)
    R = sysm.atoms[:, 0].max()
    veff = -1.5 * np.exp(-((r - R) / 1.2) ** 2) + 0.1 * np.cos(np.arange(g.N) * 0.37)
    proj = sysm.projectors()
    og = O.Grid(g.nr, g.nt, g.nz, g.xin, g.dr, g.dt, g.dz, g.twist, g.fdn, g.cell_typ)
    atoms = [O.ProjAtom(a.coords, a.gamma, [O.ProjImage(im.rc_pos, im.chi, im.coords) for im in a.images])
             for a in proj]
    oham = O.Hamiltonian(og, tuple(kpt), veff, sysm.W, atoms)
    return sysm, veff, proj, oham, dts


def to_dev(X, dt, dev):
    """(N, ns) column-major numpy block -> (ns, N) device tensor."""
    return torch.from_numpy(np.ascontiguousarray(X.T.astype(NP[dt]))).to(dev.torch_device)


def to_host(T):
    return T.cpu().numpy().T.astype(np.complex128 if T.is_complex() else np.float64)


def rand_block(N, ns, dt, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((N, ns))
    if dt in ("c128", "c64"):
        X = X + 1j * rng.standard_normal((N, ns))
    return X


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("name", list(CASES))
def test_h_apply_parity(device, name):
    from cyclix_msparc_amd.api import Hamiltonian
    sysm, veff, proj, oham, dts = make_case(name)
    ham = Hamiltonian(device, sysm, sysm.kpt, veff, proj)
    for dt in dts:
        X = rand_block(sysm.grid.N, 5, dt, 11)
        Xq = X.astype(NP[dt]).astype(X.dtype)  # the block the device sees
        HX = to_host(ham.apply(to_dev(X, dt, device)))
        ref = oham.apply(Xq)
        print(f"h_apply {name} {dt}: rel err {rel(HX, ref):.3e}", flush=True)
        assert rel(HX, ref) < TOL[dt], (name, dt)


@pytest.mark.parametrize("name", list(CASES))
def test_chebyshev_filter_parity(device, name):
    from cyclix_msparc_amd.api import Hamiltonian
    sysm, veff, proj, oham, dts = make_case(name)
    ham = Hamiltonian(device, sysm, sysm.kpt, veff, proj)
    lo, hi = O.lanczos_bounds(oham.apply, np.random.default_rng(1).random(sysm.grid.N), 300, 1e-3,
                              O.metric_weights(oham.g))
    bup, a0 = 1.01 * hi, lo - 0.1
    cut = 0.5 * (bup + a0)
    for dt in dts:
        for m in (1, 2, 7):
            X = rand_block(sysm.grid.N, 4, dt, 5)
            Xq = X.astype(NP[dt]).astype(X.dtype)
            Y = to_host(ham.chebyshev_filter(to_dev(X, dt, device), m, cut, bup, a0))
            ref = O.chebyshev_filter(oham.apply, Xq, m, cut, bup, a0)
            tol = TOL[dt] * (10 if dt in ("f32", "c64") else 1) * m
            print(f"filter {name} {dt} m={m}: rel err {rel(Y, ref):.3e}", flush=True)
            assert rel(Y, ref) < tol, (name, dt, m)


def _subspace_projector(psi, Wm, V):
    M = psi.conj().T @ (Wm[:, None] * psi)
    return psi @ np.linalg.solve(M, psi.conj().T @ (Wm[:, None] * V))


@pytest.mark.parametrize("name", ["cyclic_gamma", "helical_bloch", "helical_wide"])
def test_eigsolver_pass_parity(device, name):
    from cyclix_msparc_amd.api import Hamiltonian
    sysm, veff, proj, oham, dts = make_case(name)
    ham = Hamiltonian(device, sysm, sysm.kpt, veff, proj)
    N, ns, npl = sysm.grid.N, 8, 10
    Wm = O.metric_weights(oham.g)
    lo, hi = O.lanczos_bounds(oham.apply, np.random.default_rng(1).random(N), 300, 1e-3, Wm)
    bup, a0 = 1.01 * hi, lo - 0.1
    cut = 0.5 * (bup + a0)
    dt = "c128" if any(sysm.kpt) else "f64"
    X = rand_block(N, ns, dt, 0)
    ev_ref, psi_ref = O.eig_solver_pass(oham, X, npl, cut, bup, a0)
    psi_d = to_dev(X, dt, device)
    ev = ham.eig_pass(psi_d, npl, cut, bup, a0)
    psi = to_host(psi_d)
    print(f"eig pass {name}: eig err {np.max(np.abs(ev - ev_ref)):.3e}", flush=True)
    assert np.allclose(ev, ev_ref, rtol=1e-9, atol=1e-9)
    V = np.random.default_rng(3).standard_normal((N, 3))
    assert rel(_subspace_projector(psi, Wm, V), _subspace_projector(psi_ref, Wm, V)) < 1e-8
    # W-normalised columns (src/eigSolver.m:122-124)
    nrm = np.sum(sysm.W[:, None] * np.abs(psi) ** 2, axis=0)
    assert np.allclose(nrm, 1.0, atol=1e-12)
    # a few more passes: still the same iterates as the oracle
    for it in range(4):
        cut, a0 = ev.max() + 0.1, ev.min()
        ev = ham.eig_pass(psi_d, npl, cut, bup, a0)
        ev_ref, psi_ref = O.eig_solver_pass(oham, psi_ref, npl, cut, bup, a0)
    assert np.allclose(ev, ev_ref, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("name,dt", [("cyclic_gamma", "f64"), ("cyclic_gamma", "f32"),
                                     ("helical_bloch", "c128"), ("helical_bloch", "c64"),
                                     ("helical_wide", "c64"), ("cyclic_wide", "f32")])
def test_rayleigh_ritz_many_states(device, name, dt):
    """Rayleigh-Ritz on 70 states: the Gram products cover several 64 x 64
    output regions (Hermitian overlap: upper regions mirrored; H projection:
    all regions) with a partial last region, the rotation several column
    passes.  Eigenvalues and the spanned subspace follow the oracle's metric
    Rayleigh-Ritz (src/eigSolver.m:99-124) on the same (quantised) block."""
    from cyclix_msparc_amd.api import Hamiltonian
    sysm, veff, proj, oham, dts = make_case(name)
    ham = Hamiltonian(device, sysm, sysm.kpt, veff, proj)
    N, ns = sysm.grid.N, 70
    X = rand_block(N, ns, dt, 11)
    Xq = X.astype(NP[dt]).astype(X.dtype)
    ev_ref, psi_ref = O.rayleigh_ritz_w(oham, Xq)
    psi_d = to_dev(Xq, dt, device)
    ev = ham.rayleigh_ritz(psi_d)
    psi = to_host(psi_d)
    tol = 1e-9 if dt in ("f64", "c128") else 2e-5
    err = np.max(np.abs(ev - ev_ref)) / np.max(np.abs(ev_ref))
    print(f"rayleigh-ritz {name} {dt} ns={ns}: eigenvalue rel err {err:.3e}", flush=True)
    assert err < tol
    Wm = O.metric_weights(oham.g)
    V = np.random.default_rng(5).standard_normal((N, 4))
    assert rel(_subspace_projector(psi, Wm, V), _subspace_projector(psi_ref, Wm, V)) < (1e-8 if tol < 1e-6 else 1e-3)
    nrm = np.sum(sysm.W[:, None] * np.abs(psi) ** 2, axis=0)
    assert np.allclose(nrm, 1.0, atol=1e-12 if tol < 1e-6 else 1e-5)
    ham.close()


@pytest.mark.parametrize("name,dt", [("cyclic_gamma", "f32"), ("helical_bloch", "c64")])
def test_fp32_rotation_gate(device, name, dt):
    """FP32 blocks: the overlap runs on the FP32 matrix cores with FP64
    chunk sums; when its Cholesky factor certifies the block well
    conditioned, the projection and the rotation stay there, else all three
    products run on the FP64 cores.  A random block takes the FP32 products;
    a block with two nearly parallel columns the FP64 ones; both match the
    oracle's Rayleigh-Ritz on the same block."""
    from cyclix_msparc_amd.api import Hamiltonian
    sysm, veff, proj, oham, dts = make_case(name)
    ham = Hamiltonian(device, sysm, sysm.kpt, veff, proj)
    N, ns = sysm.grid.N, 24
    Wm = O.metric_weights(oham.g)
    V = np.random.default_rng(8).standard_normal((N, 4))
    for near, want in ((False, (1, 0)), (True, (1, 1))):
        X = rand_block(N, ns, dt, 31)
        if near:
            X[:, 5] = X[:, 4] + 1e-3 * X[:, 5]
        Xq = X.astype(NP[dt]).astype(X.dtype)
        ev_ref, psi_ref = O.rayleigh_ritz_w(oham, Xq)
        psi_d = to_dev(Xq, dt, device)
        ev = ham.rayleigh_ritz(psi_d)
        err = np.max(np.abs(ev - ev_ref)) / np.max(np.abs(ev_ref))
        sub = rel(_subspace_projector(to_host(psi_d), Wm, V), _subspace_projector(psi_ref, Wm, V))
        print(f"{name} {dt} near-parallel={near}: rotations (fp32, fp64) {ham.rotation_counts()}, "
              f"eigenvalue rel err {err:.2e}, subspace err {sub:.2e}", flush=True)
        assert ham.rotation_counts() == want
        assert err < 2e-5 and sub < 1e-3
    ham.close()


@pytest.mark.parametrize("name,dt,ns", [("cyclic_gamma", "f64", 150), ("cyclic_gamma", "f64", 256),
                                        ("helical_bloch", "c128", 150), ("helical_bloch", "c128", 256)])
def test_rayleigh_ritz_tridiagonal_paths(device, name, dt, ns, monkeypatch):
    """The dense pencil's tridiagonal reduction three ways on the same block:
    over several workgroups with the matrix in LDS (2, 4 groups for the real
    sizes; 3, 8 for the complex ones), in one workgroup, and by rocSOLVER's
    heevd.  Each matches the oracle's metric Rayleigh-Ritz
    (src/eigSolver.m:99-124) in eigenvalues and spanned subspace."""
    from cyclix_msparc_amd.api import Hamiltonian
    sysm, veff, proj, oham, dts = make_case(name)
    ham = Hamiltonian(device, sysm, sysm.kpt, veff, proj)
    N = sysm.grid.N
    X = rand_block(N, ns, dt, 23)
    ev_ref, psi_ref = O.rayleigh_ritz_w(oham, X)
    Wm = O.metric_weights(oham.g)
    V = np.random.default_rng(6).standard_normal((N, 4))
    P_ref = _subspace_projector(psi_ref, Wm, V)
    for mode in ("dist", "one", "lapack"):
        monkeypatch.setenv("CYX_TRIDIAG", mode)
        psi_d = to_dev(X, dt, device)
        ev = ham.rayleigh_ritz(psi_d)
        psi = to_host(psi_d)
        err = np.max(np.abs(ev - ev_ref)) / np.max(np.abs(ev_ref))
        print(f"rayleigh-ritz {name} {dt} ns={ns} tridiagonal={mode}: eigenvalue rel err {err:.3e}", flush=True)
        assert err < 1e-9
        assert rel(_subspace_projector(psi, Wm, V), P_ref) < 1e-8
        nrm = np.sum(sysm.W[:, None] * np.abs(psi) ** 2, axis=0)
        assert np.allclose(nrm, 1.0, atol=1e-12)
    ham.close()


@pytest.mark.parametrize("kpt,twist", [((0, 0, 0), 0.0), ((0, 1.0, 0.6), 0.05)])
def test_chefsi_converges_to_dense_eigenvalues(device, kpt, twist):
    """Repeated device passes reach the lowest eigenvalues of the dense H."""
    from cyclix_msparc_amd.api import Hamiltonian
    from cyclix_msparc_amd.system import small_system
    sysm = small_system(nr=12, nt=14, nz=13, twist=twist, kpt=kpt, nstates=6)
    g = sysm.grid
    r = np.tile(g.r, g.nt * g.nz)
    veff = -1.5 * np.exp(-((r - sysm.atoms[:, 0].max()) / 1.2) ** 2)
    proj = sysm.projectors()
    og = O.Grid(g.nr, g.nt, g.nz, g.xin, g.dr, g.dt, g.dz, g.twist, g.fdn, g.cell_typ)
    oham = O.Hamiltonian(og, tuple(kpt), veff, sysm.W,
                         [O.ProjAtom(a.coords, a.gamma, [O.ProjImage(i.rc_pos, i.chi, i.coords) for i in a.images])
                          for a in proj])
    lam = np.sort(np.real(np.linalg.eigvals(oham.dense())))
    ham = Hamiltonian(device, sysm, kpt, veff, proj)
    dt = "c128" if any(kpt) else "f64"
    psi_d = to_dev(rand_block(g.N, 6, dt, 0), dt, device)
    lo, hi = ham.lanczos_bounds(np.random.default_rng(0).random(g.N), 300, 1e-3)
    bup, a0 = 1.01 * hi, lo - 0.1
    cut = 0.5 * (bup + a0)
    for it in range(40):
        ev = ham.eig_pass(psi_d, 14, cut, bup, a0)
        a0, cut = ev.min(), ev.max() + 0.1
    assert np.allclose(ev, lam[:6], atol=1e-8)


@pytest.mark.parametrize("name", ["cyclic_gamma", "helical_bloch", "helical_wide"])
def test_lanczos_bounds_parity(device, name):
    from cyclix_msparc_amd.api import Hamiltonian
    sysm, veff, proj, oham, dts = make_case(name)
    ham = Hamiltonian(device, sysm, sysm.kpt, veff, proj)
    v0 = np.random.default_rng(4).random(sysm.grid.N)
    lo_ref, hi_ref = O.lanczos_bounds(oham.apply, v0.astype(complex) if any(sysm.kpt) else v0, 300, 1e-3,
                                      O.metric_weights(oham.g))
    lo, hi = ham.lanczos_bounds(v0, 300, 1e-3)
    assert lo == pytest.approx(lo_ref, rel=1e-6, abs=1e-6)
    assert hi == pytest.approx(hi_ref, rel=1e-6)


@pytest.mark.parametrize("helical,nt,nz", [(False, 16, 14), (True, 16, 14), (True, 70, 8)])
def test_aar_poisson_parity(device, helical, nt, nz):
    from cyclix_msparc_amd.api import Poisson
    from cyclix_msparc_amd.system import GridSpec
    g = GridSpec(30, nt, nz, 2.0, 0.25, 2 * math.pi / 6 / nt, 0.35 * 14 / nz, 0.06 if helical else 0.0, 6,
                 5 if helical else 3)
    og = O.Grid(g.nr, g.nt, g.nz, g.xin, g.dr, g.dt, g.dz, g.twist, g.fdn, g.cell_typ)
    rng = np.random.default_rng(8)
    rho = np.abs(rng.standard_normal(g.N))
    b = -rho.mean() * np.ones(g.N)
    x_ref, relres_ref, it_ref = O.poisson_solve(og, rho, b, 1e-8, None, 1000, "jacobi")
    ps = Poisson(device, g)
    dev = device.torch_device
    phi, info = ps.solve(torch.from_numpy(rho).to(dev), torch.from_numpy(b).to(dev), None, 1e-8, 1000)
    phi = phi.cpu().numpy()
    print(f"aar helical={helical}: it {info.iterations} vs {it_ref}, relres {info.relres:.2e}", flush=True)
    assert info.converged and info.relres <= 1e-8
    assert abs(info.iterations - it_ref) <= 2
    assert np.linalg.norm(phi - x_ref) / np.linalg.norm(x_ref) < 1e-7
    f = O.poisson_rhs(rho, b)
    assert np.linalg.norm(f - O.lap_matrix(og) @ phi) / np.linalg.norm(f) <= 1.0001e-8


@pytest.mark.parametrize("cap", [3, 4, 5, 6, 7, 8, 12, 13])
def test_aar_iterate_at_iteration_cap(device, cap):
    """At the iteration cap AAR returns the iterate whose residual the last
    iteration measured (src/aar.m:33-72), whichever of the rotating device
    buffers holds it and whether the last step was Richardson or Anderson
    (p = 6); the oracle's iterate after the same number of steps.  The
    Anderson step at count 6 sees 5 of the m = 7 history columns, so its
    DF'DF is singular and takes the eigen-decomposition pseudo-inverse (from
    cap 7 on: at the cap itself the step stops before solving for gamma); the
    one at count 12 has the full history and takes the certified Cholesky
    solve (cap 13)."""
    from cyclix_msparc_amd.api import Poisson
    from cyclix_msparc_amd.system import GridSpec
    g = GridSpec(30, 16, 14, 2.0, 0.25, 2 * math.pi / 6 / 16, 0.35, 0.06, 6, 5)
    og = O.Grid(g.nr, g.nt, g.nz, g.xin, g.dr, g.dt, g.dz, g.twist, g.fdn, g.cell_typ)
    rng = np.random.default_rng(9)
    rho = np.abs(rng.standard_normal(g.N))
    b = -rho.mean() * np.ones(g.N)
    x_ref, relres_ref, it_ref = O.poisson_solve(og, rho, b, 1e-14, None, cap, "jacobi")
    dev = device.torch_device
    phi, info = Poisson(device, g).solve(torch.from_numpy(rho).to(dev), torch.from_numpy(b).to(dev), None, 1e-14, cap)
    phi = phi.cpu().numpy()
    err = np.linalg.norm(phi - x_ref) / np.linalg.norm(x_ref)
    print(f"aar cap {cap}: pinv_eig {info.pinv_eig}, rel err {err:.2e}", flush=True)
    assert not info.converged and info.iterations == cap == it_ref
    assert info.relres == pytest.approx(relres_ref, rel=1e-9)
    assert info.pinv_eig == (1 if cap >= 7 else 0)
    assert err < 1e-11


def test_aar_reports_nonconvergence(device):
    from cyclix_msparc_amd.api import Poisson
    from cyclix_msparc_amd.system import GridSpec
    g = GridSpec(20, 16, 14, 2.0, 0.25, 2 * math.pi / 6 / 16, 0.35, 0.0, 6, 3)
    rho = torch.ones(g.N, dtype=torch.float64, device=device.torch_device)
    b = -0.5 * rho
    phi, info = Poisson(device, g).solve(rho, b, None, 1e-12, 3)
    assert not info.converged and info.iterations == 3


def test_errors_are_reported(device):
    from cyclix_msparc_amd import _native as nat
    from cyclix_msparc_amd.api import Hamiltonian
    sysm, veff, proj, oham, dts = make_case("helical_bloch")
    ham = Hamiltonian(device, sysm, sysm.kpt, veff, proj)
    X = torch.zeros((2, sysm.grid.N), dtype=torch.float64, device=device.torch_device)
    with pytest.raises(nat.CyxError):  # real dtype on a Bloch channel
        ham.apply(X)
