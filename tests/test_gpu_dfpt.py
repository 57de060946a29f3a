"""DFPT right-hand sides and graph-captured Sternheimer solves on the device
(src/phonon.m:1358-1444, 1516-1521), through the C-ABI:

* cyx_sternheimer_rhs against the oracle's restatement of the reference's
  right-hand side (small grids, three dtypes);
* the Sternheimer plan (iterations replayed from a HIP graph) against the
  host-polled solve and the oracle's bicgstab;
* the filter plan with the Sternheimer iterations as a second branch of its
  graph: the filtered blocks and the solutions equal those of the two run
  apart;
* the 48 right-hand sides of three random displacement patterns (seed 1) on
  the 120 x 24 x 96 helical channel, from a converged SCF: right-hand sides
  against the oracle, solutions by their true residual under the oracle's
  operator."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import cyclix_oracle as O  # noqa: E402
from tests.test_gpu_parity import make_case, rel, to_dev, to_host  # noqa: E402
from tests.test_gpu_sternheimer import occupied_states  # noqa: E402
from tests.test_oracle_dfpt import _oatoms, _odatoms  # noqa: E402

pytestmark = pytest.mark.gpu

BET = 27.211386245988 / 0.2


@pytest.mark.parametrize("name,dt", [("cyclic_gamma", "f64"), ("helical_bloch", "c128"), ("helical_bloch", "c64")])
def test_rhs_parity(device, name, dt):
    from cyclix_msparc_amd import dfpt
    from cyclix_msparc_amd.api import Hamiltonian
    sysm, veff, proj, oham, _ = make_case(name)
    g = sysm.grid
    N = g.N
    u = dfpt.random_displacements(sysm.n_atm, 1, seed=1)[0]
    dchi = sysm.projector_derivatives(u, proj)
    rng = np.random.default_rng(3)
    r = np.tile(g.r, g.nt * g.nz)
    dveff = 0.4 * np.sin(0.11 * np.arange(N)) * np.exp(-((r - sysm.atoms[:, 0].max()) / 1.5) ** 2)
    nv = 5
    cx = dt in ("c128", "c64")
    psi = rng.standard_normal((N, nv)) + (1j * rng.standard_normal((N, nv)) if cx else 0)
    lam = np.sort(rng.standard_normal(nv))
    coeff = dfpt.coeff_pv_psi(lam, np.ones(nv), BET, nv)
    ref, dl_ref = O.sternheimer_rhs(sysm.W, sysm.kpt, dveff, _oatoms(proj), _odatoms(proj, dchi),
                                    psi.astype(complex), coeff)
    dham = Hamiltonian(device, sysm, sysm.kpt, dveff, dfpt.perturbation_projectors(proj, dchi))
    B, dl = dham.sternheimer_rhs(to_dev(psi, dt, device), nv, coeff)
    tol = 1e-11 if dt in ("f64", "c128") else 5e-5
    err = rel(to_host(B), ref if cx else ref.real)
    print(f"{name} {dt}: rhs rel err {err:.2e}", flush=True)
    assert err < tol
    assert np.max(np.abs(dl - dl_ref)) <= tol * np.max(np.abs(dl_ref))
    dham.close()


def _solve_case(device):
    from cyclix_msparc_amd.api import Hamiltonian
    sysm, veff, proj, oham, _ = make_case("helical_bloch")
    N = sysm.grid.N
    ev, psi = occupied_states(oham, sysm, 4)
    lam = np.array([ev[0], ev[1], ev[2], ev[3], ev[1], ev[0]])
    rng = np.random.default_rng(8)
    B = rng.standard_normal((N, 6)) + 1j * rng.standard_normal((N, 6))
    ham = Hamiltonian(device, sysm, sysm.kpt, veff, proj)
    return sysm, veff, proj, oham, ham, psi, lam, B


def test_plan_matches_solve_and_oracle(device):
    from cyclix_msparc_amd.api import SternheimerPlan
    sysm, veff, proj, oham, ham, psi, lam, B = _solve_case(device)
    tol = 1e-9
    Bd, Pd = to_dev(B, "c128", device), to_dev(psi, "c128", device)
    X1, info1, its1, rel1 = ham.sternheimer_solve(Bd, lam, Pd, tol=tol, max_iter=3000, raise_on_noconv=True)
    sp = SternheimerPlan(ham, Bd, lam, Pd, tol=tol)
    act, info, its, relres = sp.solve(max_iter=3000, poll=4)
    assert act == 0 and info.n_converged == 6 and np.all(relres <= tol)
    assert np.array_equal(its, its1)
    assert rel(to_host(sp.X), to_host(X1)) <= 1e-13
    # frozen columns: further replays change nothing
    X_before = sp.X.clone()
    sp.iterate(4)
    torch.cuda.synchronize()
    assert torch.equal(sp.X, X_before)
    Xo, relo, itso, flo = O.sternheimer_solve(oham, lam, psi, B, tol=tol, maxit=3000)
    assert np.all(flo == 0)
    X = to_host(sp.X)
    for j in range(6):
        assert np.linalg.norm(X[:, j] - Xo[:, j]) <= 1e-5 * np.linalg.norm(Xo[:, j])
    print(f"plan iterations {its.tolist()} (oracle {itso.tolist()})", flush=True)
    sp.close()
    ham.close()


def test_filter_plan_carries_sternheimer_iterations(device):
    """One graph launch per pass: the filter steps of three channels and four
    BiCGSTAB iterations of six right-hand sides (an independent branch)."""
    from cyclix_msparc_amd.api import FilterPlan, Hamiltonian, SternheimerPlan
    sysm, veff, proj, oham, ham, psi, lam, B = _solve_case(device)
    tol = 1e-9
    Bd, Pd = to_dev(B, "c128", device), to_dev(psi, "c128", device)
    X1, _, its1, _ = ham.sternheimer_solve(Bd, lam, Pd, tol=tol, max_iter=3000, raise_on_noconv=True)
    g = sysm.grid
    kpts = [sysm.kpt, (0.0, 2.0, 0.3), (0.0, 0.0, -0.4)]
    fh = [Hamiltonian(device, sysm, k, veff, proj) for k in kpts]
    lo, hi = fh[0].lanczos_bounds(np.random.default_rng(1).random(g.N), 300, 1e-2)
    a0, bup = lo - 0.1, 1.01 * hi
    a = 0.5 * (a0 + bup)
    torch.manual_seed(2)
    init = [torch.randn((n, g.N), dtype=torch.complex128, device=device.torch_device) for n in (7, 3, 10)]
    plain = [x.clone() for x in init]
    p0 = FilterPlan(device, fh, plain, 5)
    p0.run(a, bup, a0)
    carried = [x.clone() for x in init]
    p1 = FilterPlan(device, fh, carried, 5)
    sp = SternheimerPlan(ham, Bd, lam, Pd, tol=tol)
    p1.attach_sternheimer(sp, 4)
    p1.run(a, bup, a0)
    torch.cuda.synchronize()
    for x, y in zip(carried, plain):
        assert torch.equal(x, y)
    passes = 1
    while sp.status()[0] > 0 and passes < 1000:
        p1.run(a, bup, a0)
        passes += 1
    act, info, its, relres = sp.status()
    assert act == 0 and info.n_converged == 6
    assert np.array_equal(its, its1)
    assert rel(to_host(sp.X), to_host(X1)) <= 1e-13
    print(f"{passes} combined launches, iterations {its.tolist()}", flush=True)
    p1.attach_sternheimer(None)
    p1.close()
    p0.close()
    sp.close()
    for h in fh:
        h.close()
    ham.close()


def test_dfpt_48_rhs_on_the_production_channel(device):
    """Three displacement patterns (seed 1) x 16 occupied states of the
    (10,0) helical channel at 120 x 24 x 96, from a converged FP64 SCF."""
    import time

    from cyclix_msparc_amd import dfpt
    from cyclix_msparc_amd.api import SternheimerPlan
    from cyclix_msparc_amd.scf import GroundState, SCFOptions
    from cyclix_msparc_amd.system import make_system
    sysm = make_system("cnt10_helical")
    g = sysm.grid
    gs = GroundState(device, sysm, options=SCFOptions(), precision="fp64")
    res = gs.run()
    assert res.converged
    occ = res.occ[:, 0]
    nv = dfpt.occupied_count(occ)
    print(f"occupied states (occ > 1e-5): {nv}", flush=True)
    assert 16 <= nv <= 48
    npert = math.ceil(48 / nv)
    U = dfpt.random_displacements(sysm.n_atm, npert, seed=1)
    proj = sysm.projectors()
    kpt = gs.kpts[0]
    psi = gs.psi[0]
    B, lam, dl, dveffs = dfpt.sternheimer_rhs_batch(device, sysm, kpt, psi, res.EigVal[:, 0], occ, gs.opt.bet,
                                                    gs.rho, U, proj=proj)
    assert B.shape == (npert * nv, g.N)
    coeff = dfpt.coeff_pv_psi(res.EigVal[:, 0], occ, gs.opt.bet, nv)
    P = psi[:nv].cpu().numpy().T
    for p in range(npert):
        dchi = sysm.projector_derivatives(U[p], proj)
        ref, dlr = O.sternheimer_rhs(sysm.W, kpt, dveffs[p].cpu().numpy(), _oatoms(proj), _odatoms(proj, dchi), P,
                                     coeff)
        err = rel(to_host(B[p * nv:(p + 1) * nv]), ref)
        print(f"pattern {p}: rhs rel err {err:.2e}, dlambda {np.round(dl[p][:4], 6).tolist()}", flush=True)
        assert err < 1e-10
        assert np.max(np.abs(dl[p] - dlr)) <= 1e-10 * max(np.max(np.abs(dlr)), 1e-300)
    ham = gs.hams[0]
    t0 = time.perf_counter()
    B, lam = B[:48], lam[:48]
    sp = SternheimerPlan(ham, B, lam, psi[:nv], tol=1e-8)
    act, info, its, relres = sp.solve(max_iter=3000, poll=8)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(f"48 RHS at 120x24x96: {info.max_iterations} iterations max, {dt:.2f} s", flush=True)
    assert act == 0 and info.n_converged == 48
    og = O.Grid(g.nr, g.nt, g.nz, g.xin, g.dr, g.dt, g.dz, g.twist, g.fdn, g.cell_typ)
    oham = O.Hamiltonian(og, kpt, gs.veff.cpu().numpy(), sysm.W, _oatoms(proj))
    cols = [0, 17, 33, 47]
    X = sp.X[cols].cpu().numpy().T
    Bh = B[cols].cpu().numpy().T
    R = Bh - O.sternheimer_apply(oham, lam[cols], P, X)
    true_rel = np.linalg.norm(R, axis=0) / np.linalg.norm(Bh, axis=0)
    print(f"true relative residuals of columns {cols}: {true_rel}", flush=True)
    assert np.all(true_rel <= 1e-7)
    sp.close()
    gs.close()
