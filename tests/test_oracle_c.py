"""The C restatement (oracle/cyx_oracle.c) agrees with the numpy oracle: two
independent restatements of the same reference routines."""
import math

import numpy as np
import pytest

from oracle import cyclix_oracle as O


@pytest.fixture(scope="module")
def c_oracle():
    from oracle import c_oracle
    try:
        c_oracle.load()
    except RuntimeError:
        import subprocess
        from pathlib import Path
        subprocess.run(["make", "-s", "-C", str(Path(__file__).resolve().parents[1] / "oracle")], check=True)
    return c_oracle


@pytest.mark.parametrize("twist,kpt", [(0.0, (0, 0, 0)), (0.05, (0, 1.0, 0.6)), (0.0, (0, 2.0, 0.0))])
def test_c_h_apply_and_filter_match_numpy(c_oracle, twist, kpt):
    from cyclix_msparc_amd.system import small_system
    sysm = small_system(nr=20, nt=14, nz=13, twist=twist, kpt=kpt)
    g = sysm.grid
    og = O.Grid(g.nr, g.nt, g.nz, g.xin, g.dr, g.dt, g.dz, g.twist, g.fdn, g.cell_typ)
    atoms = [O.ProjAtom(a.coords, a.gamma, [O.ProjImage(i.rc_pos, i.chi, i.coords) for i in a.images])
             for a in sysm.projectors()]
    veff = np.cos(np.arange(g.N) * 0.1)
    H = O.Hamiltonian(og, kpt, veff, sysm.W, atoms)
    Cc = c_oracle.COracle(og, kpt, veff, sysm.W, atoms)
    rng = np.random.default_rng(0)
    X = rng.standard_normal((g.N, 3)) + 1j * rng.standard_normal((g.N, 3))
    ref = H.apply(X)
    assert np.max(np.abs(Cc.apply(X) - ref)) < 1e-11 * np.max(np.abs(ref))
    Y = Cc.chebyshev_filter(X, 5, 10.0, 200.0, -1.0)
    Yr = O.chebyshev_filter(H.apply, X, 5, 10.0, 200.0, -1.0)
    assert np.max(np.abs(Y - Yr)) < 1e-10 * np.max(np.abs(Yr))


@pytest.mark.parametrize("helical", [False, True])
def test_c_aar_matches_numpy(c_oracle, helical):
    g = O.Grid(18, 14, 13, 2.0, 0.3, 2 * math.pi / 6 / 14, 0.35, 0.07 if helical else 0.0, 6, 5 if helical else 3)
    rng = np.random.default_rng(4)
    rho = np.abs(rng.standard_normal(g.N))
    b = -rho.mean() * np.ones(g.N)
    x_ref, rel_ref, it_ref = O.poisson_solve(g, rho, b, 1e-9, None, 3000, "jacobi")
    Cc = c_oracle.COracle(g, (0, 0, 0), np.zeros(g.N), O.intg_weights(g))
    x, rel, it = Cc.aar_jacobi(O.poisson_rhs(rho, b), None, 1e-9, 3000)
    assert rel <= 1e-9 and abs(it - it_ref) <= 2
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-8
