"""Shared set-up of the SCF tests: the coarse (6,0) sector system and the
oracle's inputs built from it."""
import numpy as np

from oracle import cyclix_oracle as O


def oracle_inputs(sysm):
    g0 = sysm.grid
    g = O.Grid(g0.nr, g0.nt, g0.nz, g0.xin, g0.dr, g0.dt, g0.dz, g0.twist, g0.fdn, g0.cell_typ)
    atoms = [O.ProjAtom(a.coords, a.gamma, [O.ProjImage(im.rc_pos, im.chi, im.coords) for im in a.images])
             for a in sysm.projectors()]
    return g, atoms, sysm.pseudocharge()


def oracle_scf(sysm, psi0, SCF_tol, poisson_tol, **kw):
    g, atoms, pc = oracle_inputs(sysm)
    return O.scf(g, sysm.W, atoms, [(0.0, 0.0, 0.0)], [1.0], pc["b"], pc["rho_at"], pc["Eself"], pc["E_corr"],
                 sysm.nelectron, sysm.n_atm, psi0, sysm.npl, SCF_tol, poisson_tol, **kw)


def psi0_for(sysm, seed=0):
    """(N, Nev) N(0,1) start block."""
    return np.random.default_rng(seed).standard_normal((sysm.grid.N, sysm.nstates))
