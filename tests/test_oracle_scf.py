"""Pins of the oracle's SCF loop (src/scf.m:32-306) by what a converged
self-consistent state must satisfy: the fixed point does not depend on the
mixing parameters, the occupied states are eigenpairs of the final
Hamiltonian, the density holds Nelectron and phi solves the Poisson problem
of that density."""
import numpy as np
import pytest

from oracle import cyclix_oracle as O
from tests.scf_common import oracle_inputs, oracle_scf, psi0_for


@pytest.fixture(scope="module")
def sysm():
    from cyclix_msparc_amd.system import scf_system
    return scf_system()


@pytest.fixture(scope="module")
def converged(sysm):
    return oracle_scf(sysm, [psi0_for(sysm)], 1e-7, 1e-9)


def test_scf_converges_and_counts_electrons(sysm, converged):
    assert converged["converged"]
    assert converged["history"][-1]["err"] <= 1e-7
    assert np.sum(converged["rho"] * sysm.W) == pytest.approx(sysm.nelectron, abs=1e-10)


def test_scf_fixed_point_independent_of_mixing(sysm, converged):
    other = oracle_scf(sysm, [psi0_for(sysm, seed=5)], 1e-7, 1e-9, hist=4, beta=0.2)
    assert other["converged"]
    assert other["Eatom"] == pytest.approx(converged["Eatom"], abs=1e-8)
    nocc = int(sysm.nelectron // 2)
    assert np.allclose(other["EigVal"][:nocc], converged["EigVal"][:nocc], atol=1e-6)


def test_scf_states_are_eigenpairs_of_final_hamiltonian(sysm, converged):
    g, atoms, _ = oracle_inputs(sysm)
    H = O.Hamiltonian(g, (0.0, 0.0, 0.0), converged["veff"], sysm.W, atoms)
    psi, ev = converged["psi"][0], converged["EigVal"][:, 0]
    R = H.apply(psi) - psi * ev[None, :]
    res = np.sqrt(np.sum(sysm.W[:, None] * np.abs(R) ** 2, axis=0))
    nocc = int(sysm.nelectron // 2)
    assert np.all(res[:nocc] < 1e-5)


def test_scf_phi_solves_poisson_of_final_density(sysm, converged):
    g, _, pc = oracle_inputs(sysm)
    L = O.lap_matrix(g)
    f = O.poisson_rhs(converged["rho"], pc["b"])
    assert np.linalg.norm(L @ converged["phi"] - f) / np.linalg.norm(f) <= 1e-9
