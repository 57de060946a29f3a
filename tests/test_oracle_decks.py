"""The psp8 set-up path against data the reference's own runs hold
(``tests/Cyclix/CNT_relax`` and ``tests/Cyclix/CNT_ABINITmatch``: the
``.inpt``/``.ion``/``.psp8`` files copied under ``tests/golden/decks`` with
the energies their ``C.out`` files print).

* The psp8 reader (``src/readPseudopot.m:15-96``): header values, gammas,
  the valence density integrating to Z.
* The pseudocharge radius search (``src/initialization.m:47-134``) against
  the radii the CNT_relax run prints (6.49 0.49 6.48).
* The self and correction energy E_corr - E_self
  (``src/calculate_b_guessRho_Eself.m``, printed by
  ``src/electronicGroundState.m:95``) against both runs.  CNT_relax is a
  cell relaxation whose first SCF runs at L3 = 7.98 / RELAX_MAXDILAT = 6.65
  Bohr (its ``C.geopt`` RELAXSTEP 1); the deck's geometry is scaled there.
* The real spherical harmonics (``src/sphericalHarmonics.m``): orthonormal
  on the sphere.

The whole SCF of both decks against the printed energies runs on the GPU
(``tests/test_gpu_decks.py``)."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from cyclix_msparc_amd.deck import Psp8, load_deck, read_inpt, read_ion
from cyclix_msparc_amd.system import real_ylm

DECKS = Path(__file__).resolve().parent / "golden" / "decks"
REF = json.loads((DECKS / "reference_energies.json").read_text())
# the CNT_relax run's first SCF: L3 divided by RELAX_MAXDILAT = 1.2
RELAX_CELL = (20.0, 0.349065850398866, 7.98 / 1.2)


def test_psp8_reader():
    p = Psp8.read(DECKS / "CNT_relax" / "C.psp8")
    assert p.Z == 4.0 and p.lmax == 1 and p.lloc == 4 and p.nproj == [2, 2]
    assert p.rc == pytest.approx(1.42727)
    assert np.allclose(p.gamma[0], [8.1302824913405, 0.90835387171729])
    assert np.allclose(p.gamma[1], [-2.7474117539033, -0.83725610360987])
    assert p.r.size == 600 and p.r[1] == pytest.approx(0.01) and p.r[-1] == pytest.approx(5.99)
    assert p.vloc[0] == pytest.approx(-11.224703059543)
    # r chi / r, with the r = 0 value copied from r = 0.01
    assert p.proj[0][1, 0] == pytest.approx(5.1228774022581e-02 / 0.01)
    assert p.proj[0][0, 0] == p.proj[0][1, 0]
    # the valence density (stored / 4 pi) integrates to the valence charge
    q = np.trapezoid(4 * math.pi * p.r ** 2 * p.rho_isolated, p.r)
    assert q == pytest.approx(p.Z, rel=2e-3)


def test_input_readers():
    K = read_inpt(DECKS / "CNT_ABINITmatch" / "C.inpt")
    assert K["FD_GRID"] == ["100", "25", "40"] and K["BC"] == ["d", "c", "p"] and K["KPOINT_GRID"] == ["1", "16", "1"]
    t = read_ion(DECKS / "CNT_ABINITmatch" / "C.ion")
    assert len(t) == 1 and t[0]["n"] == 4 and t[0]["frac"] and t[0]["psp"] == "C.psp8"
    assert np.allclose(t[0]["coords"][:, 0], 0.590400269541860)


def test_pseudocharge_radii_match_the_reference_run():
    D = load_deck(DECKS / "CNT_relax", "C")
    assert [round(v, 2) for v in D.system.species.rb_xyz] == REF["CNT_relax"]["pseudocharge_radii_2dp"]
    g = D.system.grid
    assert (g.nr, g.nt, g.nz) == (81, 19, 32)
    assert g.dr == pytest.approx(0.25) and g.dz == pytest.approx(0.249375)
    # the printed mesh spacing along theta is dy * (xin + xout) / 2
    assert g.dt * (2 * g.xin + 20.0) / 2 == pytest.approx(0.243723, abs=1e-6)


@pytest.mark.parametrize("deck,cell,tol", [("CNT_relax", RELAX_CELL, 1e-7), ("CNT_ABINITmatch", None, 1e-8)])
def test_self_and_correction_energy_matches_the_reference_run(deck, cell, tol):
    D = load_deck(DECKS / deck, "C", cell=cell)
    pc = D.system.pseudocharge()
    esc = pc["E_corr"] - pc["Eself"]
    assert abs(esc - REF[deck]["self_and_correction_energy"]) <= tol
    # the pseudocharge carries the valence charge
    assert np.dot(D.system.W, pc["b"]) == pytest.approx(-D.system.nelectron, abs=1e-6)


def test_channels_of_the_sixteen_point_deck():
    D = load_deck(DECKS / "CNT_ABINITmatch", "C")
    assert len(D.channels) == REF["CNT_ABINITmatch"]["symmetry_adapted_kpoints"]
    assert sum(w for _, w in D.channels) == pytest.approx(1.0)


def test_real_spherical_harmonics_are_orthonormal():
    # Gauss-Legendre in cos(theta) x uniform in phi: exact for these degrees
    ct, wt = np.polynomial.legendre.leggauss(16)
    nph = 32
    ph = np.arange(nph) * 2 * math.pi / nph
    C, P = np.meshgrid(ct, ph, indexing="ij")
    S = np.sqrt(1 - C * C)
    X, Y, Z = S * np.cos(P), S * np.sin(P), C
    w = (wt[:, None] * np.full(nph, 2 * math.pi / nph)[None, :]).ravel()
    fs = [real_ylm(l, m, X, Y, Z).ravel() for l in range(3) for m in range(-l, l + 1)]
    G = np.array([[np.sum(w * a * b) for b in fs] for a in fs])
    assert np.allclose(G, np.eye(len(fs)), atol=1e-12)
    # scale-free, and 0 at the origin for l >= 1
    assert real_ylm(1, 1, np.array([2.0]), np.array([0.0]), np.array([0.0]))[0] == pytest.approx(math.sqrt(3 / (4 * math.pi)))
    assert real_ylm(1, 0, np.zeros(1), np.zeros(1), np.zeros(1))[0] == 0.0


def test_oracle_scf_matches_the_reference_run():
    """The oracle's whole SCF (``src/scf.m``) on the CNT_relax deck at its
    first relaxation cell: free energy per atom against the reference run's
    printed value (observed 5e-8 Ha/atom; about two minutes on 8 cores)."""
    from oracle import cyclix_oracle as O
    from tests.scf_common import oracle_inputs, psi0_for
    D = load_deck(DECKS / "CNT_relax", "C", cell=RELAX_CELL)
    s = D.system
    g, atoms, pc = oracle_inputs(s)
    o = D.options
    res = O.scf(g, s.W, atoms, [(0.0, 0.0, 0.0)], [1.0], pc["b"], pc["rho_at"], pc["Eself"], pc["E_corr"],
                s.nelectron, s.n_atm, [psi0_for(s)], s.npl, o["SCF_tol"], o["poisson_tol"], bet=o["bet"],
                kind=o["elec_T_type"])
    assert res["converged"]
    assert abs(res["Eatom"] - REF["CNT_relax"]["free_energy_per_atom"]) <= 1e-6
