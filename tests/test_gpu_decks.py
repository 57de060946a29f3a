"""Whole SCF runs of the reference's own Cyclix decks on the device, against
the energies those runs printed (``tests/Cyclix/CNT_relax/C.out:86-92``,
``tests/Cyclix/CNT_ABINITmatch/C.out:82-88``; the decks and numbers are under
``tests/golden/decks``).

CNT_relax: (18,0) tube sector, 4 C, ONCV psp8 (s and p projectors), LDA-PW,
Fermi-Dirac smearing 0.001 Ha, grid 81 x 19 x 32, 12 states, Chebyshev
degree 80, Gamma channel; its first SCF runs at L3 = 6.65 Bohr (a cell
relaxation started at 7.98 / 1.2).  CNT_ABINITmatch: (16,0) sector, 4 C,
grid 101 x 25 x 40, 12 states, degree 100, KPOINT_GRID 1 16 1 (9
symmetry-adapted Bloch channels with time-reversal weights).

The budget of the energy target is 1e-5 Ha/atom; the runs agree to a few
1e-8 Ha/atom.  The Fermi level is compared for CNT_ABINITmatch only: CNT_relax
has a gap, inside which a 0.001 Ha smearing leaves the level undetermined."""
import json
from pathlib import Path

import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

DECKS = Path(__file__).resolve().parent / "golden" / "decks"
REF = json.loads((DECKS / "reference_energies.json").read_text())
RELAX_CELL = (20.0, 0.349065850398866, 7.98 / 1.2)


@pytest.mark.parametrize("deck,cell", [("CNT_relax", RELAX_CELL), ("CNT_ABINITmatch", None)])
def test_reference_deck_energies(device, deck, cell):
    from cyclix_msparc_amd.deck import load_deck
    from cyclix_msparc_amd.scf import GroundState, SCFOptions
    D = load_deck(DECKS / deck, "C", cell=cell)
    s = D.system
    gs = GroundState(device, s, channels=D.channels, options=SCFOptions(**D.options), precision="fp64")
    try:
        res = gs.run()
        rec = res.history[-1]
        ref = REF[deck]
        d = dict(Eatom=res.Eatom - ref["free_energy_per_atom"],
                 Exc=rec.Exc - ref["exchange_correlation_energy"],
                 Eband=rec.Eband - ref["band_structure_energy"],
                 Esc=(gs.Ecorr - gs.Eself) - ref["self_and_correction_energy"],
                 Ef=res.lambda_f - ref["fermi_level"])
        print(f"{deck}: {len(D.channels)} channels, {res.count_SCF} SCF iterations, differences from the "
              f"reference run: " + ", ".join(f"{k} {v:.2e}" for k, v in d.items()), flush=True)
        assert res.converged
        assert abs(d["Eatom"]) <= 1e-6
        assert abs(d["Esc"]) <= 1e-6
        assert abs(d["Exc"]) <= 1e-5
        # the band energy moves with the SCF tolerance (1e-6 on the potential)
        assert abs(d["Eband"]) <= 1e-4
        if deck == "CNT_ABINITmatch":
            assert abs(d["Ef"]) <= 1e-5
    finally:
        gs.close()


@pytest.mark.parametrize("deck,cell,tol", [("CNT_relax", RELAX_CELL, 1e-6), ("CNT_ABINITmatch", None, 5e-6)])
def test_reference_deck_forces(device, deck, cell, tol):
    """Drift-free atomic forces (``src/atomicForce.m``,
    ``src/electronicGroundState.m:104-160``) against the reference runs'
    average and maximum |F| (observed 2.5e-7 Ha/Bohr for CNT_relax; 1.6e-6
    for CNT_ABINITmatch, whose output comes from an earlier code version),
    and the per-atom forces CNT_ABINITmatch's C.static holds."""
    import numpy as np

    from cyclix_msparc_amd.deck import load_deck
    from cyclix_msparc_amd.forces import atomic_forces, drift_free
    from cyclix_msparc_amd.scf import GroundState, SCFOptions
    D = load_deck(DECKS / deck, "C", cell=cell)
    s = D.system
    gs = GroundState(device, s, channels=D.channels, options=SCFOptions(**D.options), precision="fp64")
    try:
        assert gs.run().converged
        F = drift_free(atomic_forces(gs), s.grid.cell_typ)
        mag = np.linalg.norm(F, axis=1)
        ref = REF[deck]
        print(f"{deck}: |F| average {mag.mean():.10f} (reference {ref['average_force']:.10f}), maximum "
              f"{mag.max():.10f} (reference {ref['maximum_force']:.10f})", flush=True)
        assert abs(mag.mean() - ref["average_force"]) <= tol
        assert abs(mag.max() - ref["maximum_force"]) <= tol
        if "static_forces" in ref:
            assert np.max(np.abs(F - np.array(ref["static_forces"]))) <= 1e-5
    finally:
        gs.close()


def test_forces_are_the_energy_gradient(device):
    """Away from the symmetric geometry (atoms moved by a seeded random 0.3 Bohr
    pattern), F . u equals -dE/du by central differences of the converged
    free energy along two random directions, to the discretisation gap of the
    reference's force formula (observed 0.1-0.2 %)."""
    from dataclasses import replace

    import numpy as np

    from cyclix_msparc_amd import dfpt
    from cyclix_msparc_amd.deck import load_deck
    from cyclix_msparc_amd.forces import atomic_forces
    from cyclix_msparc_amd.scf import GroundState, SCFOptions
    D = load_deck(DECKS / "CNT_relax", "C", cell=RELAX_CELL)
    opts = dict(D.options, SCF_tol=1e-9, poisson_tol=1e-11)
    s = replace(D.system, atoms=dfpt.displaced_atoms(D.system, dfpt.random_displacements(4, 1, seed=7)[0], 0.3))

    def solve(sys_):
        gs = GroundState(device, sys_, channels=D.channels, options=SCFOptions(**opts), precision="fp64")
        res = gs.run()
        assert res.converged
        return gs, res

    gs, _ = solve(s)
    F = atomic_forces(gs)
    gs.close()
    for seed in (1, 2):
        u = dfpt.random_displacements(4, 1, seed=seed)[0]
        E = []
        for eps in (1e-3, -1e-3):
            g2, r2 = solve(replace(s, atoms=dfpt.displaced_atoms(s, u, eps)))
            E.append(r2.Etotal)
            g2.close()
        fu, de = float(np.sum(F * u)), -(E[0] - E[1]) / 2e-3
        print(f"direction {seed}: F.u {fu:.6f}, -dE/du {de:.6f}", flush=True)
        assert abs(fu - de) <= 4e-3 * abs(de)
