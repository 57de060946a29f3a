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
