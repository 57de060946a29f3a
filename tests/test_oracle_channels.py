"""Bloch channels of a cyclic / helical sector (``Generate_kpts``,
``src/initialization.m:1591-1682``): the set-up code and the oracle against
the channel counts the reference's own runs report, and against each other.

The counts are those the reference decks print as "Number of symmetry
adapted k-points" (``tests/Cyclix/CNT_ABINITmatch/C.out``: 9 for a 1/16
sector with KPOINT_GRID 1 16 1; ``LDA_CNT/C.out``: 8 for a 1/15 sector with
1 10 1; ``tests/Cyclix/MoS2_CYCLIC/MoS2.inpt``: a 1/17 sector with 1 17 1,
9 channels; single-channel decks: 1).  The 1/15 sector with 10 points
exercises MATLAB's rounding of halves away from zero (nu = 1.5 j)."""
import math

import numpy as np
import pytest

from cyclix_msparc_amd.system import generate_kpts
from oracle import cyclix_oracle as O

DECKS = [
    # (L1, L2, L3, nkpt, channels reported)
    (20.0, 0.392699081698724, 7.98, (1, 16, 1), 9),      # CNT_ABINITmatch
    (24.0, 0.418879020478639, 4.647785137, (1, 10, 1), 8),  # LDA_CNT
    (29.910212616, 0.369599135716446, 6.055850878784, (1, 17, 1), 9),  # MoS2_CYCLIC
    (20.0, 0.349065850398866, 7.98, (1, 1, 1), 1),       # CNT_relax
    (24.0, 0.698131700797732, 1.523592733066272, (1, 1, 1), 1),  # CNT_chiral
]


@pytest.mark.parametrize("L1,L2,L3,nkpt,count", DECKS)
def test_channel_counts_match_reference_runs(L1, L2, L3, nkpt, count):
    grid, w = generate_kpts(nkpt, L1, L2, L3)
    chans, wo = O.bloch_channels(nkpt, L1, L2, L3)
    assert grid.shape[0] == count == len(chans)
    assert np.sum(w) == pytest.approx(1.0, abs=1e-14)
    assert np.allclose(grid, np.array(chans)) and np.allclose(w, wo)


def test_fifteen_fold_sector_rounds_halves_away_from_zero():
    grid, w = generate_kpts((1, 10, 1), 24.0, 0.418879020478639, 4.647785137)
    assert grid[:, 1].tolist() == [0, 2, 5, 8, 9, 11, 12, 14]
    # nu = 9 and 12 stand for their time-reversal partners 6 and 3
    assert np.allclose(w * 10, [1, 1, 1, 1, 2, 1, 2, 1])


@pytest.mark.parametrize("nkpt,shift", [((1, 1, 4), 0.0), ((1, 1, 5), 0.5), ((1, 6, 3), 0.0), ((1, 4, 2), 0.25)])
def test_helical_grids_agree_with_oracle(nkpt, shift):
    L1, L2, L3 = 16.0, 2 * math.pi / 6, 8.05
    grid, w = generate_kpts(nkpt, L1, L2, L3, kptshift=(0.0, 0.0, shift))
    chans, wo = O.bloch_channels(nkpt, L1, L2, L3, kptshift=(0.0, 0.0, shift))
    assert np.allclose(grid, np.array(chans)) and np.allclose(w, wo)
    assert np.sum(w) == pytest.approx(1.0, abs=1e-14)
    # the axial wavenumbers lie in (-pi/L3, pi/L3]
    assert np.all(grid[:, 2] > -math.pi / L3 - 1e-12) and np.all(grid[:, 2] <= math.pi / L3 + 1e-12)
    # every dropped channel has its partner among the kept ones
    full, _ = generate_kpts(nkpt, L1, L2, L3, kptshift=(0.0, 0.0, shift), time_rev=False)
    kept = {(round(r[1]), round(r[2], 10)) for r in grid}
    Nc = round(2 * math.pi / L2)
    for r in full:
        key = (round(r[1]), round(r[2], 10))
        partners = {((Nc - key[0]) % Nc if key[0] else 0, round(-r[2], 10) + 0.0)}
        assert key in kept or any((p[0], p[1]) in kept for p in partners)
