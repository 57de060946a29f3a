"""Device SCF pieces against the oracle, through the C-ABI: Periodic-Pulay
mixing step by step, the SCF error and Escc sums, and a whole SCF run on the
coarse (6,0) sector (energy per atom, eigenvalues, Fermi level)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import cyclix_oracle as O  # noqa: E402
from tests.scf_common import oracle_scf, psi0_for  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("hist,beta,omega,pfreq", [(7, 0.3, None, 1), (3, 0.5, 0.1, 2), (1, 0.3, None, 1)])
def test_mixer_matches_oracle(device, hist, beta, omega, pfreq):
    from cyclix_msparc_amd.api import Mixer
    N = 5000
    rng = np.random.default_rng(11)
    td = device.torch_device
    x = rng.standard_normal(N)
    A = 0.4 * rng.standard_normal(N)
    c = rng.standard_normal(N)
    mix_o = O.PeriodicPulay(N, hist, beta, omega, pfreq)
    mix_d = Mixer(device, torch.from_numpy(x).to(td), hist, beta, -1 if omega is None else omega, pfreq)
    for it in range(1, 12):
        g = np.tanh(A * x) + c                    # a nonlinear map, so the history columns are independent
        ref = mix_o(g, x, it)
        out = mix_d.step(torch.from_numpy(g).to(td), torch.from_numpy(x).to(td), it).cpu().numpy()
        assert np.max(np.abs(out - ref)) <= 1e-11 * np.max(np.abs(ref)), f"iteration {it}"
        x = ref
    mix_d.close()


def test_scf_terms_match_numpy(device):
    from cyclix_msparc_amd.api import Hamiltonian
    from cyclix_msparc_amd.system import scf_system
    sysm = scf_system()
    N = sysm.grid.N
    rng = np.random.default_rng(3)
    vo, vi, rho = rng.standard_normal(N), rng.standard_normal(N), rng.random(N)
    h = Hamiltonian(device, sysm, (0.0, 0.0, 0.0), np.zeros(N), sysm.projectors())
    td = device.torch_device
    escc, dv2, v2 = h.scf_terms(*(torch.from_numpy(a).to(td) for a in (vo, vi, rho)))
    assert escc == pytest.approx(np.sum((vo - vi) * rho * sysm.W), rel=1e-12)
    assert dv2 == pytest.approx(np.sum((vo - vi) ** 2), rel=1e-12)
    assert v2 == pytest.approx(np.sum(vo ** 2), rel=1e-12)
    h.close()


def test_scf_matches_oracle(device):
    """Same system, same psi0, same tolerances: the device SCF reaches the
    oracle's fixed point (energy per atom within 1e-7 Ha, occupied
    eigenvalues within 1e-6 Ha)."""
    from cyclix_msparc_amd.scf import GroundState, SCFOptions
    from cyclix_msparc_amd.system import scf_system
    sysm = scf_system()
    psi0 = psi0_for(sysm)
    ref = oracle_scf(sysm, [psi0], 1e-7, 1e-9)
    gs = GroundState(device, sysm, options=SCFOptions(SCF_tol=1e-7, poisson_tol=1e-9))
    res = gs.run(psi0=[psi0.T])
    assert res.converged and ref["converged"]
    assert res.Eatom == pytest.approx(ref["Eatom"], abs=1e-7)
    nocc = int(sysm.nelectron // 2)
    assert np.max(np.abs(res.EigVal[:nocc] - ref["EigVal"][:nocc])) < 1e-6
    assert res.lambda_f == pytest.approx(ref["lambda_f"], abs=1e-6)
    assert abs(res.count_SCF - ref["count_SCF"]) <= 6
    # the first SCF iteration's terms follow the same trajectory closely
    h0, r0 = res.history[0], ref["history"][0]
    assert h0.Etotal == pytest.approx(r0["Etotal"], abs=1e-6)


def test_mixed_precision_scf_error_against_fp64_oracle(device):
    """The complex64/FP32 path (FP32 psi, filter and H-applies; FP64 Gram
    accumulation, Cholesky, dense eigensolve, density and energies) converged
    as far as FP32 allows, against the FP64 oracle's fixed point: energy per
    atom within 1e-5 Ha (measured about 4e-7) and occupied eigenvalues within
    1e-6 Ha (measured about 1.5e-7)."""
    from cyclix_msparc_amd.scf import GroundState, SCFOptions
    from cyclix_msparc_amd.system import scf_system
    sysm = scf_system()
    psi0 = psi0_for(sysm)
    ref = oracle_scf(sysm, [psi0], 1e-7, 1e-9)
    gs = GroundState(device, sysm, options=SCFOptions(SCF_tol=1e-7, poisson_tol=1e-9), precision="fp32")
    assert gs.hams[0].complex is False and gs.precision == "fp32"
    res = gs.run(psi0=[psi0.T])
    assert gs.psi[0].dtype == torch.float32
    nocc = int(sysm.nelectron // 2)
    assert abs(res.Eatom - ref["Eatom"]) <= 1e-5
    assert np.max(np.abs(res.EigVal[:nocc] - ref["EigVal"][:nocc])) <= 1e-6
    print(f"mixed-precision error: dE/atom = {res.Eatom - ref['Eatom']:.3e} Ha, "
          f"max |d eig| = {np.max(np.abs(res.EigVal[:nocc] - ref['EigVal'][:nocc])):.3e} Ha")


def test_multichannel_scf_matches_oracle(device):
    """Two Bloch channels from Generate_kpts (src/initialization.m:1591-1682)
    on the 1/6 sector with KPOINT_GRID 1 3 1: nu = 0 (real, weight 1/3) and
    nu = 4 (complex, weight 2/3, standing for nu = 2 by time reversal).  The
    device SCF (density summed over channels with their weights, one Fermi
    level over both) reaches the oracle's fixed point."""
    from cyclix_msparc_amd.scf import GroundState, SCFOptions
    from cyclix_msparc_amd.system import generate_kpts, scf_system
    from tests.scf_common import oracle_inputs
    sysm = scf_system()
    g0 = sysm.grid
    L1 = (g0.nr - 1) * g0.dr
    kpts, w = generate_kpts((1, 3, 1), L1, g0.L2, g0.L3)
    assert kpts[:, 1].tolist() == [0.0, 4.0] and np.allclose(w, [1 / 3, 2 / 3])
    rng = np.random.default_rng(2)
    N, ns = g0.N, sysm.nstates
    psi0 = [rng.standard_normal((N, ns)), rng.standard_normal((N, ns)) + 1j * rng.standard_normal((N, ns))]
    g, atoms, pc = oracle_inputs(sysm)
    ref = O.scf(g, sysm.W, atoms, [tuple(k) for k in kpts], list(w), pc["b"], pc["rho_at"], pc["Eself"],
                pc["E_corr"], sysm.nelectron, sysm.n_atm, psi0, sysm.npl, 1e-7, 1e-9)
    gs = GroundState(device, sysm, channels=[(tuple(k), wk) for k, wk in zip(kpts, w)],
                     options=SCFOptions(SCF_tol=1e-7, poisson_tol=1e-9))
    res = gs.run(psi0=[p.T for p in psi0])
    assert res.converged and ref["converged"]
    assert gs.psi[0].dtype == torch.float64 and gs.psi[1].dtype == torch.complex128
    dE = abs(res.Eatom - ref["Eatom"])
    occ = res.occ > 1e-6
    dl = float(np.max(np.abs(res.EigVal - ref["EigVal"])[occ]))
    print(f"two-channel SCF: {res.count_SCF} iterations (oracle {ref['count_SCF']}), |dE/atom| {dE:.2e}, "
          f"max occupied |d lambda| {dl:.2e}, lambda_f {res.lambda_f:.8f} vs {ref['lambda_f']:.8f}")
    assert dE <= 1e-7
    assert dl <= 1e-6
    assert res.lambda_f == pytest.approx(ref["lambda_f"], abs=1e-6)
    gs.close()
