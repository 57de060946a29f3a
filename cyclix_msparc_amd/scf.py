"""Self-consistent field loop with every N-sized vector resident on the device.

Follows ``src/scf.m:32-306`` for nspin = 1, potential mixing
(MixingVariable = 1, the default of ``src/initialization.m:1031``) and
BC = 3 (Dirichlet in r, periodic in theta and z: no mean shift of Veff,
``src/scf.m:199-203``), first relaxation step (elecgs_Count = 1).  The host
side holds only scalars and Ns-sized arrays: eigenvalues, occupations, the
Fermi level and the energy terms, as the MATLAB host does.

Per SCF iteration on the device: one CheFSI pass per channel
(``cyx_eigsolver_pass``), electron density (``cyx_density_accumulate``), the
AAR Poisson solve (``cyx_poisson_solve``), LDA-PW Vxc and Veff
(``cyx_lda_pw``), the energy grid sums (``cyx_energy_terms``,
``cyx_scf_terms``) and Periodic-Pulay mixing (``cyx_mixer_step``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from .api import Device, EigState, Hamiltonian, Mixer, Poisson, eigSolver, lda_pw
from .system import CyclixSystem

HA_PER_EV = 1.0 / 27.211386245988    # 1 / S.Cst (src/initialization.m:947)


def scf_tol_for_energy(target, nelectron, n_atom):
    """SCF_tol from target_energy_accuracy (``src/initialization.m:516-545``)."""
    b = -0.25440 + 1.0 * math.log10(nelectron / n_atom)
    return 10.0 ** ((math.log10(target) - b) / 2.02)


def smearing_occ(ev, lam_f, bet, kind):
    """``src/occupations.m:11-15``: kind 0 Fermi-Dirac, 1 Gaussian."""
    ev = np.asarray(ev, dtype=float)
    if kind == 0:
        from scipy.special import expit  # 1 / (1 + e^x) without overflow at large beta
        return expit(-bet * (ev - lam_f))
    from scipy.special import erf
    return 0.5 * (1.0 - erf(bet * (ev - lam_f)))


def occupations(EigVal, wkpt, bet, nelectron, kind=1, occfac=2.0):
    """Fermi level and occupations (``src/occupations.m:3-36``): the root of
    occfac * sum_k w_k sum_n occ(e_nk) = Nelectron, by bracketed Brent in
    place of fzero.  EigVal: (Nev, nch)."""
    from scipy.optimize import brentq

    ev = np.asarray(EigVal, dtype=float)
    w = np.asarray(wkpt, dtype=float)

    def f(lf):
        return occfac * float(np.sum(w[None, :] * smearing_occ(ev, lf, bet, kind))) - nelectron

    lo = float(ev.min()) - 10.0 / bet - 1.0
    hi = float(ev.max()) + 10.0 / bet + 1.0
    lam_f = brentq(f, lo, hi, xtol=1e-15, rtol=4 * np.finfo(float).eps, maxiter=500)
    return lam_f, smearing_occ(ev, lam_f, bet, kind)


def band_entropy_energy(EigVal, occ, wkpt, lam_f, bet, kind=1, occfac=2.0):
    """(Eband, Eent) of ``src/evaluateTotalEnergy.m:12-20,58-74``."""
    ev = np.asarray(EigVal, dtype=float)
    w = np.asarray(wkpt, dtype=float)[None, :]
    Eband = occfac * float(np.sum(w * ev * occ))
    if kind == 0:
        with np.errstate(divide="ignore", invalid="ignore"):
            v = occfac / bet * (occ * np.log(occ) + (1 - occ) * np.log(1 - occ))
        v[np.isnan(v)] = 0.0
    else:
        v = -occfac / bet / (2 * math.sqrt(math.pi)) * np.exp(-(bet * (ev - lam_f)) ** 2)
    return Eband, float(np.sum(w * v))


@dataclass
class SCFOptions:
    """The ``S`` fields the SCF loop reads (defaults: ``src/initialization.m:1011-1038``)."""

    target_energy_accuracy: float = 1e-5   # Ha/atom; sets SCF_tol when SCF_tol < 0
    SCF_tol: float = -1.0
    poisson_tol: float = -1.0              # default 0.01 * SCF_tol (src/initialization.m:549-552)
    MAXIT_SCF: int = 100
    MINIT_SCF: int = 3
    rhoTrigger: int = 4
    MixingHistory: int = 7
    MixingParameter: float = 0.3
    MixingParameterSimple: float = -1.0    # < 0: equal to MixingParameter
    PulayFrequency: int = 1
    elec_T_type: int = 1                   # Gaussian smearing
    bet: float = -1.0                      # < 0: 1 / (0.2 eV) for Gaussian, 1 / (0.1 eV) for Fermi-Dirac
    TOL_LANCZOS: float = 1e-2
    lanczos_maxit: int = 300
    MAXIT_POISSON: int = 1000
    # AAR preconditioner of the Poisson solves: "spectral" (the exact inverse
    # of Lap_std, in the place of the reference's ILU factors S.LapPreconL/U,
    # src/poissonSolve.m:36) or "jacobi" (the diagonal of Lap_std)
    poisson_precond: str = "spectral"
    rhotol: float = 1e-14
    verbose: bool = False

    def resolved(self, nelectron, n_atom):
        o = SCFOptions(**self.__dict__)
        if o.SCF_tol < 0:
            o.SCF_tol = scf_tol_for_energy(o.target_energy_accuracy, nelectron, n_atom)
        if o.poisson_tol < 0:
            o.poisson_tol = 0.01 * o.SCF_tol
        if o.bet < 0:
            smear_ev = 0.1 if o.elec_T_type == 0 else 0.2          # src/initialization.m:265-272
            o.bet = 1.0 / (smear_ev * HA_PER_EV)
        if o.MixingParameterSimple < 0:
            o.MixingParameterSimple = o.MixingParameter
        return o


@dataclass
class SCFRecord:
    count_SCF: int
    Etotal: float
    Eband: float
    Exc: float
    Exc_dc: float
    Eelec_dc: float
    Eent: float
    Escc: float
    lambda_f: float
    err: float
    aar_iterations: int


@dataclass
class SCFResult:
    converged: bool
    count_SCF: int                 # SCF iterations done
    Etotal: float
    Eatom: float
    lambda_f: float
    EigVal: np.ndarray             # (Nev, nch)
    occ: np.ndarray
    history: list = field(default_factory=list)


class GroundState:
    """Device state of one ground-state calculation: the channels'
    Hamiltonians and psi blocks, rho, phi, Vxc, Veff, the Poisson operator
    and the mixer.  ``channels`` is a list of (kpt, weight); kpt follows
    ``S.kptgrid`` rows (0, nu, eta)."""

    def __init__(self, dev: Device, system: CyclixSystem, channels=None, nstates=None, npl=None,
                 options: SCFOptions | None = None, precision="fp64"):
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.precision = precision
        self.dev = dev
        self.system = system
        self.g = system.grid
        self.N = self.g.N
        if channels is None:
            channels = [(tuple(system.kpt), 1.0)]
        self.kpts = [tuple(float(v) for v in k) for k, _ in channels]
        self.wkpt = np.array([float(w) for _, w in channels])
        self.Nev = int(nstates or system.nstates)
        self.npl = int(npl or system.npl)
        self.nelectron = float(system.nelectron)
        self.n_atom = int(len(system.atoms))
        self.opt = (options or SCFOptions()).resolved(self.nelectron, self.n_atom)
        pc = system.pseudocharge()
        self.Eself, self.Ecorr = float(pc["Eself"]), float(pc["E_corr"])
        td = dev.torch_device
        self.b = torch.from_numpy(np.ascontiguousarray(pc["b"], dtype=float)).to(td)
        self.rho = torch.from_numpy(np.ascontiguousarray(pc["rho_at"], dtype=float)).to(td)
        if self.opt.poisson_precond not in ("spectral", "jacobi"):
            raise ValueError(f"poisson_precond {self.opt.poisson_precond!r} is not 'spectral' or 'jacobi'")
        self.poisson = Poisson(dev, self.g, spectral=self.opt.poisson_precond == "spectral")
        # initial potential (src/scf.m:32-39)
        self.phi, info = self.poisson.solve(self.rho, self.b, None, self.opt.poisson_tol, self.opt.MAXIT_POISSON)
        self.vxc, self.veff = lda_pw(dev, self.rho, self.phi, self.opt.rhotol)
        proj = system.projectors()
        self.hams = [Hamiltonian(dev, system, k, np.zeros(self.N), proj) for k in self.kpts]
        for h in self.hams:
            h.set_veff(self.veff)
        self.psi = None
        self.work = None
        self._psi0 = None
        # the state the SCF loop starts from (src/scf.m:32-39), for restart()
        self._start = tuple(t.clone() for t in (self.rho, self.phi, self.vxc, self.veff))

    def restart(self):
        """Return to the state before the first SCF iteration: the atomic
        density and its potential, and the psi0 blocks of init_psi (device
        copies; nothing crosses PCIe)."""
        if self._psi0 is None:
            raise RuntimeError("init_psi has not been called")
        for dst, src in zip((self.rho, self.phi, self.vxc, self.veff), self._start):
            dst.copy_(src)
        for p, p0 in zip(self.psi, self._psi0):
            p.copy_(p0)
        for h in self.hams:
            h.set_veff(self.veff)

    def residuals(self):
        """FP64 residual check of the current Ritz pairs: per channel, the
        relative W-norm residual of every state (``cyx_ritz_residuals``)."""
        return [h.ritz_residuals(self.psi[ks], self.EigVal[:, ks], self.work[ks]) for ks, h in enumerate(self.hams)]

    def close(self):
        """Destroy the device handles (Hamiltonians, Poisson operator)."""
        for h in self.hams:
            h.close()
        self.poisson.close()

    def init_psi(self, psi0=None, seed=0):
        """psi blocks (Nev, N) per channel.  psi0: list of host arrays, or
        None for N(0,1) columns from ``numpy.random.default_rng(seed)``
        (real for real channels, complex otherwise)."""
        td = self.dev.torch_device
        rng = np.random.default_rng(seed)
        self.psi, self.work = [], []
        for ks, h in enumerate(self.hams):
            if psi0 is not None:
                X = np.asarray(psi0[ks])
            else:
                X = rng.standard_normal((self.Nev, self.N))
                if h.complex:
                    X = X + 1j * rng.standard_normal((self.Nev, self.N))
            if self.precision == "fp64":
                dt = torch.complex128 if h.complex else torch.float64
            else:
                dt = torch.complex64 if h.complex else torch.float32
            if not h.complex and np.iscomplexobj(X):
                raise ValueError("complex psi0 for a real channel")
            p = torch.from_numpy(np.ascontiguousarray(X)).to(td).to(dt).contiguous()
            self.psi.append(p)
            self.work.append(torch.empty_like(p))
        self._psi0 = [p.clone() for p in self.psi]

    def run(self, psi0=None, seed=0) -> SCFResult:
        """``src/scf.m:88-306``: iterate to SCF_tol (at least MINIT_SCF, at
        most MAXIT_SCF SCF iterations)."""
        for _ in self.iterations(psi0, seed):
            pass
        return self.result

    def iterations(self, psi0=None, seed=0):
        """The SCF loop of ``src/scf.m:88-306`` as a generator: yields the
        SCFRecord of each SCF iteration (the first rhoTrigger - 1 passes are
        Chebyshev cycles without a density update).  ``self.result`` holds
        the SCFResult when the generator is exhausted."""
        o = self.opt
        if self.psi is None or psi0 is not None:
            self.init_psi(psi0, seed)
        nch = len(self.hams)
        S = EigState(hams=self.hams, npl=self.npl, Nev=self.Nev, EigVal=np.zeros((self.Nev, nch)),
                     rhoTrigger=o.rhoTrigger, TOL_LANCZOS=o.TOL_LANCZOS, lanczos_maxit=o.lanczos_maxit,
                     work=self.work)
        ubgv = [None] * nch
        EigVal = np.zeros((self.Nev, nch))
        a0, bup, cut = np.zeros(nch), np.zeros(nch), np.zeros(nch)
        mixer = Mixer(self.dev, self.veff, o.MixingHistory, o.MixingParameter, o.MixingParameterSimple,
                      o.PulayFrequency)
        veff_in = self.veff.clone()
        count, count_scf, err = 1, 1, 100.0
        history = []
        lam_f, occ = 0.0, None
        Etot = 0.0
        max_scf = o.MAXIT_SCF
        min_scf = min(o.MINIT_SCF, max_scf)
        self.result = None
        try:
            while (err > o.SCF_tol and count_scf <= max_scf) or count_scf <= min_scf:
                ubgv, self.psi, EigVal, a0, bup, cut = eigSolver(S, count, ubgv, self.psi, EigVal, a0, bup, cut)
                self.EigVal = EigVal
                self.bounds = (a0.copy(), bup.copy(), cut.copy())
                lam_f, occ = occupations(EigVal, self.wkpt, o.bet, self.nelectron, o.elec_T_type)
                if count >= o.rhoTrigger:
                    # electron density (src/electronDensity.m:13-17)
                    self.rho.zero_()
                    for ks, h in enumerate(self.hams):
                        h.density_accumulate(self.rho, self.psi[ks], occ[:, ks], self.wkpt[ks])
                    self.phi, info = self.poisson.solve(self.rho, self.b, self.phi, o.poisson_tol, o.MAXIT_POISSON)
                    self.vxc, veff_out = lda_pw(self.dev, self.rho, self.phi, o.rhotol)
                    h0 = self.hams[0]
                    Exc, Exc_dc, Eelec_dc = h0.energy_terms(self.rho, self.vxc, self.phi, self.b, o.rhotol)
                    Eband, Eent = band_entropy_energy(EigVal, occ, self.wkpt, lam_f, o.bet, o.elec_T_type)
                    Escc, dv2, v2 = h0.scf_terms(veff_out, veff_in, self.rho)
                    Etot = Eband + Exc - Exc_dc + Eelec_dc - self.Eself + self.Ecorr + Eent + Escc
                    err = math.sqrt(dv2) / math.sqrt(v2)
                    rec = SCFRecord(count_scf, Etot, Eband, Exc, Exc_dc, Eelec_dc, Eent, Escc, lam_f, err,
                                    int(info.iterations))
                    history.append(rec)
                    if o.verbose:
                        print(f" SCF {count_scf:3d}  Etot = {Etot:.10f}  Eatom = {Etot / self.n_atom:.10f}  "
                              f"err = {err:.4e}  AAR {info.iterations}", flush=True)
                    if err > o.SCF_tol or count_scf < o.MINIT_SCF:
                        mixer.step(veff_out, veff_in, count_scf, out=veff_out)
                        veff_in.copy_(veff_out)
                    self.veff = veff_out
                    for h in self.hams:
                        h.set_veff(self.veff)
                    count_scf += 1
                    yield rec
                count += 1
        finally:
            mixer.close()
        converged = not (count_scf == max_scf + 1 and err > o.SCF_tol)
        self.EigVal = EigVal
        self.chefsi_passes = count - 1
        self.result = SCFResult(converged=converged, count_SCF=count_scf - 1, Etotal=Etot, Eatom=Etot / self.n_atom,
                                lambda_f=lam_f, EigVal=EigVal.copy(), occ=occ, history=history)
