"""Host-side mirror of the reference's hot-path interface, over the C-ABI.

Names, argument meaning and error behaviour follow the MATLAB routines they
replace (paths relative to the reference root ``Cyclix_MS_200220/``):

* ``h_nonlocal_vector_mult``  - ``src/h_nonlocal_vector_mult.m:1``
* ``chebyshev_filter``        - ``src/chebyshev_filter.m:1``
* ``eigSolver``               - ``src/eigSolver.m:1`` (serial channel loop)
* ``poissonSolve``            - ``src/poissonSolve.m:1``
* ``aar``                     - ``src/aar.m:1``

The sparse 1-D factors (DL11..DG3), ``S.Lap_std`` and the ILU factors that the
MATLAB signatures carry are replaced by a device-resident operator built from
the grid scalars (``Hamiltonian`` / ``Poisson``); the Jacobi diagonal stands in
for ILU.  Blocks of states are torch tensors on the GPU of shape (ns, N): row
s is column s of MATLAB's N x ns ``psi`` (the same bytes, column-major).
torch is only the allocator and stream here; every operation runs in the
HIP library.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _native as nat
from .system import CyclixSystem, GridSpec, ProjAtom

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None

TORCH_DTYPES = {}
if torch is not None:
    TORCH_DTYPES = {nat.CYX_F64: torch.float64, nat.CYX_C128: torch.complex128,
                    nat.CYX_F32: torch.float32, nat.CYX_C64: torch.complex64}


def grid_desc(g: GridSpec) -> nat.GridDesc:
    return nat.GridDesc(g.nr, g.nt, g.nz, g.fdn, g.cell_typ, 0, g.xin, g.dr, g.dt, g.dz, g.twist)


class ProjectorPack:
    """``S.Atom(J).rcImage(img)`` flattened into a cyx_projector_desc (the
    numpy arrays are kept alive with the descriptor)."""

    def __init__(self, atoms: list[ProjAtom]):
        self.angnum = np.array([a.gamma.size for a in atoms], dtype=np.int32)
        self.gamma = np.ascontiguousarray(np.concatenate([a.gamma for a in atoms]) if atoms else np.zeros(0))
        img_atom, nn, disp, pos, chi = [], [], [], [], []
        for J, a in enumerate(atoms):
            for im in a.images:
                img_atom.append(J)
                nn.append(im.rc_pos.size)
                disp.append(a.coords - im.coords)
                pos.append(im.rc_pos.astype(np.int32))
                chi.append(np.asarray(im.chi, dtype=float).ravel(order="F"))
        self.image_atom = np.array(img_atom, dtype=np.int32)
        self.image_nnodes = np.array(nn, dtype=np.int32)
        self.image_disp = np.ascontiguousarray(np.array(disp, dtype=float).reshape(-1))
        self.rc_pos = np.ascontiguousarray(np.concatenate(pos) if pos else np.zeros(0, np.int32))
        self.chi = np.ascontiguousarray(np.concatenate(chi) if chi else np.zeros(0))
        self.desc = nat.ProjectorDesc(len(atoms), len(img_atom), nat.iptr(self.angnum), nat.dptr(self.gamma),
                                      nat.iptr(self.image_atom), nat.iptr(self.image_nnodes),
                                      nat.dptr(self.image_disp), nat.iptr(self.rc_pos), nat.dptr(self.chi))


_LIVE_DEVICES = weakref.WeakSet()


@atexit.register
def _close_live_devices():
    """Close every open context, and its handles in reverse creation order,
    while the HIP runtime is still up (interpreter exit runs this before the
    runtime's own teardown)."""
    for d in list(_LIVE_DEVICES):
        try:
            d.close()
        except Exception:
            pass


class Device:
    """A HIP context bound to torch's current stream on ``device``.  Handles
    made from it register themselves; ``close()`` destroys them in reverse
    creation order, then the context."""

    def __init__(self, device: int = 0):
        if torch is None or not torch.cuda.is_available():
            raise RuntimeError("cyclix_msparc_amd needs a ROCm GPU (torch.cuda is unavailable)")
        self.lib = nat.load()
        self.index = device
        h = C.c_void_p()
        nat.check(self.lib.cyx_ctx_create(device, C.byref(h)))
        self.h = h
        self._children = []
        _LIVE_DEVICES.add(self)
        self.torch_device = torch.device("cuda", device)
        stream = torch.cuda.current_stream(self.torch_device)
        nat.check(self.lib.cyx_ctx_set_stream(self.h, C.c_void_p(stream.cuda_stream)))

    def synchronize(self):
        nat.check(self.lib.cyx_ctx_synchronize(self.h))

    PHASES = ("stencil", "h_apply", "kb_gather", "gram", "rotate", "wnorm", "dense_eig", "aar", "rr_apply")

    def set_timing(self, on=True):
        nat.check(self.lib.cyx_ctx_set_timing(self.h, 1 if on else 0))

    def reset_timing(self):
        nat.check(self.lib.cyx_ctx_reset_timing(self.h))

    def timing(self):
        """{phase: (total_ms, launches)} from the device event timers."""
        out = {}
        for ph in self.PHASES:
            ms, n = C.c_double(), C.c_int64()
            nat.check(self.lib.cyx_ctx_read_timing(self.h, ph.encode(), C.byref(ms), C.byref(n)))
            out[ph] = (ms.value, n.value)
        return out

    def register(self, handle):
        """Record a handle (an object with ``close()``) made on this context."""
        self._children.append(weakref.ref(handle))

    def close(self):
        if self.h:
            self.synchronize()
            for ref in reversed(self._children):
                obj = ref()
                if obj is not None:
                    obj.close()
            self._children.clear()
            self.lib.cyx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _dtype_code(t) -> int:
    for k, v in TORCH_DTYPES.items():
        if v == t.dtype:
            return k
    raise TypeError(f"unsupported dtype {t.dtype}")


def _check_block(t, N):
    if not (t.is_cuda and t.is_contiguous() and t.dim() == 2 and t.shape[1] == N):
        raise ValueError(f"expected a contiguous CUDA block of shape (ns, {N}), got {tuple(t.shape)}")


class Hamiltonian:
    """Device operator of one Bloch channel (``blochLaplacian_1d`` +
    ``S.Veff`` + ``S.Atom`` projectors)."""

    def __init__(self, dev: Device, system: CyclixSystem, kpt, veff, projectors=None):
        self.dev = dev
        self.lib = dev.lib
        self.system = system
        self.g = system.grid
        self.N = self.g.N
        self.kpt = np.ascontiguousarray(np.asarray(kpt, dtype=float))
        self.complex = bool(self.kpt[1] != 0 or self.kpt[2] != 0)
        veff = np.ascontiguousarray(np.asarray(veff.cpu() if hasattr(veff, "cpu") else veff, dtype=float))
        W = np.ascontiguousarray(system.W, dtype=float)
        self._gd = grid_desc(self.g)
        self._proj = ProjectorPack(projectors) if projectors else None
        h = C.c_void_p()
        nat.check(self.lib.cyx_ham_create(dev.h, C.byref(self._gd), nat.dptr(self.kpt), nat.dptr(veff), nat.dptr(W),
                                          C.byref(self._proj.desc) if self._proj else None, C.byref(h)))
        self.h = h
        self.dev.register(self)

    def close(self):
        if getattr(self, "h", None):
            self.lib.cyx_ham_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_veff(self, veff):
        if hasattr(veff, "is_cuda") and veff.is_cuda:
            v = veff.contiguous().to(torch.float64)
            nat.check(self.lib.cyx_ham_set_veff(self.h, C.c_void_p(v.data_ptr()), 1))
        else:
            v = np.ascontiguousarray(np.asarray(veff, dtype=float))
            nat.check(self.lib.cyx_ham_set_veff(self.h, v.ctypes.data_as(C.c_void_p), 0))

    def set_tuning(self, tr=0, zb=0, nontemporal=False, no_xcd_remap=False, tile_major=False):
        """Launch shape of the fused stencil (results do not depend on it)."""
        t = nat.Tuning(int(tr), int(zb), (1 if nontemporal else 0) | (2 if no_xcd_remap else 0)
                       | (4 if tile_major else 0), 0)
        nat.check(self.lib.cyx_ham_set_tuning(self.h, C.byref(t)))

    def new_block(self, ns, dtype_code):
        return torch.empty((ns, self.N), dtype=TORCH_DTYPES[dtype_code], device=self.dev.torch_device)

    def apply(self, X):
        """HX = H X  (``src/h_nonlocal_vector_mult.m:21-43``)."""
        _check_block(X, self.N)
        HX = torch.empty_like(X)
        nat.check(self.lib.cyx_h_apply(self.h, _dtype_code(X), C.c_void_p(X.data_ptr()), C.c_void_p(HX.data_ptr()),
                                       X.shape[0], self.N))
        return HX

    def chebyshev_filter(self, X, m, a, b, a0, overwrite=False):
        """``src/chebyshev_filter.m:24-42``; X is preserved unless overwrite."""
        _check_block(X, self.N)
        Xw = X if overwrite else X.clone()
        Y = torch.empty_like(X)
        nat.check(self.lib.cyx_chebyshev_filter(self.h, _dtype_code(X), C.c_void_p(Xw.data_ptr()),
                                                C.c_void_p(Y.data_ptr()), X.shape[0], self.N, int(m), float(a),
                                                float(b), float(a0)))
        return Y

    def eig_pass(self, psi, npl, lambda_cutoff, bup, a0, work=None):
        """One CheFSI pass in place on psi; returns the ascending Ritz values
        (``src/eigSolver.m:98-124``)."""
        _check_block(psi, self.N)
        ns = psi.shape[0]
        if work is None:
            work = torch.empty_like(psi)
        ev = np.zeros(ns)
        nat.check(self.lib.cyx_eigsolver_pass(self.h, _dtype_code(psi), C.c_void_p(psi.data_ptr()),
                                              C.c_void_p(work.data_ptr()), ns, int(npl), float(lambda_cutoff),
                                              float(bup), float(a0), nat.dptr(ev)))
        return ev

    def rayleigh_ritz(self, psi, work=None):
        _check_block(psi, self.N)
        ns = psi.shape[0]
        if work is None:
            work = torch.empty_like(psi)
        ev = np.zeros(ns)
        nat.check(self.lib.cyx_rayleigh_ritz(self.h, _dtype_code(psi), C.c_void_p(psi.data_ptr()),
                                             C.c_void_p(work.data_ptr()), ns, nat.dptr(ev)))
        return ev

    def rotation_counts(self):
        """(FP32, FP64) matrix-core Rayleigh-Ritz rotations of FP32 blocks so far."""
        a, b = C.c_int64(), C.c_int64()
        nat.check(self.lib.cyx_ham_rotation_counts(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def ritz_residuals(self, psi, eigval, work=None):
        """||H psi_s - eigval_s psi_s||_W / ||psi_s||_W per state, H applied in
        FP64 arithmetic whatever psi's storage type (FP64 residual check of
        the pairs ``src/eigSolver.m:110-124`` returns)."""
        _check_block(psi, self.N)
        ns = psi.shape[0]
        if work is None:
            work = torch.empty_like(psi)
        ev = np.ascontiguousarray(np.asarray(eigval, dtype=float))
        if ev.size != ns:
            raise ValueError("one eigenvalue per state")
        res = np.zeros(ns)
        nat.check(self.lib.cyx_ritz_residuals(self.h, _dtype_code(psi), C.c_void_p(psi.data_ptr()),
                                              C.c_void_p(work.data_ptr()), ns, nat.dptr(ev), nat.dptr(res)))
        return res

    def density_accumulate(self, rho, psi, occ, weight=1.0):
        """rho += 2 w sum_s occ_s |psi_s|^2  (``src/electronDensity.m:13-17``)."""
        _check_block(psi, self.N)
        occ = np.ascontiguousarray(np.asarray(occ, dtype=float))
        nat.check(self.lib.cyx_density_accumulate(self.h, _dtype_code(psi), C.c_void_p(psi.data_ptr()), psi.shape[0],
                                                  nat.dptr(occ), float(weight), C.c_void_p(rho.data_ptr())))
        return rho

    def density_response(self, drho, psi, dpsi, occ, dg=None, weight=1.0):
        """drho += 2 w sum_s (dg_s |psi_s|^2 + 2 occ_s Re(conj(psi_s) dpsi_s))
        (``src/phonon.m:1232-1240``, q = 0) over the first len(occ) states."""
        occ = np.ascontiguousarray(np.asarray(occ, dtype=float))
        ns = occ.size
        _check_block(psi, self.N)
        _check_block(dpsi, self.N)
        if psi.dtype != dpsi.dtype or psi.shape[0] < ns or dpsi.shape[0] < ns:
            raise ValueError("psi and dpsi must share a dtype and hold len(occ) states")
        if drho.dtype != torch.float64 or drho.numel() != self.N or not drho.is_contiguous():
            raise ValueError("drho must be a contiguous float64 vector of length N")
        dgp = None
        if dg is not None:
            dg = np.ascontiguousarray(np.asarray(dg, dtype=float))
            if dg.size != ns:
                raise ValueError("dg must have len(occ) entries")
            dgp = nat.dptr(dg)
        nat.check(self.lib.cyx_density_response(self.h, _dtype_code(psi), C.c_void_p(psi.data_ptr()),
                                                C.c_void_p(dpsi.data_ptr()), ns, nat.dptr(occ), dgp, float(weight),
                                                C.c_void_p(drho.data_ptr())))
        return drho

    def energy_terms(self, rho, vxc, phi, b, rhotol=1e-14):
        """(Exc, Exc_dc, Eelec_dc) grid sums (``src/evaluateTotalEnergy.m:27-52,72``)."""
        out = np.zeros(3)
        nat.check(self.lib.cyx_energy_terms(self.h, C.c_void_p(rho.data_ptr()), C.c_void_p(vxc.data_ptr()),
                                            C.c_void_p(phi.data_ptr()), C.c_void_p(b.data_ptr()), float(rhotol),
                                            nat.dptr(out)))
        return tuple(out)

    def scf_terms(self, vout, vin, rho):
        """(Escc, ||vout - vin||^2, ||vout||^2) for potential mixing
        (``src/scf.m:160-162,182-184``)."""
        out = np.zeros(3)
        nat.check(self.lib.cyx_scf_terms(self.h, C.c_void_p(vout.data_ptr()), C.c_void_p(vin.data_ptr()),
                                         C.c_void_p(rho.data_ptr()), nat.dptr(out)))
        return tuple(out)

    def sternheimer_apply(self, X, lam, psi_v=None):
        """Y_j = (H - lam_j) X_j + psi_v psi_v^H (W .* X_j)  (``src/phonon.m:2139-2157``);
        X: (ncol, N) block, psi_v: (nv, N) block of the same dtype or None."""
        _check_block(X, self.N)
        lam = np.ascontiguousarray(np.asarray(lam, dtype=float))
        if lam.size != X.shape[0]:
            raise ValueError("one shift per column")
        nv = 0 if psi_v is None else psi_v.shape[0]
        if psi_v is not None:
            _check_block(psi_v, self.N)
            if psi_v.dtype != X.dtype:
                raise ValueError("psi_v and X must share a dtype")
        Y = torch.empty_like(X)
        nat.check(self.lib.cyx_sternheimer_apply(self.h, _dtype_code(X), C.c_void_p(X.data_ptr()),
                                                 C.c_void_p(Y.data_ptr()), X.shape[0], self.N, nat.dptr(lam),
                                                 C.c_void_p(psi_v.data_ptr()) if nv else None, self.N, nv))
        return Y

    def sternheimer_solve(self, B, lam, psi_v=None, X0=None, tol=1e-8, max_iter=3000, precond=True,
                          raise_on_noconv=False):
        """Batched BiCGSTAB on the Sternheimer systems, one per column
        (``src/phonon.m:1516-1521``).  Returns (X, info, iterations, relres)."""
        _check_block(B, self.N)
        lam = np.ascontiguousarray(np.asarray(lam, dtype=float))
        if lam.size != B.shape[0]:
            raise ValueError("one shift per column")
        nv = 0 if psi_v is None else psi_v.shape[0]
        if psi_v is not None:
            _check_block(psi_v, self.N)
            if psi_v.dtype != B.dtype:
                raise ValueError("psi_v and B must share a dtype")
        X = X0.clone().contiguous() if X0 is not None else torch.empty_like(B)
        info = nat.SternInfo()
        its = np.zeros(B.shape[0], dtype=np.int32)
        rel = np.zeros(B.shape[0])
        allow = () if raise_on_noconv else (nat.CYX_ERR_NOCONV,)
        nat.check(self.lib.cyx_sternheimer_solve(self.h, _dtype_code(B), C.c_void_p(B.data_ptr()),
                                                 C.c_void_p(X.data_ptr()), B.shape[0], self.N, nat.dptr(lam),
                                                 C.c_void_p(psi_v.data_ptr()) if nv else None, self.N, nv,
                                                 1 if X0 is not None else 0, float(tol), int(max_iter),
                                                 precond_code(precond), C.byref(info), nat.iptr(its),
                                                 nat.dptr(rel)), allow=allow)
        return X, info, its, rel

    def sternheimer_rhs(self, psi, nv, coeff, out=None):
        """Right-hand sides of one perturbation (``src/phonon.m:1358-1444``) on
        this handle as the first-order operator: its potential is dVeff and its
        projectors are the pairs (chi +- dchi)/sqrt(2) with gammas +-gamma
        (``dfpt.perturbation_projectors``).  psi: (>= nv, N) block of the
        channel's states; coeff: nv x nv (``dfpt.coeff_pv_psi``).  Returns the
        (nv, N) block B and dlambda (S.Dlambda_dR) on the host."""
        _check_block(psi, self.N)
        nv = int(nv)
        if not 1 <= nv <= psi.shape[0]:
            raise ValueError("bad number of occupied states")
        cf = np.asfortranarray(np.asarray(coeff, dtype=float))
        if cf.shape != (nv, nv):
            raise ValueError("coeff must be nv x nv")
        B = torch.empty((nv, self.N), dtype=psi.dtype, device=psi.device) if out is None else out
        _check_block(B, self.N)
        if B.dtype != psi.dtype or B.shape[0] < nv:
            raise ValueError("out must share psi's dtype and hold nv columns")
        dl = np.zeros(nv)
        cfc = np.ascontiguousarray(cf.ravel(order="F"))
        nat.check(self.lib.cyx_sternheimer_rhs(self.h, _dtype_code(psi), C.c_void_p(psi.data_ptr()), self.N, nv,
                                               nat.dptr(cfc), C.c_void_p(B.data_ptr()), self.N, nat.dptr(dl)))
        return B, dl

    def weighted_gram(self, A, B):
        """A^H diag(W) B (FP64 accumulation) of two (n, N) device blocks of
        one dtype: the reference's PsiW * X products.  Returns an (na, nb)
        complex128 or float64 device tensor."""
        _check_block(A, self.N)
        _check_block(B, self.N)
        if A.dtype != B.dtype:
            raise ValueError("A and B must share a dtype")
        out_dt = torch.complex128 if A.is_complex() else torch.float64
        M = torch.empty((B.shape[0], A.shape[0]), dtype=out_dt, device=A.device)  # column-major na x nb
        nat.check(self.lib.cyx_weighted_gram(self.h, _dtype_code(A), C.c_void_p(A.data_ptr()), self.N, A.shape[0],
                                             C.c_void_p(B.data_ptr()), self.N, B.shape[0], C.c_void_p(M.data_ptr())))
        return M.T

    def lanczos_bounds(self, v0, maxit=300, tol=1e-3):
        v0 = np.ascontiguousarray(np.asarray(v0, dtype=float))
        lo, hi = C.c_double(), C.c_double()
        nat.check(self.lib.cyx_lanczos_bounds(self.h, nat.dptr(v0), int(maxit), float(tol), C.byref(lo), C.byref(hi)))
        return lo.value, hi.value


class SternheimerPlan:
    """Batched BiCGSTAB on the Sternheimer systems with its iterations
    captured in a HIP graph (cyx_sternheimer_plan_*): converged columns
    freeze on the device, the host only reads the active count between
    replays.  B: (ncol, N) right-hand sides; lam: one shift per column; psi_v:
    (nv, N) occupied states or None.  The solution is ``self.X``."""

    def __init__(self, ham: Hamiltonian, B, lam, psi_v=None, X0=None, tol=1e-8, precond=True):
        self.dev = ham.dev
        self.lib = ham.lib
        self.ham = ham
        _check_block(B, ham.N)
        lam = np.ascontiguousarray(np.asarray(lam, dtype=float))
        if lam.size != B.shape[0]:
            raise ValueError("one shift per column")
        nv = 0 if psi_v is None else psi_v.shape[0]
        if psi_v is not None:
            _check_block(psi_v, ham.N)
            if psi_v.dtype != B.dtype:
                raise ValueError("psi_v and B must share a dtype")
        # the captured graph holds these addresses: keep the tensors alive
        self.B, self.psi_v, self.lam = B, psi_v, lam
        self.X = X0.clone().contiguous() if X0 is not None else torch.empty_like(B)
        self.ncol = B.shape[0]
        self.tol = float(tol)
        h = C.c_void_p()
        nat.check(self.lib.cyx_sternheimer_plan_create(ham.h, _dtype_code(B), C.c_void_p(B.data_ptr()),
                                                       C.c_void_p(self.X.data_ptr()), self.ncol, ham.N,
                                                       nat.dptr(lam), C.c_void_p(psi_v.data_ptr()) if nv else None,
                                                       ham.N, nv, 1 if X0 is not None else 0, self.tol,
                                                       precond_code(precond), C.byref(h)))
        self.h = h
        self.dev.register(self)

    def iterate(self, iters=4):
        nat.check(self.lib.cyx_sternheimer_plan_iterate(self.h, int(iters)))

    def status(self):
        """(active columns, SternInfo, per-column iterations, relative residuals)."""
        act = C.c_int32()
        info = nat.SternInfo()
        its = np.zeros(self.ncol, dtype=np.int32)
        rel = np.zeros(self.ncol)
        nat.check(self.lib.cyx_sternheimer_plan_status(self.h, C.byref(act), C.byref(info), nat.iptr(its),
                                                       nat.dptr(rel)))
        return act.value, info, its, rel

    def solve(self, max_iter=3000, poll=8):
        """Replay until every column meets tol or max_iter iterations ran."""
        done = 0
        act = self.status()[0]
        while act > 0 and done < max_iter:
            k = min(poll, max_iter - done)
            self.iterate(k)
            done += k
            act = self.status()[0]
        return self.status()

    def close(self):
        if getattr(self, "h", None):
            self.lib.cyx_sternheimer_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FilterPlan:
    """Batched Chebyshev filter over several channels: one launch per
    recurrence step for all channels, the pass captured in a HIP graph
    (cyx_filter_plan_*).  blocks: list of (ncol, N) device tensors (one per
    channel, same dtype); each is filtered in place (the result is copied
    back into it when the recurrence ends in the work block)."""

    def __init__(self, dev: Device, hams, blocks, m, autotune=False, bounds=None):
        self.dev = dev
        self.lib = dev.lib
        self.hams = list(hams)
        self.blocks = list(blocks)
        if len(self.hams) != len(self.blocks) or not self.hams:
            raise ValueError("one block per Hamiltonian")
        dt = self.blocks[0].dtype
        for h, X in zip(self.hams, self.blocks):
            _check_block(X, h.N)
            if X.dtype != dt:
                raise ValueError("all blocks must share a dtype")
        self.dtype_code = _dtype_code(self.blocks[0])
        self.work = [torch.empty_like(X) for X in self.blocks]
        n = len(self.hams)
        hv = (C.c_void_p * n)(*[h.h.value for h in self.hams])
        av = (C.c_void_p * n)(*[X.data_ptr() for X in self.blocks])
        bv = (C.c_void_p * n)(*[W.data_ptr() for W in self.work])
        nc = (C.c_int32 * n)(*[X.shape[0] for X in self.blocks])
        ld = (C.c_int64 * n)(*[h.N for h in self.hams])
        a = b = a0 = None
        if autotune:
            a, b, a0 = (np.ascontiguousarray(np.asarray(v, dtype=float)) for v in bounds)
        h = C.c_void_p()
        nat.check(self.lib.cyx_filter_plan_create(self.dev.h, n, hv, self.dtype_code, av, bv, nc, ld, int(m),
                                                  1 if autotune else 0, nat.dptr(a) if a is not None else None,
                                                  nat.dptr(b) if b is not None else None,
                                                  nat.dptr(a0) if a0 is not None else None, C.byref(h)))
        self.h = h
        self.dev.register(self)
        self.m = int(m)

    def tuning(self):
        t = nat.Tuning()
        nat.check(self.lib.cyx_filter_plan_tuning(self.h, C.byref(t)))
        return {"tr": t.tr, "flags": t.flags}

    def attach_sternheimer(self, sp, iters=4):
        """Run ``iters`` iterations of the Sternheimer plan ``sp`` in every
        pass's graph launch, as an independent branch (None detaches)."""
        nat.check(self.lib.cyx_filter_plan_attach_sternheimer(self.h, sp.h if sp is not None else None,
                                                              int(iters)))
        self.stern = sp

    def run(self, a, b, a0):
        """Filter every block in place with per-channel bounds."""
        a, b, a0 = (np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=float), (len(self.hams),)))
                    for v in (a, b, a0))
        res = C.c_int32()
        nat.check(self.lib.cyx_filter_plan_run(self.h, nat.dptr(a), nat.dptr(b), nat.dptr(a0), C.byref(res)))
        if res.value:
            for X, W in zip(self.blocks, self.work):
                X.copy_(W)
        return self.blocks

    def close(self):
        if getattr(self, "h", None):
            self.lib.cyx_filter_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lda_pw(dev: Device, rho, phi=None, rhotol=1e-14):
    """(vxc, veff): LDA-PW potential (``src/exchangeCorrelationPotential.m:82-106``)
    and, with phi, Veff = phi + Vxc (``src/scf.m:39``); device FP64 vectors."""
    vxc = torch.empty_like(rho)
    veff = torch.empty_like(rho) if phi is not None else None
    nat.check(dev.lib.cyx_lda_pw(dev.h, rho.numel(), C.c_void_p(rho.data_ptr()),
                                 C.c_void_p(phi.data_ptr()) if phi is not None else None,
                                 C.c_void_p(vxc.data_ptr()), C.c_void_p(veff.data_ptr()) if veff is not None else None,
                                 float(rhotol)))
    return vxc, veff


def lda_pw_fxc(dev: Device, rho, drho=None, dphi=None, rhotol=1e-14, fxc=True):
    """(fxc, dveff): the LDA-PW exchange-correlation kernel dVxc/drho at rho
    (S.Dvxc_drho, ``src/phonon.m:215``) and, with drho and dphi,
    DVeff = Dphi + fxc Drho (``src/phonon.m:303-306``); device FP64 vectors."""
    if rho.dtype != torch.float64 or not rho.is_contiguous():
        raise ValueError("rho must be a contiguous float64 vector")
    for t in (drho, dphi):
        if t is not None and (t.dtype != torch.float64 or t.numel() != rho.numel() or not t.is_contiguous()):
            raise ValueError("drho and dphi must be contiguous float64 vectors of rho's length")
    want_dveff = drho is not None and dphi is not None
    if not (fxc or want_dveff):
        raise ValueError("nothing to compute")
    f = torch.empty_like(rho) if fxc else None
    dveff = torch.empty_like(rho) if want_dveff else None
    ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    nat.check(dev.lib.cyx_lda_pw_fxc(dev.h, rho.numel(), ptr(rho), ptr(drho), ptr(dphi), ptr(f), ptr(dveff),
                                     float(rhotol)))
    return f, dveff


class Mixer:
    """Periodic-Pulay mixing state (``src/mixing.m:20-116``) for FP64 device
    vectors of length N; ``x0`` is the initial input iterate
    (``src/scf.m:42-56``)."""

    def __init__(self, dev: Device, x0, history=7, beta=0.3, omega=None, pulay_frequency=1):
        self.dev = dev
        self.lib = dev.lib
        x0 = x0.contiguous()
        if x0.dtype != torch.float64 or not x0.is_cuda:
            raise ValueError("x0 must be a float64 device tensor")
        self.N = x0.numel()
        h = C.c_void_p()
        omega = beta if omega is None or omega < 0 else omega   # src/initialization.m:573-575
        nat.check(self.lib.cyx_mixer_create(dev.h, self.N, int(history), float(beta), float(omega),
                                            int(pulay_frequency), C.c_void_p(x0.data_ptr()), C.byref(h)))
        self.h = h
        self.dev.register(self)

    def close(self):
        if getattr(self, "h", None):
            self.lib.cyx_mixer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, g_k, x_k, it, out=None):
        """x_{k+1} = mixing(S, g_k, x_k, iter) (``src/mixing.m:20``)."""
        for t in (g_k, x_k):
            if t.dtype != torch.float64 or t.numel() != self.N or not t.is_contiguous():
                raise ValueError("mixing operands must be contiguous float64 of length N")
        out = torch.empty_like(x_k) if out is None else out
        if out.data_ptr() == x_k.data_ptr():
            raise ValueError("x_out may not alias x_k")
        nat.check(self.lib.cyx_mixer_step(self.h, C.c_void_p(g_k.data_ptr()), C.c_void_p(x_k.data_ptr()),
                                          C.c_void_p(out.data_ptr()), int(it)))
        return out


class Poisson:
    """``S.Lap_std`` (k = 0) on the device, with the Jacobi diagonal or, with
    ``spectral=True``, the exact inverse of Lap_std (theta/z FFT and one band
    solve per mode) as the AAR preconditioner, in the place of the
    reference's ILU factors (``src/poissonSolve.m:36``)."""

    def __init__(self, dev: Device, grid: GridSpec, spectral=False):
        self.dev = dev
        self.lib = dev.lib
        self.N = grid.N
        self._gd = grid_desc(grid)
        h = C.c_void_p()
        nat.check(self.lib.cyx_poisson_create(dev.h, C.byref(self._gd), C.byref(h)))
        self.h = h
        self.dev.register(self)
        self.spectral = False
        if spectral:
            self.set_spectral(True)

    def set_spectral(self, on=True):
        nat.check(self.lib.cyx_poisson_set_spectral(self.h, 1 if on else 0))
        self.spectral = bool(on)

    def spectral_apply(self, r):
        """z = Lap_std^-1 r exactly (FP64, (N,) or (ncol, N))."""
        r = r.contiguous()
        z = torch.empty_like(r)
        ncol = 1 if r.dim() == 1 else r.shape[0]
        nat.check(self.lib.cyx_poisson_spectral_apply(self.h, C.c_void_p(r.data_ptr()), C.c_void_p(z.data_ptr()),
                                                      ncol, self.N))
        return z

    def close(self):
        if getattr(self, "h", None):
            self.lib.cyx_poisson_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply(self, x):
        """y = Lap_std x."""
        x = x.contiguous()
        y = torch.empty_like(x)
        nat.check(self.lib.cyx_poisson_apply(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr())))
        return y

    def aar(self, f, x_guess=None, tol=1e-8, max_iter=1000, omega=0.6, beta=0.6, m=7, p=6):
        """``src/aar.m:33-72``; returns (x, info).  Raises on non-convergence
        only through ``info.converged`` (the reference prints and returns)."""
        f = f.contiguous()
        x = x_guess.clone().contiguous() if x_guess is not None else torch.empty_like(f)
        info = nat.AarInfo()
        nat.check(self.lib.cyx_aar(self.h, C.c_void_p(f.data_ptr()), C.c_void_p(x.data_ptr()),
                                   1 if x_guess is not None else 0, float(tol), int(max_iter), float(omega),
                                   float(beta), int(m), int(p), C.byref(info)), allow=(nat.CYX_ERR_NOCONV,))
        return x, info

    def solve(self, rho, b, phi_guess=None, tol=1e-8, max_iter=1000):
        """``src/poissonSolve.m:17-36`` (Cyclix branch)."""
        phi = phi_guess.clone().contiguous() if phi_guess is not None else torch.empty_like(rho)
        info = nat.AarInfo()
        nat.check(self.lib.cyx_poisson_solve(self.h, C.c_void_p(rho.contiguous().data_ptr()),
                                             C.c_void_p(b.contiguous().data_ptr()), C.c_void_p(phi.data_ptr()),
                                             1 if phi_guess is not None else 0, float(tol), int(max_iter),
                                             C.byref(info)), allow=(nat.CYX_ERR_NOCONV,))
        return phi, info


def precond_code(precond):
    """The Sternheimer preconditioner argument of the C-ABI: "spectral" (2,
    the exact Lap_std^-1 in the place of the reference's ILU factors),
    "jacobi" or True (1), "none" or False (0)."""
    if precond is True or precond == "jacobi":
        return 1
    if precond is False or precond is None or precond == "none":
        return 0
    if precond == "spectral":
        return 2
    raise ValueError(f"precond {precond!r}: 'spectral', 'jacobi' or 'none'")


# --------------------------------------------------------------------------
# reference-named entry points
# --------------------------------------------------------------------------
def h_nonlocal_vector_mult(ham: Hamiltonian, X):
    """Hnlx = h_nonlocal_vector_mult(DL11,DL22,DL33,DG1,DG2,DG3,Veff,X,S,kptvec);
    the channel operator (1-D factors, Veff, S, kptvec) is ``ham``."""
    return ham.apply(X)


def chebyshev_filter(X, m, a, b, a0, ham: Hamiltonian):
    """Y = chebyshev_filter(X,m,a,b,a0,DL11,DL22,DL33,DG1,DG2,DG3,Veff,S,kpt_vec)."""
    return ham.chebyshev_filter(X, m, a, b, a0)


@dataclass
class EigState:
    """The fields of ``S`` that eigSolver reads and writes."""

    hams: list                    # one Hamiltonian per channel (S.kptgrid rows)
    npl: int                      # S.npl
    Nev: int                      # S.Nev
    EigVal: np.ndarray = None     # (Nev, nch), S.EigVal
    elecgs_Count: int = 1
    rhoTrigger: int = 4
    chefsibound_flag: int = 0
    TOL_LANCZOS: float = 1e-2
    lanczos_maxit: int = 300
    work: list = field(default_factory=list)


def eigSolver(S: EigState, count, upper_bound_guess_vecs, psi, EigVal, a0, bup, lambda_cutoff):
    """[ubgv,psi,EigVal,a0,bup,lambda_cutoff] = eigSolver(S,count,...)
    ``src/eigSolver.m:26-124`` (serial branch).  psi: list of device blocks
    (one per channel), updated in place.  The spectral bounds use the
    library's Lanczos estimator (``cyx_lanczos_bounds``) in place of eigs."""
    if S.elecgs_Count > 1:
        S.rhoTrigger = 1
    nch = len(S.hams)
    for ks in range(nch):
        ham = S.hams[ks]
        if count == 1:
            v0 = np.random.default_rng(ks + 1).random(ham.N)
            lo, hi = ham.lanczos_bounds(v0, S.lanczos_maxit, S.TOL_LANCZOS)
            bup[ks] = 1.01 * hi
            upper_bound_guess_vecs[ks] = None
            a0[ks] = (lo - 0.1) if S.elecgs_Count == 1 else np.min(S.EigVal[:, ks])
        elif count > S.rhoTrigger:
            if S.chefsibound_flag == 1:
                v0 = np.random.default_rng(ks + 1).random(ham.N)
                _, hi = ham.lanczos_bounds(v0, S.lanczos_maxit, S.TOL_LANCZOS)
                bup[ks] = 1.01 * hi
            a0[ks] = np.min(S.EigVal[:, ks])
        if count == 1 and S.elecgs_Count == 1:
            lambda_cutoff[ks] = 0.5 * (bup[ks] + a0[ks])
        else:
            lambda_cutoff[ks] = np.max(S.EigVal[:, ks]) + 0.10
        work = S.work[ks] if ks < len(S.work) else None
        EigVal[:, ks] = ham.eig_pass(psi[ks], S.npl, lambda_cutoff[ks], bup[ks], a0[ks], work=work)
    S.EigVal = EigVal
    return upper_bound_guess_vecs, psi, EigVal, a0, bup, lambda_cutoff


def aar(A: Poisson, b, x_guess, tol, max_iter, omega, beta, m, p, L=None, U=None):
    """x = aar(A,b,x_guess,tol,max_iter,omega,beta,m,p,L,U); ``A`` is the
    device Poisson operator, the preconditioner is its Jacobi diagonal (L, U
    are accepted and ignored).  Prints the reference's convergence message."""
    x, info = A.aar(b, x_guess, tol, max_iter, omega, beta, m, p)
    if not info.converged:
        print(f" AAR exceeded maximum iterations and converged to a relative residual of {info.relres:g}. ")
    return x


def poissonSolve(S, poisson_tol, Isguess):
    """S = poissonSolve(S,poisson_tol,Isguess) for Cyclix: S.phi from
    f = -4 pi (S.rho + S.b) (``src/poissonSolve.m:19-36``).  ``S`` carries
    ``poisson`` (a Poisson operator), ``rho``, ``b``, ``phi`` (device
    vectors) and ``MAXIT_POISSON``."""
    guess = S.phi if Isguess else None
    S.phi, S.poisson_info = S.poisson.solve(S.rho, S.b, guess, poisson_tol, getattr(S, "MAXIT_POISSON", 1000))
    return S
