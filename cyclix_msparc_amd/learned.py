"""Learned part of the AAR Poisson preconditioner.

A linear three-level 3-D convolutional U-Net U (cyx_unet.hip) approximates
Lap_std^{-1}: d = D^{-1} r + U(D^{-1} r) (D the Jacobi diagonal) is the
correction for a residual r.  Inside AAR it is applied after every Anderson
iteration as x += delta d with the exact line search delta = <Lap d, r> /
<Lap d, Lap d>, which can only lower the residual norm.  The plain
Jacobi-AAR iteration stays the fallback and the judge: every solve stops on
the true relative residual ||f - Lap_std x|| / ||f|| of src/aar.m.

Training data come from the method's own solves: random densities of 4-12
Gaussians (sigma 0.5-2 Bohr, total charge 32 e), f = -4 pi rho, solved by
Jacobi-AAR; each solve gives the pair (f, phi) and, from the same solve
stopped early, residual/error pairs (r_K, e_K) with Lap_std e_K = r_K (see
generate_pairs).  The network is fitted to e - D^{-1} r from the input
D^{-1} r (each pair scaled by 1 / rms(D^{-1} r), which a linear map does not
notice) with the device's own forward and backward kernels and Adam.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from pathlib import Path

import numpy as np
import torch

from . import _native as nat
from .api import Device, Poisson, grid_desc
from .system import GridSpec, to_cart

# (mode, cin, cout) of e0 d1 e1 d2 e2 m2 u1 f1 u0 f0; mode 0 conv3, 1 down, 2 up
SPEC = [(0, 1, 16), (1, 16, 32), (0, 32, 32), (1, 32, 64), (0, 64, 64),
        (0, 64, 64), (2, 64, 32), (0, 32, 32), (2, 32, 16), (0, 16, 1)]
WEIGHTS_DIR = Path(__file__).resolve().parent / "weights"


def _ntap(mode):
    return 27 if mode == 0 else 8


def param_offsets(nr):
    off = [0]
    for mode, ci, co in SPEC:
        off.append(off[-1] + _ntap(mode) * ci * co)
    return off, off[-1] + nr


class UNet:
    """Device U-Net of one grid (cyx_unet_*)."""

    def __init__(self, dev: Device, grid: GridSpec):
        self.dev = dev
        self.lib = dev.lib
        self.grid = grid
        self.N = grid.N
        self._gd = grid_desc(grid)
        h = C.c_void_p()
        nat.check(self.lib.cyx_unet_create(dev.h, C.byref(self._gd), C.byref(h)))
        self.h = h
        dev.register(self)
        self.nparam = int(self.lib.cyx_unet_num_params(h))
        self.attached = []

    def close(self):
        if getattr(self, "h", None):
            for ps in self.attached:
                if getattr(ps, "h", None):
                    self.lib.cyx_poisson_set_preconditioner(ps.h, None)
            self.lib.cyx_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters
    def init_params(self, seed=0):
        """Variance-preserving random weights for every layer but the last,
        which starts at zero (U = 0: M is the Jacobi preconditioner), and a
        unit radial scale."""
        off, n = param_offsets(self.grid.nr)
        rng = np.random.default_rng(seed)
        p = np.zeros(n, dtype=np.float32)
        for l, (mode, ci, co) in enumerate(SPEC[:-1]):
            fan_in = _ntap(mode) * ci if mode != 2 else ci
            p[off[l]:off[l + 1]] = rng.standard_normal(off[l + 1] - off[l]) / math.sqrt(fan_in)
        p[off[-1]:] = 1.0
        self.set_params(p, reset_optimizer=True)
        return p

    def set_params(self, p, reset_optimizer=False):
        p = np.ascontiguousarray(np.asarray(p, dtype=np.float32))
        if p.size != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters, got {p.size}")
        nat.check(self.lib.cyx_unet_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float)),
                                               1 if reset_optimizer else 0))

    def get_params(self):
        p = np.zeros(self.nparam, dtype=np.float32)
        nat.check(self.lib.cyx_unet_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    def save(self, path):
        g = self.grid
        np.savez(path, params=self.get_params(), dims=np.array([g.nr, g.nt, g.nz]))

    def load(self, path):
        d = np.load(path, allow_pickle=False)
        g = self.grid
        if tuple(int(v) for v in d["dims"]) != (g.nr, g.nt, g.nz):
            raise ValueError(f"{path} holds weights for grid {tuple(d['dims'])}, not {(g.nr, g.nt, g.nz)}")
        self.set_params(d["params"], reset_optimizer=True)

    # ---- compute
    def _check(self, t, dtype):
        if not (t.is_cuda and t.is_contiguous() and t.dtype == dtype and t.dim() == 2 and t.shape[1] == self.N):
            raise ValueError(f"expected a contiguous (nb, {self.N}) {dtype} device tensor")

    def forward(self, x):
        self._check(x, torch.float32)
        u = torch.empty_like(x)
        nat.check(self.lib.cyx_unet_forward(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(u.data_ptr()), x.shape[0]))
        return u

    def gradient(self, x, t):
        self._check(x, torch.float32)
        self._check(t, torch.float32)
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = C.c_double()
        nat.check(self.lib.cyx_unet_gradient(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()), x.shape[0],
                                             g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)))
        return loss.value, g

    def train_step(self, x, t, lr):
        loss = C.c_double()
        nat.check(self.lib.cyx_unet_train_step(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                               x.shape[0], float(lr), C.byref(loss)))
        return loss.value

    def input_of(self, poisson: Poisson, r, scale=None):
        """x_b = scale_b r_b / D for FP64 residuals r (nb, N)."""
        self._check(r, torch.float64)
        x = torch.empty(r.shape, dtype=torch.float32, device=r.device)
        sc = None
        if scale is not None:
            sc = scale.to(torch.float32).contiguous()
        nat.check(self.lib.cyx_unet_input(self.h, poisson.h, C.c_void_p(r.data_ptr()),
                                          C.c_void_p(sc.data_ptr()) if sc is not None else None, r.shape[0],
                                          C.c_void_p(x.data_ptr())))
        return x

    def attach(self, poisson: Poisson):
        """Use U in every later solve of this Poisson handle."""
        nat.check(self.lib.cyx_poisson_set_preconditioner(poisson.h, self.h))
        self.attached.append(poisson)

    def detach(self, poisson: Poisson):
        nat.check(self.lib.cyx_poisson_set_preconditioner(poisson.h, None))
        self.attached = [p for p in self.attached if p is not poisson]


# --------------------------------------------------------------------------
# training data
# --------------------------------------------------------------------------
class DensitySampler:
    """Random densities of 4-12 Gaussians (sigma 0.5-2 Bohr) centred at
    random points of the domain, with periodic images in theta and z, scaled
    to a total charge; evaluated on the device."""

    def __init__(self, system, device, total_charge=32.0):
        g = system.grid
        self.g = g
        r = np.tile(g.r, g.nt * g.nz)
        th = np.repeat(np.tile(np.arange(g.nt) * g.dt, g.nz), g.nr)
        z = np.repeat(np.arange(g.nz) * g.dz, g.nr * g.nt)
        self.xp = torch.from_numpy(to_cart(g, np.stack([r, th, z], 1))).to(device)
        self.W = torch.from_numpy(np.ascontiguousarray(system.W, dtype=float)).to(device)
        self.q = total_charge

    def sample(self, rng):
        g = self.g
        L2, L3 = g.nt * g.dt, g.nz * g.dz
        ng = int(rng.integers(4, 13))
        rho = torch.zeros(g.N, dtype=torch.float64, device=self.xp.device)
        r_lo, r_hi = g.xin + 1.0, g.xin + (g.nr - 1) * g.dr - 1.0
        for _ in range(ng):
            c = (rng.uniform(r_lo, r_hi), rng.uniform(0, L2), rng.uniform(0, L3))
            sig = rng.uniform(0.5, 2.0)
            amp = rng.uniform(0.5, 1.5)
            for dth in (-L2, 0.0, L2):
                for dz in (-L3, 0.0, L3):
                    xc = torch.from_numpy(to_cart(g, [c[0], c[1] + dth, c[2] + dz])[0]).to(rho.device)
                    d2 = ((self.xp - xc) ** 2).sum(1)
                    rho += amp * torch.exp(-d2 / (2 * sig * sig))
        return rho * (self.q / float((rho * self.W).sum()))


def generate_pairs(dev: Device, system, n_densities, seed=0, tol=1e-6, max_iter=8000, snapshots=(),
                   noise_fraction=0.0, log=None):
    """Residual/error pairs from the method's own Jacobi-AAR solves.

    For each random density (f = -4 pi rho) the solve to tol gives x*; AAR
    runs stopped after K iterations (K in snapshots, cold start) give x_K,
    and each (r_K, e_K) = (f - Lap_std x_K, x* - x_K) is a pair, as is
    (f, x*).  These are the residuals the preconditioner sees inside AAR,
    with the corrections that would solve them exactly.  A fraction of the
    right-hand sides are white noise (pair (f, x*) only), so that the rough
    end of the spectrum is represented too.  Pairs are stored FP32 (their
    rounding is far below tol = 1e-6).  Returns (R, E) as (n, N) float32
    device tensors and the Jacobi iteration counts of the full solves."""
    g = system.grid
    td = dev.torch_device
    ps = Poisson(dev, g)
    rng = np.random.default_rng(seed)
    sampler = DensitySampler(system, td)
    gen = torch.Generator(device=td)
    gen.manual_seed(seed)
    is_noise = np.zeros(n_densities, dtype=bool)
    is_noise[rng.permutation(n_densities)[:int(round(noise_fraction * n_densities))]] = True
    npairs = int(np.sum(np.where(is_noise, 1, 1 + len(snapshots))))
    R = torch.empty((npairs, g.N), dtype=torch.float32, device=td)
    E = torch.empty_like(R)
    its = np.zeros(n_densities, dtype=int)
    t0 = time.time()
    row = 0
    for k in range(n_densities):
        if is_noise[k]:
            f = torch.randn(g.N, dtype=torch.float64, device=td, generator=gen).contiguous()
        else:
            f = (-4 * np.pi * sampler.sample(rng)).contiguous()
        xs, info = ps.aar(f, None, tol, max_iter)
        if not info.converged:
            raise RuntimeError(f"training solve {k} did not converge (relres {info.relres:.3e})")
        its[k] = info.iterations
        R[row] = f.to(torch.float32)
        E[row] = xs.to(torch.float32)
        row += 1
        if not is_noise[k]:
            for K in snapshots:
                xk, _ = ps.aar(f, None, 1e-300, K)
                R[row] = (f - ps.apply(xk)).to(torch.float32)
                E[row] = (xs - xk).to(torch.float32)
                row += 1
        if log and (k + 1) % 128 == 0:
            log(f"densities {k + 1}/{n_densities}: {time.time() - t0:.1f} s, mean AAR iterations {its[:k + 1].mean():.0f}")
    ps.close()
    return R, E, its


def training_tensors(unet: UNet, poisson: Poisson, R, E):
    """Inputs D^{-1} r and targets e - D^{-1} r, each pair scaled by
    1 / rms(D^{-1} r)."""
    X = unet.input_of(poisson, R.to(torch.float64).contiguous())
    scale = 1.0 / torch.sqrt(torch.mean(X.to(torch.float64) ** 2, dim=1))
    X = (X * scale[:, None].to(torch.float32)).contiguous()
    T = (E * scale[:, None].to(torch.float32) - X).contiguous()
    return X, T


def train(unet: UNet, X, T, epochs=20, batch=8, lr=2e-3, seed=0, val=None, log=None):
    """Adam with cosine decay over epochs of shuffled mini-batches; returns
    the per-epoch mean training loss (and validation loss when val=(Xv, Tv))."""
    n = X.shape[0]
    rng = np.random.default_rng(seed)
    steps = epochs * (n // batch)
    hist = []
    step = 0
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(n)).to(X.device)
        tot = 0.0
        for b0 in range(0, n - batch + 1, batch):
            idx = perm[b0:b0 + batch]
            lr_t = 0.5 * lr * (1 + math.cos(math.pi * step / max(steps, 1)))
            tot += unet.train_step(X[idx].contiguous(), T[idx].contiguous(), max(lr_t, 1e-6))
            step += 1
        rec = {"epoch": ep + 1, "loss": tot / (n // batch)}
        if val is not None:
            Xv, Tv = val
            rec["val_loss"] = evaluate(unet, Xv, Tv)
        hist.append(rec)
        if log:
            log(f"epoch {ep + 1}: " + ", ".join(f"{k} {v:.4e}" for k, v in rec.items() if k != "epoch"))
    return hist


def evaluate(unet: UNet, X, T, batch=8):
    """Mean relative squared error of U on (X, T)."""
    tot = 0.0
    for b0 in range(0, X.shape[0], batch):
        x, t = X[b0:b0 + batch].contiguous(), T[b0:b0 + batch].contiguous()
        u = unet.forward(x)
        e = ((u - t).to(torch.float64) ** 2).sum(1) / (t.to(torch.float64) ** 2).sum(1)
        tot += float(e.sum())
    return tot / X.shape[0]


def default_weights_path(grid: GridSpec) -> Path:
    return WEIGHTS_DIR / f"unet_{grid.nr}x{grid.nt}x{grid.nz}.npz"


def compare_solves(dev: Device, system, unet: UNet, n=16, seed=12345, tol=1e-8, max_iter=4000, omega=0.6,
                   beta=0.6, p_learned=6):
    """Jacobi-AAR vs learned-preconditioner AAR on n fresh random densities
    (cold start, f = -4 pi rho): per-solve iterations and milliseconds."""
    g = system.grid
    ps = Poisson(dev, g)
    sampler = DensitySampler(system, dev.torch_device)
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        f = (-4 * np.pi * sampler.sample(rng)).contiguous()
        row = {}
        for mode in ("jacobi", "learned"):
            if mode == "learned":
                unet.attach(ps)
            pp = p_learned if mode == "learned" else 6
            ps.aar(f, None, tol, 20, omega, beta, 7, pp)   # warm the launch path
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x, info = ps.aar(f, None, tol, max_iter, omega, beta, 7, pp)
            torch.cuda.synchronize()
            row[mode] = {"iterations": int(info.iterations), "ms": (time.perf_counter() - t0) * 1e3,
                         "relres": float(info.relres), "converged": bool(info.converged),
                         "precond": int(info.precond)}
            if mode == "learned":
                unet.detach(ps)
        out.append(row)
    ps.close()
    return out


def main(argv=None):
    """Train the network of a workload's grid on the device and save it."""
    import argparse
    import json

    from .system import make_system

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--workload", default="cnt10_helical")
    ap.add_argument("--densities", type=int, default=4096, help="solves; each gives 1 + len(snapshots) pairs")
    ap.add_argument("--snapshots", default="", help="comma-separated AAR iteration counts for residual/error pairs")
    ap.add_argument("--val", type=int, default=32, help="densities held out for validation")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--compare", type=int, default=16)
    ap.add_argument("--noise", type=float, default=0.0, help="fraction of white-noise right-hand sides")
    ap.add_argument("--omega", type=float, default=0.6)
    ap.add_argument("--learned-p", type=int, default=6, help="AAR Anderson period with the learned correction")
    ap.add_argument("--compare-maxit", type=int, default=4000)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    system = make_system(args.workload)
    dev = Device(0)
    t0 = time.time()
    snaps = tuple(int(v) for v in args.snapshots.split(",") if v)
    F, PHI, its = generate_pairs(dev, system, args.densities, seed=args.seed, snapshots=snaps,
                                 noise_fraction=args.noise, log=log)
    Fv, PHIv, _ = generate_pairs(dev, system, args.val, seed=args.seed + 1, snapshots=snaps,
                                 noise_fraction=args.noise)
    log(f"generated {F.shape[0]} training and {Fv.shape[0]} validation pairs in {time.time() - t0:.1f} s "
        f"(Jacobi-AAR iterations: mean {its.mean():.0f}, max {its.max()})")
    unet = UNet(dev, system.grid)
    unet.init_params(seed=args.seed)
    ps = Poisson(dev, system.grid)
    Xt, Tt = training_tensors(unet, ps, F, PHI)
    Xv, Tv = training_tensors(unet, ps, Fv, PHIv)
    del F, PHI, Fv, PHIv
    log(f"validation loss before training: {evaluate(unet, Xv, Tv):.4e}")
    t1 = time.time()
    hist = train(unet, Xt, Tt, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed, val=(Xv, Tv),
                 log=log)
    log(f"trained {args.epochs} epochs in {time.time() - t1:.1f} s")
    out = Path(args.out) if args.out else default_weights_path(system.grid)
    out.parent.mkdir(parents=True, exist_ok=True)
    unet.save(out)
    x1 = Xv[:1].contiguous()
    unet.forward(x1)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(20):
        unet.forward(x1)
    torch.cuda.synchronize()
    log(f"forward (one sample): {(time.perf_counter() - t2) / 20 * 1e3:.3f} ms")
    cmp = compare_solves(dev, system, unet, n=args.compare, max_iter=args.compare_maxit, omega=args.omega,
                         beta=args.omega, p_learned=args.learned_p) if args.compare > 0 else []
    for row in cmp:
        log(json.dumps(row))
    if cmp:
        mj = np.mean([r["jacobi"]["iterations"] for r in cmp])
        ml = np.mean([r["learned"]["iterations"] for r in cmp])
        tj = np.mean([r["jacobi"]["ms"] for r in cmp])
        tl = np.mean([r["learned"]["ms"] for r in cmp])
        log(f"mean iterations jacobi {mj:.0f} learned {ml:.0f}; mean ms jacobi {tj:.1f} learned {tl:.1f}")
    with open(out.with_suffix(".json"), "w") as fh:
        json.dump({"args": vars(args), "history": hist, "generation_iterations_mean": float(its.mean()),
                   "compare": cmp}, fh, indent=1)
    log(f"saved {out}")


if __name__ == "__main__":
    main()
