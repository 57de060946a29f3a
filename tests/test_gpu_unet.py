"""Device U-Net (forward, gradients) against the FP64 oracle with the same
BF16-rounded weights, and its use as the AAR preconditioner."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import unet_oracle as U  # noqa: E402

pytestmark = pytest.mark.gpu


def small_grid():
    from cyclix_msparc_amd.system import GridSpec
    return GridSpec(16, 8, 12, 3.0, 0.35, 2 * np.pi / 6 / 8, 0.5, 0.0, 6, 3)


def rand_params(nr, seed):
    off, n = U.offsets(nr)
    rng = np.random.default_rng(seed)
    p = np.zeros(n)
    for l, (mode, ci, co) in enumerate(U.SPEC):
        fan = (27 * ci) if mode == U.CONV3 else (8 * ci if mode == U.DOWN else ci)
        p[off[l]:off[l + 1]] = rng.standard_normal(off[l + 1] - off[l]) / np.sqrt(fan)
    p[off[-1]:] = 1.0 + 0.2 * rng.standard_normal(nr)
    return p.astype(np.float32)


def oracle_params(p, nr):
    off, _ = U.offsets(nr)
    q = p.astype(np.float64)
    q[:off[-1]] = U.bf16_round(p[:off[-1]])
    return q


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_unet_forward_matches_oracle(device):
    from cyclix_msparc_amd.learned import UNet
    g = small_grid()
    dims = (g.nr, g.nt, g.nz)
    net = UNet(device, g)
    p = rand_params(g.nr, 1)
    net.set_params(p)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((3, g.N)).astype(np.float32)
    u = net.forward(torch.from_numpy(X).to(device.torch_device)).cpu().numpy()
    q = oracle_params(p, g.nr)
    for b in range(3):
        ref, _ = U.forward(q, X[b].astype(np.float64), dims)
        # BF16 rounding of the matrix-core inputs: a few parts in 1e3 per layer
        assert relerr(u[b], ref) < 3e-2, b
    net.close()


def test_unet_gradient_matches_oracle(device):
    from cyclix_msparc_amd.learned import UNet
    g = small_grid()
    dims = (g.nr, g.nt, g.nz)
    net = UNet(device, g)
    p = rand_params(g.nr, 3)
    net.set_params(p)
    rng = np.random.default_rng(4)
    X = rng.standard_normal((2, g.N)).astype(np.float32)
    T = rng.standard_normal((2, g.N)).astype(np.float32)
    td = device.torch_device
    loss, grad = net.gradient(torch.from_numpy(X).to(td), torch.from_numpy(T).to(td))
    L, gref = U.loss_and_grad(oracle_params(p, g.nr), X.astype(np.float64), T.astype(np.float64), dims)
    assert loss == pytest.approx(L, rel=2e-2)
    off, n = U.offsets(g.nr)
    for l in range(len(U.SPEC)):
        sl = slice(off[l], off[l + 1])
        assert relerr(grad[sl], gref[sl]) < 5e-2, f"layer {l}"
    assert relerr(grad[off[-1]:], gref[off[-1]:]) < 5e-2
    net.close()


def test_unet_training_reduces_loss(device):
    from cyclix_msparc_amd.learned import UNet
    g = small_grid()
    net = UNet(device, g)
    net.init_params(seed=0)
    rng = np.random.default_rng(5)
    td = device.torch_device
    X = torch.from_numpy(rng.standard_normal((8, g.N)).astype(np.float32)).to(td)
    # a linear target the network can represent: a smoothed copy of the input
    T = 0.5 * X + 0.25 * (torch.roll(X, 1, 1) + torch.roll(X, -1, 1))
    l0 = net.train_step(X, T, 1e-3)
    for _ in range(60):
        last = net.train_step(X, T, 1e-3)
    assert last < 0.5 * l0
    net.close()


def _poisson_case(device):
    from cyclix_msparc_amd.api import Poisson
    g = small_grid()
    rng = np.random.default_rng(6)
    f = torch.from_numpy(rng.standard_normal(g.N)).to(device.torch_device)
    return g, Poisson(device, g), f


def test_learned_aar_with_zero_network_converges_to_jacobi_solution(device):
    """U = 0 (last layer zero): the correction is a Jacobi step with exact
    line search after each Anderson iteration; the solve reaches the Jacobi
    solution and never needs more iterations."""
    from cyclix_msparc_amd.learned import UNet
    g, ps, f = _poisson_case(device)
    x_j, info_j = ps.aar(f, None, 1e-10, 3000)
    net = UNet(device, g)
    net.init_params(seed=1)
    net.attach(ps)
    x_l, info_l = ps.aar(f, None, 1e-10, 3000)
    assert info_j.converged and info_l.converged and info_l.precond == 1
    assert info_l.iterations <= info_j.iterations
    assert torch.max(torch.abs(x_l - x_j)) <= 1e-6 * torch.max(torch.abs(x_j))
    net.close()
    ps.close()


def test_learned_aar_converges_with_random_network(device):
    """Any U only changes the path (the line search cannot raise the
    residual): the solve still stops on the true residual and reaches the
    Jacobi solution."""
    from cyclix_msparc_amd.learned import UNet
    g, ps, f = _poisson_case(device)
    x_j, info_j = ps.aar(f, None, 1e-10, 5000)
    net = UNet(device, g)
    net.set_params(rand_params(g.nr, 7))
    net.attach(ps)
    x_l, info_l = ps.aar(f, None, 1e-10, 5000)
    assert info_l.converged and info_l.relres <= 1e-10 and info_l.precond == 1
    assert torch.max(torch.abs(x_l - x_j)) <= 1e-6 * torch.max(torch.abs(x_j))
    net.close()
    ps.close()


def test_trained_network_on_the_production_grid(device):
    """The committed network for the 120 x 24 x 96 grid: on fresh random
    densities its correction D^{-1} f + U(D^{-1} f) approximates the Poisson
    solution (relative error well below the Jacobi-only 100 %), and inside AAR
    the solve still reaches 1e-8 with the Jacobi solution.  The iteration
    counts of both are printed: on this grid the learned correction does not
    lower them, so the test asserts convergence and agreement only."""
    from cyclix_msparc_amd.api import Poisson
    from cyclix_msparc_amd.learned import DensitySampler, UNet, default_weights_path
    from cyclix_msparc_amd.system import make_system
    sysm = make_system("cnt10_helical")
    g = sysm.grid
    wp = default_weights_path(g)
    assert wp.exists(), f"missing trained weights {wp}"
    net = UNet(device, g)
    net.load(wp)
    ps = Poisson(device, g)
    sampler = DensitySampler(sysm, device.torch_device)
    rng = np.random.default_rng(777)
    for k in range(2):
        f = (-4 * np.pi * sampler.sample(rng)).contiguous()
        xj, ij = ps.aar(f, None, 1e-8, 5000)
        assert ij.converged
        x = net.input_of(ps, f[None, :].contiguous())
        approx = x[0].to(torch.float64) + net.forward(x)[0].to(torch.float64)
        rel_err = float(torch.linalg.norm(approx - xj) / torch.linalg.norm(xj))
        assert rel_err < 0.5, rel_err
        net.attach(ps)
        xl, il = ps.aar(f, None, 1e-8, 5000)
        net.detach(ps)
        assert il.converged and il.precond == 1
        assert float(torch.max(torch.abs(xl - xj)) / torch.max(torch.abs(xj))) < 1e-6
        print(f"density {k}: network relative error {rel_err:.3f}; AAR iterations jacobi {ij.iterations}, "
              f"learned {il.iterations}")
    net.close()
    ps.close()
