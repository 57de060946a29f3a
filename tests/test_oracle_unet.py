"""Pins of the U-Net oracle: adjoint identities of its three operators,
linearity, and gradients against central finite differences."""
import numpy as np
import pytest

from oracle import unet_oracle as U

DIMS = (8, 4, 8)  # nr, nt, nz (divisible by 4)


def rand_params(seed, dims=DIMS, scale=0.3):
    nr = dims[0]
    off, n = U.offsets(nr)
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(n) * scale
    p[off[-1]:] = 1.0 + 0.1 * rng.standard_normal(nr)
    return p


@pytest.mark.parametrize("op,adj,shape_in,wshape", [
    (U.conv3, U.conv3_adj, (4, 4, 6, 3), (27, 5, 3)),
    (U.down, U.down_adj, (4, 4, 6, 3), (8, 5, 3)),
    (U.up, U.up_adj, (2, 2, 3, 3), (8, 5, 3)),
])
def test_adjoint_identity(op, adj, shape_in, wshape):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape_in)
    W = rng.standard_normal(wshape)
    y = op(x, W)
    g = rng.standard_normal(y.shape)
    assert np.sum(y * g) == pytest.approx(np.sum(x * adj(g, W)), rel=1e-12)


def test_conv3_zero_padding_in_r_and_periodic_theta_z():
    x = np.zeros((4, 4, 5, 1))
    x[0, 0, 0, 0] = 1.0
    W = np.zeros((27, 1, 1))
    W[13 - 1] = 1.0             # tap (0, 0, -1): y[i] = x[i-1]
    y = U.conv3(x, W)
    assert y[0, 0, 1, 0] == 1.0 and np.sum(y) == 1.0
    W[:] = 0
    W[13 - 9] = 1.0             # tap (-1, 0, 0): y[k] = x[k-1], periodic
    y = U.conv3(x, W)
    assert y[1, 0, 0, 0] == 1.0
    x2 = np.zeros((4, 4, 5, 1))
    x2[3, 0, 0, 0] = 1.0
    assert U.conv3(x2, W)[0, 0, 0, 0] == 1.0
    W[:] = 0
    W[13 + 1] = 1.0             # y[i] = x[i+1]: the last radius reads zero
    x3 = np.zeros((4, 4, 5, 1))
    x3[..., 4, 0] = 1.0
    assert np.all(U.conv3(x3, W)[..., 4, 0] == 0.0) and np.all(U.conv3(x3, W)[..., 3, 0] == 1.0)


def test_network_is_linear():
    p = rand_params(2)
    rng = np.random.default_rng(3)
    N = int(np.prod(DIMS))
    x, y = rng.standard_normal(N), rng.standard_normal(N)
    ux, _ = U.forward(p, x, DIMS)
    uy, _ = U.forward(p, y, DIMS)
    uxy, _ = U.forward(p, 2.0 * x - 0.5 * y, DIMS)
    assert np.allclose(uxy, 2.0 * ux - 0.5 * uy, rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences():
    p = rand_params(4)
    rng = np.random.default_rng(5)
    N = int(np.prod(DIMS))
    X = rng.standard_normal((2, N))
    T = rng.standard_normal((2, N))
    L, g = U.loss_and_grad(p, X, T, DIMS)
    off, n = U.offsets(DIMS[0])
    idx = [off[l] + int(rng.integers(off[l + 1] - off[l])) for l in range(len(U.SPEC))] + [off[-1] + 3]
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (U.loss_and_grad(pp, X, T, DIMS)[0] - U.loss_and_grad(pm, X, T, DIMS)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9), k


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -3.14159265])
    r = U.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7
    assert abs(r[3] - v[3]) <= 2 ** -8 * abs(v[3])
