"""Pins the numpy restatement of the learned preconditioner's network
(oracle/unet.py): convolution against scipy.ndimage.correlate, the
backward passes by adjointness and finite differences, the grid transfers
against the C oracle's nn=1 transformers, the exact shifted inverse against
the FD operator."""
import numpy as np
import pytest
import scipy.ndimage

import oracle as O
import unet as U
from gpawdft_amd.fd_operators import guc_laplace_stencil
from gpawdft_amd.grid_descriptor import GridDescriptor
from gpawdft_amd.transformers import transformer_paddings


def test_conv3d_matches_scipy_correlate():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 6, 5, 7))
    w = rng.standard_normal((4, 3, 27))
    b = rng.standard_normal(4)
    y = U.conv3d(x, w, b)
    for bb in range(2):
        for o in range(4):
            ref = b[o] + sum(
                scipy.ndimage.correlate(x[bb, i], w[o, i].reshape(3, 3, 3),
                                        mode='wrap') for i in range(3))
            assert np.abs(y[bb, o] - ref).max() < 1e-12


def test_conv3d_backward_by_adjointness():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 3, 4, 6, 5))
    gy = rng.standard_normal((2, 5, 4, 6, 5))
    w = rng.standard_normal((5, 3, 27))
    lhs = np.sum(U.conv3d(x, w) * gy)
    assert abs(lhs - np.sum(x * U.conv3d_backward_data(gy, w))) < 1e-10
    gw, gb = U.conv3d_backward_weight(x, gy)
    assert abs(lhs - np.sum(w * gw)) < 1e-10
    assert np.allclose(gb, gy.sum(axis=(0, 2, 3, 4)))


@pytest.mark.parametrize('interp', [False, True])
def test_grid_transfers_match_c_oracle(interp):
    n = 8 if not interp else 4
    gd = GridDescriptor((n, n, n), (5.0, 5.0, 5.0), True)
    other = gd.coarsen() if not interp else gd.refine()
    pad, npad, skip, _ = (transformer_paddings(gd, other, 1))
    x = np.random.default_rng(3).standard_normal((2, n, n, n))
    ref = O.transform(2, interp, pad, skip, other.n_c, gd.pbc_c, x)
    got = U.interp2(x) if interp else U.restrict2(x)
    assert np.abs(got - ref).max() < 1e-14


def test_transfer_adjointness():
    rng = np.random.default_rng(4)
    a = rng.standard_normal((1, 8, 8, 8))
    c = rng.standard_normal((1, 4, 4, 4))
    assert abs(np.sum(U.restrict2(a) * c) -
               np.sum(a * U.interp2(c)) / 8) < 1e-12


def test_network_gradients_by_finite_differences():
    rng = np.random.default_rng(5)
    p = U.init_params(7)
    for k in p:
        if k.endswith('.b'):
            p[k] = 0.1 * rng.standard_normal(p[k].shape)
    x = rng.standard_normal((2, 2, 8, 8, 8))
    gout = rng.standard_normal((2, 1, 8, 8, 8))
    y, cache = U.forward(p, x)
    assert y.shape == (2, 1, 8, 8, 8)
    grads, gx = U.backward(p, cache, gout)

    def f(pp, xx):
        return np.sum(U.forward(pp, xx)[0] * gout)

    # small steps: a random direction moves many pre-activations at once,
    # and a step across a leaky-ReLU kink would spoil the difference
    h = 1e-7
    for name in ['e0.w', 'e1.b', 'e2a.w', 'e2b.w', 'd1.w', 'out.w', 'out.b']:
        v = rng.standard_normal(p[name].shape)
        pp = dict(p)
        pm = dict(p)
        pp[name] = p[name] + h * v
        pm[name] = p[name] - h * v
        fd = (f(pp, x) - f(pm, x)) / (2 * h)
        an = np.sum(grads[name] * v)
        assert abs(fd - an) < 1e-5 * max(1.0, abs(an)), name
    v = rng.standard_normal(x.shape)
    fd = (f(p, x + h * v) - f(p, x - h * v)) / (2 * h)
    assert abs(fd - np.sum(gx * v)) < 1e-5 * max(1.0, abs(fd))


def test_shifted_inverse_solves_the_fd_equation():
    gd = GridDescriptor((12, 10, 8), (4.0, 3.5, 3.0), True)
    coefs, offs, _ = guc_laplace_stencil(gd, -0.5, 3)
    R = np.random.default_rng(6).standard_normal((2, 12, 10, 8))
    eps = np.array([-0.3, -1.1])
    d = U.shifted_inverse(R, eps, gd.n_c, gd.h_cv.diagonal())
    Td = O.fd(coefs, offs, gd.pbc_c, d)
    for n in range(2):
        # (T - eps) d = -R
        assert np.abs(Td[n] - eps[n] * d[n] + R[n]).max() < 1e-10
