"""numpy FP64 restatement of the learned preconditioner's network --
TEST INFRASTRUCTURE.

No counterpart in the reference: the network stands beside the multigrid
preconditioner (gpaw/preconditioner.py) and plugs in at its
__call__(residuals, kpt, ekin) interface.  This module is the CPU oracle
for the device kernels (gpawdft_amd/csrc/conv.hip) and the device network
(gpawdft_amd/learned.py): the same layer list, the same parameter layout,
forward and backward written out by hand.  Tests pin it by finite
differences, by scipy.ndimage.correlate for the convolution, by
adjointness, and by the C oracle's nn=1 transformers for the grid
transfers.

Conventions: grids (B, C, D, H, W), periodic.  Tap t of a 3x3x3 kernel has
offset d_t = (t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1); the convolution is
    y[b, o, p] = bias[o] + sum_{i, t} w[o, i, t] x[b, i, p + d_t].
"""
import numpy as np

TAPS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]
SLOPE = 0.1

# (name, kind, in channels, out channels); 'conv' layers are followed by a
# leaky ReLU except the last one.  Levels: 0 = input grid, 1 = half, 2 =
# quarter; 'down' restricts (nn=1 transformer), 'up' interpolates and adds
# the saved skip of that level.
LAYERS = [
    ('e0', 'conv', 2, 16),
    ('save0', 'save', None, None),
    ('down1', 'down', None, None),
    ('e1', 'conv', 16, 32),
    ('save1', 'save', None, None),
    ('down2', 'down', None, None),
    ('e2a', 'conv', 32, 64),
    ('e2b', 'conv', 64, 32),
    ('up1', 'up', None, None),
    ('d1', 'conv', 32, 16),
    ('up0', 'up', None, None),
    ('out', 'conv', 16, 1),
]


def conv_layers():
    return [(n, ci, co) for n, k, ci, co in LAYERS if k == 'conv']


def init_params(seed=0, scale=1.0):
    """He-style initial weights (o, i, 27) and zero biases, FP64."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, ci, co in conv_layers():
        std = scale * np.sqrt(2.0 / (27 * ci))
        p[name + '.w'] = rng.standard_normal((co, ci, 27)) * std
        p[name + '.b'] = np.zeros(co)
    return p


def _shift(x, d):
    """x[..., p + d] with periodic wrap over the last three axes."""
    return np.roll(x, (-d[0], -d[1], -d[2]), axis=(-3, -2, -1))


def conv3d(x, w, b=None):
    B, C = x.shape[:2]
    y = np.zeros((B, w.shape[0]) + x.shape[2:])
    for t, d in enumerate(TAPS):
        xs = _shift(x, d)
        y += np.einsum('oi,bi...->bo...', w[:, :, t], xs)
    if b is not None:
        y += b[None, :, None, None, None]
    return y


def conv3d_backward_data(gy, w):
    gx = np.zeros((gy.shape[0], w.shape[1]) + gy.shape[2:])
    for t, d in enumerate(TAPS):
        gx += _shift(np.einsum('oi,bo...->bi...', w[:, :, t], gy),
                     (-d[0], -d[1], -d[2]))
    return gx


def conv3d_backward_weight(x, gy):
    B, O_, I_ = gy.shape[0], gy.shape[1], x.shape[1]
    gw = np.empty((O_, I_, 27))
    g2 = gy.reshape(B, O_, -1)
    for t, d in enumerate(TAPS):
        gw[:, :, t] = np.einsum('bop,bip->oi', g2,
                                _shift(x, d).reshape(B, I_, -1))
    return gw, gy.sum(axis=(0, 2, 3, 4))


def lrelu(x):
    return np.where(x > 0, x, SLOPE * x)


def lrelu_backward(x, gy):
    return np.where(x > 0, gy, SLOPE * gy)


def restrict2(x):
    """nn=1 restriction on periodic grids (c/bmgs/restrict.c with K=2):
    per axis out[i] = 0.5 * (x[2i] + 0.5 * (x[2i-1] + x[2i+1]))."""
    for ax in (-3, -2, -1):
        xm = np.roll(x, 1, axis=ax)
        xp = np.roll(x, -1, axis=ax)
        y = 0.5 * (x + 0.5 * (xm + xp))
        sl = [slice(None)] * x.ndim
        sl[ax] = slice(0, None, 2)
        x = y[tuple(sl)]
    return x


def interp2(x):
    """nn=1 interpolation on periodic grids (c/bmgs/interpolate.c with
    K=2): per axis out[2i] = x[i], out[2i+1] = 0.5 (x[i] + x[i+1])."""
    for ax in (-3, -2, -1):
        n = x.shape[ax]
        shp = list(x.shape)
        shp[ax] = 2 * n
        y = np.empty(shp)
        ev = [slice(None)] * x.ndim
        od = [slice(None)] * x.ndim
        ev[ax] = slice(0, None, 2)
        od[ax] = slice(1, None, 2)
        y[tuple(ev)] = x
        y[tuple(od)] = 0.5 * (x + np.roll(x, -1, axis=ax))
        x = y
    return x


def forward(p, x):
    """Returns (output (B, 1, ...), cache for backward)."""
    cache = []
    saved = {}
    h = x
    convs = [n for n, k, ci, co in LAYERS if k == 'conv']
    for name, kind, ci, co in LAYERS:
        if kind == 'conv':
            z = conv3d(h, p[name + '.w'], p[name + '.b'])
            cache.append((name, 'conv', h, z))
            h = z if name == convs[-1] else lrelu(z)
        elif kind == 'save':
            saved[name[-1]] = h
        elif kind == 'down':
            h = restrict2(h)
        else:
            h = interp2(h) + saved[name[-1]]
    return h, cache


def backward(p, cache, gout):
    """Gradients of sum(gout * forward(x)) w.r.t. parameters and x."""
    grads = {}
    g = gout
    convs = [n for n, k, ci, co in LAYERS if k == 'conv']
    ci_iter = iter(reversed(cache))
    gskip = {}
    for name, kind, ci, co in reversed(LAYERS):
        if kind == 'conv':
            cname, _, hin, z = next(ci_iter)
            assert cname == name
            gz = g if name == convs[-1] else lrelu_backward(z, g)
            grads[name + '.w'], grads[name + '.b'] = \
                conv3d_backward_weight(hin, gz)
            g = conv3d_backward_data(gz, p[name + '.w'])
        elif kind == 'up':
            gskip[name[-1]] = g
            # interpolation's adjoint is 8x the restriction
            g = 8.0 * restrict2(g)
        elif kind == 'down':
            g = interp2(g) / 8.0
        else:
            g = g + gskip[name[-1]]
    return grads, g


def fd_symbol(n_c, h_c, coefs=(-49.0 / 18, 1.5, -0.15, 1.0 / 90)):
    """Eigenvalues of the kinetic FD operator -0.5 * Laplace (nn=3 GPAW
    coefficients by default) on a periodic grid, on numpy's FFT frequency
    grid (gpaw/fd_operators.py:21-28)."""
    lam = 0.0
    for c, (n, h) in enumerate(zip(n_c, h_c)):
        k = 2 * np.pi * np.fft.fftfreq(n)
        s = coefs[0] + 2 * sum(coefs[j] * np.cos(j * k)
                               for j in range(1, len(coefs)))
        shape = [1, 1, 1]
        shape[c] = n
        lam = lam + (-0.5 * s / h**2).reshape(shape)
    return lam


def shifted_inverse(R, eps, n_c, h_c):
    """-(T - eps)^-1 R on a periodic grid, per band (R: (nb, n0, n1, n2)),
    exact in Fourier space."""
    lam = fd_symbol(n_c, h_c)
    out = np.empty_like(R)
    for n in range(R.shape[0]):
        F = np.fft.fftn(R[n])
        out[n] = -np.real(np.fft.ifftn(F / (lam - eps[n])))
    return out
