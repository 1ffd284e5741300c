"""Learned preconditioner: a periodic 3-D U-Net beside the multigrid one.

No counterpart in the reference.  It plugs in where
gpaw/preconditioner.py:10-60 does (`__call__(residuals, kpt, ekin=None)`,
`calculate_kinetic_energy`), and the multigrid V-cycle stays both its base
and its fallback:

    P(R) = MG(R) + s * N([R / s, v - eps])        s = rms(R) per band

N is a U-Net with 3 levels (16, 32, 64 channels, 3x3x3 periodic
convolutions, leaky ReLU 0.1, nn=1 restriction / interpolation between
levels, additive skips); it is trained to map the scaled residual and the
shifted local potential to (-(T - eps)^-1 R - MG(R)) / s, the part of the
exact shifted inverse the V-cycle misses.  Everything runs in this
library's gfx950 kernels (conv3d forward / backward-data / backward-weight,
leaky ReLU, the nn=1 transformers, Adam); torch only holds device memory.

`candidates()` returns both the multigrid and the learned correction so
the eigensolver can keep, band by band, whichever reduces the residual norm
more (the multigrid fallback).
"""
import numpy as np
import torch

from gpawdft_amd._lib import check, lib
from gpawdft_amd.device import stream_ptr
from gpawdft_amd.grid_descriptor import GridDescriptor
from gpawdft_amd.transformers import Transformer

SLOPE = 0.1

# (name, kind, in channels, out channels): 'conv' layers are followed by a
# leaky ReLU except the last; 'save' keeps the activation of a level,
# 'down' restricts to the next level, 'up' interpolates back and adds the
# saved activation of that level.
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
CONVS = [(n, ci, co) for n, k, ci, co in LAYERS if k == 'conv']


def _f32(t):
    assert t.dtype == torch.float32 and t.is_contiguous() and t.is_cuda
    return t.data_ptr()


class UNet:
    """Parameters live in one flat FP32 device buffer, layer by layer,
    weights (out, in, 27) then biases (out)."""

    def __init__(self, params=None, seed=0, device='cuda'):
        self.device = torch.device(device)
        self.offsets = {}
        n = 0
        for name, ci, co in CONVS:
            self.offsets[name + '.w'] = (n, (co, ci, 27))
            n += co * ci * 27
            self.offsets[name + '.b'] = (n, (co,))
            n += co
        self.nparams = n
        self.flat = torch.zeros(n, dtype=torch.float32, device=self.device)
        self.grad = torch.zeros_like(self.flat)
        self.m = torch.zeros_like(self.flat)
        self.v = torch.zeros_like(self.flat)
        self.t = 0
        if params is None:
            rng = np.random.default_rng(seed)
            params = {}
            for name, ci, co in CONVS:
                # the last layer starts at zero: an untrained network adds
                # nothing to the multigrid correction
                std = 0.0 if name == CONVS[-1][0] else np.sqrt(2.0 / (27 * ci))
                params[name + '.w'] = rng.standard_normal((co, ci, 27)) * std
                params[name + '.b'] = np.zeros(co)
        self.set_params(params)
        self._transfers = {}

    # parameters ---------------------------------------------------------
    def param(self, key, grad=False):
        off, shape = self.offsets[key]
        buf = self.grad if grad else self.flat
        return buf[off:off + int(np.prod(shape))].view(shape)

    def set_params(self, params):
        for key, (off, shape) in self.offsets.items():
            a = np.asarray(params[key], np.float32).reshape(shape)
            self.param(key).copy_(torch.from_numpy(a))

    def params_numpy(self):
        return {k: self.param(k).double().cpu().numpy()
                for k in self.offsets}

    def grads_numpy(self):
        return {k: self.param(k, True).double().cpu().numpy()
                for k in self.offsets}

    def save(self, path):
        np.savez(path, **{k.replace('.', '__'): v.astype(np.float32)
                          for k, v in self.params_numpy().items()})

    @classmethod
    def load(cls, path, device='cuda'):
        with np.load(path, allow_pickle=False) as f:
            params = {k.replace('__', '.'): f[k] for k in f.files}
        return cls(params, device=device)

    # grid transfers -----------------------------------------------------
    def _transfer(self, shape, down):
        key = (tuple(shape), down)
        if key not in self._transfers:
            n = np.array(shape)
            gd = GridDescriptor(n, n * 1.0, True)
            other = gd.coarsen() if down else gd.refine()
            self._transfers[key] = Transformer(gd, other, 1)
        return self._transfers[key]

    def _restrict(self, h):
        B, C = h.shape[:2]
        out = torch.empty((B, C) + tuple(s // 2 for s in h.shape[2:]),
                          dtype=h.dtype, device=h.device)
        self._transfer(h.shape[2:], True).apply(
            h.view((B * C,) + tuple(h.shape[2:])),
            out.view((B * C,) + tuple(out.shape[2:])))
        return out

    def _interpolate(self, h):
        B, C = h.shape[:2]
        out = torch.empty((B, C) + tuple(2 * s for s in h.shape[2:]),
                          dtype=h.dtype, device=h.device)
        self._transfer(h.shape[2:], False).apply(
            h.view((B * C,) + tuple(h.shape[2:])),
            out.view((B * C,) + tuple(out.shape[2:])))
        return out

    @staticmethod
    def _axpby(alpha, x, beta, y):
        """y = alpha x + beta y (in place on y)."""
        check(lib().gpd_scale_add(y.numel(), float(alpha), _f32(x),
                                  float(beta), _f32(y), stream_ptr()),
              'scale_add')

    # forward / backward -------------------------------------------------
    def forward(self, x, keep=False, bf16=False):
        """x: (B, 2, D, H, W) FP32 on the device, D, H, W divisible by 4.
        Returns (B, 1, D, H, W); keep=True stores what backward needs.
        bf16=True (inference) runs the convolutions whose shapes allow it
        on MFMA with bfloat16 operands, the others in FP32."""
        assert x.ndim == 5 and x.shape[1] == 2
        assert all(s % 4 == 0 for s in x.shape[2:]), 'grid must divide by 4'
        x = x.contiguous()
        st = stream_ptr()
        cache = []
        saved = {}
        h = x
        last = CONVS[-1][0]
        for name, kind, ci, co in LAYERS:
            if kind == 'conv':
                B, _, D, H, W = h.shape
                z = torch.empty((B, co, D, H, W), dtype=torch.float32,
                                device=h.device)
                rc = -3
                if bf16 and not keep:
                    rc = lib().gpd_conv3d_forward_bf16(
                        B, ci, co, D, H, W, _f32(h),
                        _f32(self.param(name + '.w')),
                        _f32(self.param(name + '.b')), _f32(z), st)
                    if rc not in (0, -3):
                        check(rc, 'conv3d_forward_bf16')
                if rc != 0:
                    check(lib().gpd_conv3d_forward(
                        B, ci, co, D, H, W, _f32(h),
                        _f32(self.param(name + '.w')),
                        _f32(self.param(name + '.b')), _f32(z), 0, st),
                        'conv3d_forward')
                if name != last:
                    check(lib().gpd_leaky_relu(z.numel(), SLOPE, _f32(z),
                                               _f32(z), st), 'leaky_relu')
                if keep:
                    cache.append((name, h, z))
                h = z
            elif kind == 'save':
                saved[name[-1]] = h
            elif kind == 'down':
                h = self._restrict(h)
            else:
                h = self._interpolate(h)
                self._axpby(1.0, saved[name[-1]], 1.0, h)
        self._cache = cache if keep else None
        return h

    def backward(self, gout, need_input_grad=False):
        """Accumulates nothing: overwrites self.grad with the gradients of
        sum(gout * forward(x)) from the last forward(keep=True).  Returns
        the input gradient when asked."""
        assert self._cache is not None, 'forward(keep=True) first'
        st = stream_ptr()
        g = gout.contiguous()
        last = CONVS[-1][0]
        first = CONVS[0][0]
        it = iter(reversed(self._cache))
        gskip = {}
        for name, kind, ci, co in reversed(LAYERS):
            if kind == 'conv':
                cname, hin, a = next(it)
                assert cname == name
                B, _, D, H, W = hin.shape
                if name != last:
                    gz = torch.empty_like(g)
                    check(lib().gpd_leaky_relu_backward(
                        g.numel(), SLOPE, _f32(a), _f32(g), _f32(gz), st),
                        'leaky_relu_backward')
                else:
                    gz = g
                check(lib().gpd_conv3d_backward_weight(
                    B, ci, co, D, H, W, _f32(hin), _f32(gz),
                    _f32(self.param(name + '.w', True)),
                    _f32(self.param(name + '.b', True)), st),
                    'conv3d_backward_weight')
                if name == first and not need_input_grad:
                    g = None
                    continue
                gx = torch.empty((B, ci, D, H, W), dtype=torch.float32,
                                 device=g.device)
                check(lib().gpd_conv3d_backward_data(
                    B, ci, co, D, H, W, _f32(gz),
                    _f32(self.param(name + '.w')), _f32(gx), 0, st),
                    'conv3d_backward_data')
                g = gx
            elif kind == 'up':
                gskip[name[-1]] = g
                # the interpolation's adjoint is 8x the restriction
                g = self._restrict(g)
                self._axpby(0.0, g, 8.0, g)
            elif kind == 'down':
                g = self._interpolate(g)
                self._axpby(0.0, g, 0.125, g)
            else:
                self._axpby(1.0, gskip[name[-1]], 1.0, g)
        self._cache = None
        return g

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.t += 1
        lr_t = lr * np.sqrt(1 - beta2**self.t) / (1 - beta1**self.t)
        check(lib().gpd_adam_step(self.nparams, float(lr_t), float(beta1),
                                  float(beta2), float(eps), _f32(self.flat),
                                  _f32(self.grad), _f32(self.m),
                                  _f32(self.v), stream_ptr()), 'adam_step')

    def flops_per_point(self):
        """Multiply-adds x 2 per input grid point of one forward pass."""
        f = 0.0
        level = 0
        for name, kind, ci, co in LAYERS:
            if kind == 'conv':
                f += 2 * 27 * ci * co / 8**level
            elif kind == 'down':
                level += 1
            elif kind == 'up':
                level -= 1
        return f


def network_input(R, v, eps, out=None):
    """(B, 2, ...) FP32 input: R / rms(R) and v - eps per band; returns
    (x, s) with s the per-band rms."""
    B = R.shape[0]
    R32 = R.to(torch.float32)
    s = torch.sqrt(torch.mean(R32.reshape(B, -1)**2, dim=1)).clamp_min(1e-30)
    x = out if out is not None else torch.empty(
        (B, 2) + tuple(R.shape[1:]), dtype=torch.float32, device=R.device)
    x[:, 0] = R32 / s.view((B,) + (1,) * (R.ndim - 1))
    e = torch.as_tensor(np.asarray(eps, np.float32), device=R.device)
    x[:, 1] = v.to(torch.float32)[None] - e.view((B,) + (1,) * (R.ndim - 1))
    return x, s


class LearnedPreconditioner:
    """MG(R) + s N(R / s, v - eps); multigrid fallback via candidates()."""

    uses_eigenvalues = True

    def __init__(self, mg, net, v_G, bf16=True):
        self.mg = mg
        self.net = net
        self.bf16 = bf16
        self.block = mg.block
        self.v = torch.as_tensor(np.asarray(v_G, np.float32)).to(
            net.device)

    def calculate_kinetic_energy(self, psit_xG, kpt):
        return None

    def learned_part(self, residuals, eps_x):
        x, s = network_input(residuals, self.v, eps_x)
        y = self.net.forward(x, bf16=self.bf16)[:, 0]
        B = residuals.shape[0]
        return (y * s.view((B,) + (1,) * (residuals.ndim - 1))).to(
            residuals.dtype)

    def candidates(self, residuals, kpt, eps_x):
        """[multigrid correction, learned correction] for the bands."""
        d_mg = self.mg(residuals, kpt).clone()
        d = d_mg + self.learned_part(residuals, eps_x)
        return [d_mg, d]

    def __call__(self, residuals, kpt=None, ekin=None, out=None):
        if ekin is None:
            raise ValueError('the learned preconditioner needs the band '
                             'eigenvalues (passed as ekin)')
        d = self.mg(residuals, kpt, out=out)
        d += self.learned_part(residuals, ekin)
        return d
