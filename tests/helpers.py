"""Shared test helpers: bitwise comparison and oracle-side set-up."""
import numpy as np


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64),
                                                 b.view(np.int64))


def max_ulp_diff(a, b):
    a = np.ascontiguousarray(a, np.float64).view(np.int64).ravel()
    b = np.ascontiguousarray(b, np.float64).view(np.int64).ravel()
    return int(np.abs(a - b).max()) if a.size else 0


def rel_err(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300))


def to_np(t):
    return t.detach().cpu().numpy()


def transformer_args(t):
    """(k, pad_cd, skip_cd, n_in, n_out) of a gpawdft_amd Transformer for
    the oracle."""
    return (2 * t.nn, t.pad_cd, t.skip_cd, t.gdin.n_c, t.gdout.n_c)


def stencil_args(op):
    return (op.coef_p, op.offset_pc, op.gd.n_c)
