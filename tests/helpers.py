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


def oracle_preconditioner_args(gd):
    """Arguments of oracle.precondition for the reference's two-level
    preconditioner on gd (third-order kinetic stencil on gd, first order on
    the coarse grids; gpaw/preconditioner.py:10-29)."""
    from gpawdft_amd.fd_operators import guc_laplace_stencil
    from gpawdft_amd.transformers import transformer_paddings
    gd1 = gd.coarsen()
    gd2 = gd1.coarsen()
    c0, o0, _ = guc_laplace_stencil(gd, -0.5, 3)
    c1, o1, _ = guc_laplace_stencil(gd1, -0.5, 1)
    c2, o2, _ = guc_laplace_stencil(gd2, -0.5, 1)

    def targs(a, b):
        pad, npad, skip, interp = transformer_paddings(a, b, 1)
        return (2, pad, skip, a.n_c, b.n_c)

    return ((c0, o0, gd.n_c), (c1, o1, gd1.n_c), (c2, o2, gd2.n_c),
            targs(gd, gd1), targs(gd1, gd2), targs(gd2, gd1), targs(gd1, gd),
            gd.pbc_c, 0.66666666 / c0[0])


def oracle_preconditioner(gd):
    """Callable R[n, G] -> oracle preconditioned R (flattened bands)."""
    import oracle as O
    args = oracle_preconditioner_args(gd)
    shape = tuple(gd.n_c)

    def pc(r):
        return O.precondition(*args, r.reshape((-1,) + shape)).reshape(
            r.shape[0], -1)
    return pc


def dense_hamiltonian(sysm, pt):
    """H of a synthetic system as a dense matrix: the oracle's H psi
    (kinetic stencil of order sysm.nn + local potential + projector term)
    applied to unit vectors."""
    import oracle as O
    from gpawdft_amd.fd_operators import guc_laplace_stencil
    from gpawdft_amd.hamiltonian import unpack
    gd = sysm.gd
    coefs, offsets, _ = guc_laplace_stencil(gd, -0.5, sysm.nn)
    G = int(np.prod(gd.n_c))
    E = np.eye(G).reshape((G,) + tuple(gd.n_c))
    t = O.fd(coefs, offsets, gd.pbc_c, E)
    t = O.add_local_potential(E, sysm.vt_G, t)
    P = O.lfc_integrate(pt.A_Wgm, pt.M_W, pt.G_B, pt.W_B, gd.dv, E, pt.Mmax)
    c = O.projector_coefs(P, [(pt.M_a[a], unpack(sysm.dH_asp[a][0]))
                              for a in pt.atom_indices])
    return O.lfc_add(pt.A_Wgm, pt.M_W, pt.G_B, pt.W_B, c, t).reshape(G, G)


def near_eigenvector_start(H, nb, gd, noise=0.3, seed=5):
    """The lowest nb eigenvectors of H (grid-normalised) plus smooth noise of
    relative size `noise`: a start inside RMM-DIIS's basin of convergence."""
    from gpawdft_amd.synthetic import smooth_bands_numpy
    w, V = np.linalg.eigh(0.5 * (H + H.T))
    psi = V[:, :nb].T / np.sqrt(gd.dv)
    r = smooth_bands_numpy(nb, gd.n_c, seed).reshape(nb, -1)
    r *= noise / np.sqrt(gd.dv * np.sum(r**2, axis=1))[:, None]
    return w, psi + r
