"""Davidson eigensolver (gpaw/eigensolvers/davidson.py) on the device:
the FP64 generalized eigensolver against scipy, the driver against the
oracle's numpy restatement iteration by iteration, and convergence to the
dense eigenvalues of the FD PAW Hamiltonian."""
import numpy as np
import pytest
import scipy.linalg

import davidson as DV
import rmmdiis as RM
from helpers import dense_hamiltonian, near_eigenvector_start, \
    oracle_preconditioner

pytestmark = pytest.mark.gpu


def test_general_diagonalize_matches_scipy(cuda):
    import torch
    from gpawdft_amd.hs_operators import general_diagonalize
    rng = np.random.default_rng(4)
    n = 70
    X = rng.standard_normal((n, n))
    A = X + X.T
    Y = rng.standard_normal((n, n))
    B = Y @ Y.T / n + np.eye(n)
    a = torch.from_numpy(np.tril(A) + np.triu(np.full((n, n), 7.0), 1)).cuda()
    b = torch.from_numpy(np.tril(B) + np.triu(np.full((n, n), -3.0), 1)).cuda()
    w = torch.empty(n, dtype=torch.float64, device='cuda')
    general_diagonalize(a, w, b)
    e = scipy.linalg.eigh(A, B, eigvals_only=True)
    W = w.cpu().numpy()
    V = a.cpu().numpy()               # eigenvectors in rows
    assert np.abs(W - e).max() < 1e-11 * np.abs(e).max()
    assert np.abs(V @ B @ V.T - np.eye(n)).max() < 1e-10
    assert np.abs(V @ A @ V.T - np.diag(W)).max() < 1e-9 * np.abs(e).max()


def _setup(dtype, nb):
    import torch
    from gpawdft_amd.eigensolvers import WaveFunctions
    from gpawdft_amd.hamiltonian import Hamiltonian
    from gpawdft_amd.synthetic import System
    sysm = System(16, 0.3, 2, seed=11, unit_projectors=True)
    kin, pt = sysm.operators()
    H = dense_hamiltonian(sysm, pt)
    tdt = torch.float64 if dtype == 'f64' else torch.float32
    wfs = WaveFunctions(sysm.gd, kin, pt, nb, tdt)
    ham = Hamiltonian(sysm.gd, [sysm.vt_G], sysm.dH_asp)
    return sysm, H, wfs, ham, tdt


def test_davidson_f64_follows_restatement(cuda):
    import torch
    from gpawdft_amd.eigensolvers import Davidson, KPoint
    from gpawdft_amd.synthetic import smooth_bands_numpy
    nb = 12
    sysm, H, wfs, ham, tdt = _setup('f64', nb)
    gd = sysm.gd
    psi0 = smooth_bands_numpy(nb, gd.n_c, 3)
    kpt = KPoint(torch.from_numpy(psi0).cuda())
    es = Davidson()
    pc = oracle_preconditioner(gd)
    psi = RM.orthonormalize(psi0.reshape(nb, -1).copy(), gd.dv)
    for it in range(6):
        e_dev = es.iterate(ham, wfs, kpt)
        e_ref, eps_ref, psi = DV.davidson_iteration(H, psi, gd.dv, pc)
        assert abs(e_dev - e_ref) < 1e-9 * e_ref, (it, e_dev, e_ref)
        assert np.abs(kpt.eps_n - eps_ref).max() < 1e-8, it
    # the device bands stay orthonormal without a separate step
    p = kpt.psit_nG.reshape(nb, -1).cpu().numpy()
    assert np.abs(gd.dv * p @ p.T - np.eye(nb)).max() < 1e-10


@pytest.mark.parametrize('dtype', ['f64', 'f32'])
def test_davidson_converges_to_dense_eigenvalues(cuda, dtype):
    import torch
    from gpawdft_amd.eigensolvers import Davidson, KPoint
    nb = 12
    sysm, H, wfs, ham, tdt = _setup(dtype, nb)
    w, psi0 = near_eigenvector_start(H, nb, sysm.gd)
    kpt = KPoint(torch.from_numpy(psi0.reshape((nb,) + tuple(sysm.gd.n_c)))
                 .cuda().to(tdt))
    es = Davidson()
    errors = [es.iterate(ham, wfs, kpt) for it in range(40)]
    assert errors[-1] < 1e-4 * errors[0], errors[-3:]
    tol = 1e-9 if dtype == 'f64' else 2e-5
    assert np.abs(kpt.eps_n[:6] - w[:6]).max() < tol, kpt.eps_n[:6] - w[:6]
