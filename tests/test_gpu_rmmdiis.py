"""RMM-DIIS on the device operators (gpaw/eigensolvers/rmmdiis.py:52-344).

* In FP64 the device driver follows the oracle's numpy restatement
  (oracle/rmmdiis.py) iteration by iteration from the same start: same
  residual error and eigenvalues up to rounding amplified by the iteration.
* From a start inside the basin of convergence it reaches the lowest
  eigenvalues of the dense finite-difference PAW Hamiltonian (brute-force
  diagonalisation of the oracle's H).
"""
import numpy as np
import pytest

import rmmdiis as RM
from helpers import dense_hamiltonian, near_eigenvector_start, \
    oracle_preconditioner

pytestmark = pytest.mark.gpu


def _setup(dtype, nb):
    import torch
    from gpawdft_amd.eigensolvers import WaveFunctions
    from gpawdft_amd.hamiltonian import Hamiltonian
    from gpawdft_amd.synthetic import System
    sysm = System(16, 0.3, 2, seed=11, unit_projectors=True)
    kin, pt = sysm.operators()
    H = dense_hamiltonian(sysm, pt)
    # the FD Hamiltonian with a symmetric dH is symmetric
    assert np.abs(H - H.T).max() < 1e-10 * np.abs(H).max()
    tdt = torch.float64 if dtype == 'f64' else torch.float32
    wfs = WaveFunctions(sysm.gd, kin, pt, nb, tdt)
    ham = Hamiltonian(sysm.gd, [sysm.vt_G], sysm.dH_asp)
    return sysm, H, wfs, ham, tdt


def test_rmmdiis_f64_follows_restatement(cuda):
    import torch
    from gpawdft_amd.eigensolvers import KPoint, RMMDIIS
    from gpawdft_amd.synthetic import smooth_bands_numpy
    nb = 12
    sysm, H, wfs, ham, tdt = _setup('f64', nb)
    gd = sysm.gd
    psi0 = smooth_bands_numpy(nb, gd.n_c, 3)
    kpt = KPoint(torch.from_numpy(psi0).cuda())
    es = RMMDIIS(blocksize=4)
    pc = oracle_preconditioner(gd)
    psi = RM.orthonormalize(psi0.reshape(nb, -1).copy(), gd.dv)
    for it in range(6):
        e_dev = es.iterate(ham, wfs, kpt)
        e_ref, eps_ref, psi = RM.rmmdiis_iteration(H, psi, gd.dv, pc,
                                                   blocksize=4)
        assert abs(e_dev - e_ref) < 1e-10 * e_ref, (it, e_dev, e_ref)
        assert np.abs(kpt.eps_n - eps_ref).max() < 1e-8, it
    dev = kpt.psit_nG.reshape(nb, -1).cpu().numpy()
    # same orthonormal bands (up to sign)
    s = np.sign(np.sum(dev * psi, axis=1))
    assert np.abs(dev - s[:, None] * psi).max() < 1e-6 * np.abs(psi).max()


@pytest.mark.parametrize('dtype', ['f64', 'f32'])
def test_rmmdiis_converges_to_dense_eigenvalues(cuda, dtype):
    import torch
    from gpawdft_amd.eigensolvers import KPoint, RMMDIIS
    nb = 12
    sysm, H, wfs, ham, tdt = _setup(dtype, nb)
    w, psi0 = near_eigenvector_start(H, nb, sysm.gd)
    kpt = KPoint(torch.from_numpy(psi0.reshape((nb,) + tuple(sysm.gd.n_c)))
                 .cuda().to(tdt))
    es = RMMDIIS(blocksize=4)
    errors = [es.iterate(ham, wfs, kpt) for it in range(30)]
    # the error falls by orders of magnitude; the top bands, next to
    # unconverged partners above the block, converge slowly
    assert errors[-1] < 1e-5 * errors[0], errors[-3:]
    # the lowest nine eigenvalues are converged (FP32: to single precision
    # of |H| ~ 13 Hartree)
    tol = 1e-11 if dtype == 'f64' else 2e-5
    assert np.abs(kpt.eps_n[:9] - w[:9]).max() < tol, kpt.eps_n[:9] - w[:9]
