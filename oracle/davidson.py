"""CPU restatement of the Davidson eigensolver -- TEST INFRASTRUCTURE.

numpy version of gpaw/eigensolvers/davidson.py:68-262 with
eigensolver.py:82-172 for a dense symmetric H (the oracle's FD PAW
Hamiltonian on a small grid) with S = 1 and a preconditioner callable;
scipy.linalg.eigh(H, S) stands for general_diagonalize (dsygvd).  Pinned by
convergence to the dense spectrum (tests/test_oracle_math.py).
"""
import numpy as np
import scipy.linalg

from rmmdiis import subspace_diagonalize


def davidson_iteration(H, psi, dv, precondition, weights=None, niter=1,
                       normalize=True):
    """One Davidson sweep; psi rows are orthonormal bands.  Returns
    (error, eps, new psi)."""
    nb = psi.shape[0]
    eps, psi, Hpsi = subspace_diagonalize(H, psi, dv)
    R = Hpsi - eps[:, None] * psi
    w = np.ones(nb) if weights is None else np.asarray(weights)
    error = 0.0
    for nit in range(niter):
        error = float(np.sum(w * dv * np.sum(R * R, axis=1)))
        psi2 = precondition(R)
        if normalize:
            norm = dv * np.sum(psi2 * psi2, axis=1)
            psi2 = psi2 / np.sqrt(norm)[:, None]
        Hpsi2 = psi2 @ H.T
        H2 = np.zeros((2 * nb, 2 * nb))
        S2 = np.zeros((2 * nb, 2 * nb))
        H2[:nb, :nb] = np.diag(eps)
        S2[:nb, :nb] = np.eye(nb)
        H2[nb:, :nb] = dv * psi2 @ Hpsi.T
        H2[nb:, nb:] = dv * 0.5 * (psi2 @ Hpsi2.T + Hpsi2 @ psi2.T)
        S2[nb:, :nb] = dv * psi2 @ psi.T
        S2[nb:, nb:] = dv * psi2 @ psi2.T
        H2 = np.tril(H2) + np.tril(H2, -1).T
        S2 = np.tril(S2) + np.tril(S2, -1).T
        e2, V = scipy.linalg.eigh(H2, S2)
        X = V.T[:nb]
        psi = X[:, :nb] @ psi + X[:, nb:] @ psi2
        eps = e2[:nb]
        if nit < niter - 1:
            Hpsi = psi @ H.T
            R = Hpsi - eps[:, None] * psi
    return error, eps, psi
