"""CPU restatement of the RMM-DIIS eigensolver -- TEST INFRASTRUCTURE.

numpy version of gpaw/eigensolvers/rmmdiis.py:52-344 with
eigensolver.py:82-172 and overlap.py:29-101 for a dense symmetric H (the
oracle's finite-difference PAW Hamiltonian on a small grid) with S = 1 and
a preconditioner callable.  Used to check the device driver's
convergence behaviour; not bit-comparable (BLAS/LAPACK order).
"""
import numpy as np
import scipy.linalg


def orthonormalize(psi, dv):
    S = dv * psi @ psi.T
    L = np.linalg.cholesky(S)
    C = np.linalg.inv(L)
    return C @ psi


def subspace_diagonalize(H, psi, dv):
    Hpsi = psi @ H.T
    Hnn = dv * 0.5 * (psi @ Hpsi.T + Hpsi @ psi.T)
    eps, U = scipy.linalg.eigh(Hnn)
    return eps, U.T @ psi, U.T @ Hpsi


def rmmdiis_iteration(H, psi, dv, precondition, blocksize=10, niter=3,
                      trial_step=0.1):
    """One iteration; psi rows are bands on the flattened grid.  Returns
    (error, eps, new orthonormal psi)."""
    eps, psi, R = subspace_diagonalize(H, psi, dv)
    R = R - eps[:, None] * psi
    nb = psi.shape[0]
    error = 0.0
    for n1 in range(0, nb, blocksize):
        n2 = min(n1 + blocksize, nb)
        p = psi[n1:n2]
        r = R[n1:n2]
        e = eps[n1:n2]
        error += dv * np.sum(r * r)
        pd = [p.copy()]
        rd = [r.copy()]
        dp = precondition(r)
        dr = dp @ H.T - e[:, None] * dp
        lam = -np.sum(dr * r, axis=1) / np.sum(dr * dr, axis=1)
        p = p + lam[:, None] * dp
        r = r + lam[:, None] * dr
        for nit in range(1, niter):
            pd.append(p.copy())
            rd.append(r.copy())
            if nit > 1:
                m = nit + 1
                for ib in range(n2 - n1):
                    Rm = np.array([[dv * rd[i][ib] @ rd[j][ib]
                                    for j in range(m)] for i in range(m)])
                    A = -np.ones((m + 1, m + 1))
                    A[:m, :m] = Rm
                    A[-1, -1] = 0.0
                    x = np.zeros(m + 1)
                    x[-1] = -1.0
                    a = np.linalg.solve(A, x)[:-1]
                    p[ib] = sum(a[i] * pd[i][ib] for i in range(m))
                    r[ib] = sum(a[i] * rd[i][ib] for i in range(m))
            if nit < niter - 1:
                dp = precondition(r)
                p = p + lam[:, None] * dp
                r = p @ H.T - e[:, None] * p
        dp = precondition(r)
        p = p + trial_step * dp
        psi[n1:n2] = p
    return error, eps, orthonormalize(psi, dv)
