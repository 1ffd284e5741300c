"""Pin the oracle by mathematics independent of the reference: closed
forms, invariants and brute force.  Each check is chosen so that a
plausible slip (dropped term, wrong sign, wrong index, transposed axis)
fails it."""
import numpy as np
import pytest

import oracle as O
from gpawdft_amd.fd_operators import guc_laplace_stencil, laplace
from gpawdft_amd.grid_descriptor import GridDescriptor
from gpawdft_amd.transformers import transformer_paddings

Bohr = 0.52917721067


def transform(gdin, gdout, nn, a):
    pad, npad, skip, interp = transformer_paddings(gdin, gdout, nn)
    return O.transform(2 * nn, interp, pad, skip, gdout.n_c, gdin.pbc_c, a)


@pytest.mark.parametrize('nn', [1, 2, 3, 4, 6])
def test_fd_plane_wave_symbol(nn):
    """On a periodic grid plane waves are eigenvectors of the stencil with
    eigenvalue -1/2 sum_c h_c^-2 (a0 + 2 sum_j a_j cos(j theta_c))."""
    N = np.array([16, 20, 24])
    cell = np.array([3.1, 3.9, 4.6])
    gd = GridDescriptor(N, cell, True)
    coefs, offs, _ = guc_laplace_stencil(gd, -0.5, nn)
    k = np.array([3, -2, 5])
    x = [np.arange(n) for n in N]
    th = 2 * np.pi * k / N
    psi = np.cos(th[0] * x[0][:, None, None] + th[1] * x[1][None, :, None] +
                 th[2] * x[2][None, None, :] + 0.3)
    h = cell / N
    lam = 0.0
    for c in range(3):
        s = laplace[nn][0] + 2 * sum(laplace[nn][j] * np.cos(j * th[c])
                                     for j in range(1, nn + 1))
        lam += -0.5 * s / h[c]**2
    out = O.fd(coefs, offs, gd.pbc_c, psi[None])[0]
    assert np.abs(out - lam * psi).max() < 1e-10 * abs(lam)


@pytest.mark.parametrize('nn', [1, 2, 3])
def test_fd_polynomial_exactness(nn):
    """The O(h^2n) Laplacian is exact on polynomials of degree <= 2n+1 away
    from the zero boundary (cf. gpaw/test/fd_ops/laplace.py)."""
    N = 24
    gd = GridDescriptor((N, N, N), (6.0, 6.0, 6.0), False)
    coefs, offs, _ = guc_laplace_stencil(gd, 1.0, nn)
    h = 6.0 / N
    g = (np.arange(1, N) * h)
    X, Y, Z = np.meshgrid(g, g, g, indexing='ij')
    d = 2 * nn
    f = X**d + X * Y**(d - 1) + Z**2 * Y
    lap = d * (d - 1) * X**(d - 2) + X * (d - 1) * (d - 2) * Y**(d - 3) \
        + 2 * Y
    out = O.fd(coefs, offs, gd.pbc_c, f[None])[0]
    inner = (slice(nn, -nn),) * 3
    assert np.abs(out[inner] - lap[inner]).max() < 1e-7 * np.abs(lap).max()


@pytest.mark.parametrize('nn', [1, 2, 3, 4])
@pytest.mark.parametrize('pbc', [True, False])
def test_transformers_conserve_integral(nn, pbc):
    """Restriction and interpolation conserve the integral
    (gpaw/test/fd_ops/transformations.py asserts 1e-14)."""
    gd = GridDescriptor((20, 20, 20), (1.0, 1.0, 1.0), pbc)
    gdc = gd.coarsen()
    rng = np.random.default_rng(8)
    a = rng.standard_normal(tuple(gd.n_c))
    if not pbc:      # zero near the boundary so the halo is exact
        a[:2 * nn] = a[-2 * nn:] = 0
        a[:, :2 * nn] = a[:, -2 * nn:] = 0
        a[:, :, :2 * nn] = a[:, :, -2 * nn:] = 0
    r = transform(gd, gdc, nn, a[None])[0]
    assert abs(r.sum() * gdc.dv - a.sum() * gd.dv) < 1e-13 * np.abs(a).sum() * gd.dv
    if pbc:
        b = rng.standard_normal(tuple(gdc.n_c))
        i = transform(gdc, gd, nn, b[None])[0]
        assert abs(i.sum() * gd.dv - b.sum() * gdc.dv) < 1e-13 * np.abs(b).sum() * gdc.dv


@pytest.mark.parametrize('nn', [1, 2, 3, 4])
def test_interpolation_reproduces_polynomials(nn):
    """Interpolation with k = 2nn taps is exact for polynomials of degree
    < k; even (coincident) points are copied."""
    N = 32
    gd = GridDescriptor((N, N, N), (1.0, 1.0, 1.0), False)
    gdc = gd.coarsen()
    xc = np.arange(1, N // 2) * 2.0 / N
    xf = np.arange(1, N) * 1.0 / N
    deg = 2 * nn - 1
    fc = (xc[:, None, None]**deg + 0.5 * xc[None, :, None]**(deg - 1) +
          xc[None, None, :]) * np.ones((1, 1, 1))
    ff = (xf[:, None, None]**deg + 0.5 * xf[None, :, None]**(deg - 1) +
          xf[None, None, :])
    i = transform(gdc, gd, nn, fc[None])[0]
    m = 2 * nn + 1
    inner = (slice(m, -m),) * 3
    assert np.abs(i[inner] - ff[inner]).max() < 1e-12 * np.abs(ff).max()


def test_restriction_of_constants_and_linear():
    """Restriction weights sum to one; linear functions restrict to their
    values at the coarse points."""
    gd = GridDescriptor((16, 16, 16), (1.0, 1.0, 1.0), True)
    gdc = gd.coarsen()
    for nn in (1, 2, 3, 4):
        r = transform(gd, gdc, nn, np.full((1, 16, 16, 16), 2.5))
        assert np.abs(r - 2.5).max() < 1e-15
        x = np.arange(16)
        a = np.cos(2 * np.pi * x / 16)[:, None, None] * np.ones((16, 16, 16))
        r = transform(gd, gdc, nn, a[None])[0]
        # a smooth mode is damped by the restriction filter but stays a
        # cosine on the coarse grid, sampled at the even fine points
        ratio = r[:, 3, 5] / np.cos(2 * np.pi * 2 * np.arange(8) / 16)
        assert np.ptp(ratio[np.abs(np.cos(2 * np.pi * 2 * np.arange(8) / 16)) > 0.1]) < 1e-12


def test_spline_interpolates_knots():
    rc, rmax = 1.2 / Bohr, 2.5 / Bohr
    rg = np.linspace(0, rmax, 201)
    f = np.exp(-rg**2 / rc**2)
    s = O.Spline(0, rmax, f)
    dr = rmax / 200
    got = np.array([s(b * dr) for b in range(200)])
    ref = np.array([f[b] for b in range(200)])
    assert np.abs(got - ref).max() < 1e-15
    # smooth interpolant of a Gaussian between knots, zero past rcut
    r = np.random.default_rng(3).uniform(0, 0.95 * rmax, 100)
    assert np.abs(np.array([s(x) for x in r]) - np.exp(-r**2 / rc**2)).max() < 1e-6
    assert s(1.01 * rmax) == 0.0
    # zero slope at the origin
    assert abs(s(1e-6) - s(0.0)) < 1e-9


def real_solid_harmonics(l, x, y, z):
    """r^l Y_lm, m ordered -l..l, from the standard definitions."""
    r2 = x * x + y * y + z * z
    pi = np.pi
    if l == 0:
        return [np.full_like(x, 0.5 / np.sqrt(pi))]
    if l == 1:
        c = np.sqrt(3 / (4 * pi))
        return [c * y, c * z, c * x]
    c = np.sqrt(15 / (4 * pi))
    return [c * x * y, c * y * z, np.sqrt(5 / (16 * pi)) * (3 * z * z - r2),
            c * x * z, 0.5 * c * (x * x - y * y)]


def test_lfc_brute_force_and_adjointness():
    """add() of unit coefficients equals the sum over periodic images of
    f(r) r^l Y_lm evaluated directly at every grid point (pins images,
    boxes and harmonics), and <a, add(c)> = <integrate(a), c> / dv."""
    from gpawdft_amd.lfc import LFC
    from gpawdft_amd.spline import Spline
    N = 24
    L = N * 0.18 / Bohr
    gd = GridDescriptor((N, N, N), (L, L, L), True)
    rc, rmax = 1.2 / Bohr, 2.5 / Bohr
    rg = np.linspace(0, rmax, 201)
    f = np.exp(-rg**2 / rc**2)
    splines = [Spline(l, rmax, f) for l in range(3)]
    ospl = O.Spline(0, rmax, f)
    spos = np.array([[0.05, 0.5, 0.97], [0.6, 0.02, 0.3]])
    pt = LFC(gd, [splines] * 2)
    pt.set_positions(spos)
    g = np.arange(N) * L / N
    X, Y, Z = np.meshgrid(g, g, g, indexing='ij')
    # the sphere (diameter 9.4 Bohr) is larger than the cell (8.2 Bohr):
    # every image within rcut contributes
    shifts = [(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1)
              for k in (-1, 0, 1)]
    for a_ in range(2):
        M0 = pt.M_a[a_]
        want = np.zeros((9, N, N, N))
        for s in shifts:
            d = [X - (spos[a_, 0] + s[0]) * L, Y - (spos[a_, 1] + s[1]) * L,
                 Z - (spos[a_, 2] + s[2]) * L]
            r = np.sqrt(d[0]**2 + d[1]**2 + d[2]**2)
            rad = np.vectorize(ospl)(r)
            M = 0
            for l in range(3):
                for Y_ in real_solid_harmonics(l, *d):
                    want[M] += rad * Y_
                    M += 1
        for M in range(9):
            c = np.zeros((1, pt.Mmax))
            c[0, M0 + M] = 1.0
            got = O.lfc_add(pt.A_Wgm, pt.M_W, pt.G_B, pt.W_B, c,
                            np.zeros((1, N, N, N)))[0]
            assert np.abs(got - want[M]).max() < 1e-12
    rng = np.random.default_rng(4)
    a = rng.standard_normal((2, N, N, N))
    c = rng.standard_normal((2, pt.Mmax))
    lhs = (a * O.lfc_add(pt.A_Wgm, pt.M_W, pt.G_B, pt.W_B, c,
                         np.zeros_like(a))).sum()
    P = O.lfc_integrate(pt.A_Wgm, pt.M_W, pt.G_B, pt.W_B, gd.dv, a, pt.Mmax)
    assert abs(lhs - (P * c).sum() / gd.dv) < 1e-10 * abs(lhs)


def test_preconditioner_is_linear():
    from gpawdft_amd.transformers import transformer_paddings as tp
    gd0 = GridDescriptor((24, 24, 24), (8.0, 8.0, 8.0), True)
    gd1 = gd0.coarsen()
    gd2 = gd1.coarsen()
    c0, o0, _ = guc_laplace_stencil(gd0, -0.5, 3)
    c1, o1, _ = guc_laplace_stencil(gd1, -0.5, 1)
    c2, o2, _ = guc_laplace_stencil(gd2, -0.5, 1)

    def targs(a, b):
        pad, npad, skip, interp = tp(a, b, 1)
        return (2, pad, skip, a.n_c, b.n_c)

    args = ((c0, o0, gd0.n_c), (c1, o1, gd1.n_c), (c2, o2, gd2.n_c),
            targs(gd0, gd1), targs(gd1, gd2), targs(gd2, gd1),
            targs(gd1, gd0), gd0.pbc_c, 0.66666666 / c0[0])
    rng = np.random.default_rng(5)
    R1 = rng.standard_normal((1, 24, 24, 24))
    R2 = rng.standard_normal((1, 24, 24, 24))
    P1 = O.precondition(*args, R1)
    P2 = O.precondition(*args, R2)
    P12 = O.precondition(*args, 2.0 * R1 - 0.5 * R2)
    assert np.abs(P12 - (2.0 * P1 - 0.5 * P2)).max() < 1e-12 * np.abs(P12).max()
    # one V-cycle approximately solves T d = -R: for smooth R the
    # residual of that equation drops to ~0.6 |R| (one smoothing sweep per
    # level), and d points along -R
    x = np.arange(24) * 2 * np.pi / 24
    for kx in (1, 4):
        Rs = (np.cos(kx * x)[:, None, None] * np.cos(2 * x)[None, :, None] *
              np.ones(24))
        d = O.precondition(*args, Rs[None])[0]
        Td = O.fd(c0, o0, gd0.pbc_c, d[None])[0]
        assert np.linalg.norm(Td + Rs) < 0.7 * np.linalg.norm(Rs)
        assert (d * Rs).sum() < 0


def test_rmmdiis_restatement_converges_to_dense_spectrum():
    """The numpy RMM-DIIS (oracle/rmmdiis.py) started inside its basin
    converges the lowest bands to the dense eigenvalues of the oracle's FD
    PAW Hamiltonian; a wrong sign of the line-search step, a DIIS
    coefficient error or a dropped projector term would not."""
    import rmmdiis as RM
    from helpers import dense_hamiltonian, near_eigenvector_start, \
        oracle_preconditioner
    from gpawdft_amd.lfc import LFC
    from gpawdft_amd.synthetic import System
    sysm = System(12, 0.3, 2, seed=11, unit_projectors=True)
    pt = LFC(sysm.gd, [sysm.splines] * sysm.natoms)
    pt.set_positions(sysm.spos_ac)
    H = dense_hamiltonian(sysm, pt)
    w, psi = near_eigenvector_start(H, 8, sysm.gd)
    pc = oracle_preconditioner(sysm.gd)
    psi = RM.orthonormalize(psi, sysm.gd.dv)
    errs = []
    for it in range(25):
        err, eps, psi = RM.rmmdiis_iteration(H, psi, sysm.gd.dv, pc,
                                             blocksize=4)
        errs.append(err)
    assert errs[-1] < 1e-7 * errs[0]
    assert np.abs(eps[:6] - w[:6]).max() < 1e-12
    S = sysm.gd.dv * psi @ psi.T
    assert np.abs(S - np.eye(8)).max() < 1e-12


def test_local_potential_matches_the_reference_expression():
    """orc_add_local_potential is bit-for-bit the reference's numpy
    expression `Htpsit_G += psit_G * vt_G` (gpaw/hamiltonian.py:419-424):
    one rounded product, one rounded sum."""
    rng = np.random.default_rng(41)
    psi = rng.standard_normal((3, 6, 5, 4))
    v = rng.standard_normal((6, 5, 4))
    h = rng.standard_normal((3, 6, 5, 4))
    ref = h.copy()
    for n in range(3):
        ref[n] += psi[n] * v
    got = O.add_local_potential(psi, v, h)
    assert np.array_equal(got.view(np.int64), ref.view(np.int64))


def test_projector_coefs_is_P_times_dH():
    """c_a = P_a dH_a (gpaw/hamiltonian.py:456-458, np.dot(P_xi, dH_ii)),
    checked against numpy with a non-symmetric dH so that a transposed
    operand fails; the reference's BLAS order is not pinned, so to 1e-14."""
    rng = np.random.default_rng(42)
    P = rng.standard_normal((5, 14))
    dH0 = rng.standard_normal((9, 9))
    dH1 = rng.standard_normal((5, 5))
    c = O.projector_coefs(P, [(0, dH0), (9, dH1)])
    ref = np.zeros_like(P)
    ref[:, 0:9] = P[:, 0:9] @ dH0
    ref[:, 9:14] = P[:, 9:14] @ dH1
    assert np.abs(c - ref).max() < 1e-14 * np.abs(ref).max()
    assert np.abs(c[:, :9] - P[:, :9] @ dH0.T).max() > 1e-3


def test_poisson_restatement_converges_to_the_fd_solution():
    """The oracle's multigrid Poisson solver (oracle/poisson.py) converges
    to the Fourier-space solution of the same FD equation
    -1/(4 pi) L phi = rho - <rho> on a periodic grid (zero mean)."""
    import poisson as P
    gd = GridDescriptor((16, 16, 16), (6.0, 6.0, 6.0), True)
    rng = np.random.default_rng(51)
    x = np.arange(16) * 2 * np.pi / 16
    rho = (np.cos(x)[:, None, None] * np.sin(2 * x)[None, :, None] *
           np.cos(3 * x)[None, None, :] + 0.1 * rng.standard_normal((16,) * 3))
    rho -= rho.mean()
    ps = P.Poisson(gd, eps=1e-24)
    phi = np.zeros((16, 16, 16))
    niter = ps.solve_neutral(phi, rho)
    assert niter < 200
    c, o, _ = guc_laplace_stencil(gd, -0.25 / np.pi, 3)
    k = [2 * np.pi * np.fft.fftfreq(16)] * 3
    lam = 0.0
    for ax in range(3):
        sel = [i for i, oo in enumerate(o) if oo[ax] != 0]
        sym = sum(c[i] * np.cos(o[i][ax] * k[ax]) for i in sel)
        shape = [1, 1, 1]
        shape[ax] = 16
        lam = lam + sym.reshape(shape)
    lam = lam + c[0]
    F = np.fft.fftn(rho)
    lam.flat[0] = 1.0
    F.flat[0] = 0.0
    phi_ref = np.real(np.fft.ifftn(F / lam))
    assert np.abs(phi - phi_ref).max() < 1e-9 * np.abs(phi_ref).max()


def test_davidson_restatement_converges_to_dense_spectrum():
    """The numpy Davidson (oracle/davidson.py) from a start inside its basin
    converges the lowest bands to the dense eigenvalues of the oracle's FD
    PAW Hamiltonian, and keeps the bands orthonormal."""
    import davidson as DV
    import rmmdiis as RM
    from helpers import dense_hamiltonian, near_eigenvector_start, \
        oracle_preconditioner
    from gpawdft_amd.lfc import LFC
    from gpawdft_amd.synthetic import System
    sysm = System(12, 0.3, 2, seed=11, unit_projectors=True)
    pt = LFC(sysm.gd, [sysm.splines] * sysm.natoms)
    pt.set_positions(sysm.spos_ac)
    H = dense_hamiltonian(sysm, pt)
    w, psi = near_eigenvector_start(H, 8, sysm.gd)
    pc = oracle_preconditioner(sysm.gd)
    psi = RM.orthonormalize(psi, sysm.gd.dv)
    errs = []
    for it in range(30):
        err, eps, psi = DV.davidson_iteration(H, psi, sysm.gd.dv, pc)
        errs.append(err)
    # niter = 1 Davidson converges more slowly than RMM-DIIS; the top
    # bands lag
    assert errs[-1] < 1e-5 * errs[0]
    assert np.abs(eps[:5] - w[:5]).max() < 1e-9
    S = sysm.gd.dv * psi @ psi.T
    assert np.abs(S - np.eye(8)).max() < 1e-10


def test_complex_stencil_bloch_plane_waves():
    """With phases exp(2 pi i k_c) for the image above and exp(-2 pi i k_c)
    below, a plane wave exp(i (k + G) . r) sampled on a periodic grid is an
    eigenvector of the FD operator with eigenvalue
    sum_c coef_c exp(i (k + G) . h off_c): checks the phase convention and
    direction independently of the reference (orthorhombic and hexagonal
    cells)."""
    import oracle as O
    from gpawdft_amd.fd_operators import guc_laplace_stencil
    from gpawdft_amd.grid_descriptor import GridDescriptor
    for N, cell, nn in (((12, 10, 14), np.diag([4.0, 3.5, 5.0]), 3),
                        ((10, 10, 12), np.array([[4.0, 0, 0],
                                                 [2.0, 3.4641016, 0],
                                                 [0, 0, 5.0]]), 2)):
        gd = GridDescriptor(N, cell, True)
        c, o, _ = guc_laplace_stencil(gd, -0.5, nn)
        k_c = np.array([0.25, -0.1, 0.4])
        G_c = np.array([1, -2, 3])
        phase = np.exp(2j * np.pi * np.array([[1, -1]] * 3) * k_c[:, None])
        idx = np.indices(N).reshape(3, -1).T            # grid indices
        frac = idx / np.array(N)                        # scaled positions
        psi = np.exp(2j * np.pi * frac @ (k_c + G_c)).reshape(N)
        out = O.fdz(c, o, gd.pbc_c, phase, psi[None])[0]
        off = np.asarray(o).reshape(-1, 3) / np.array(N)
        lam = np.sum(np.asarray(c) * np.exp(2j * np.pi * off @ (k_c + G_c)))
        assert np.abs(out - lam * psi).max() < 1e-11 * np.abs(lam)
        # the wrong direction convention (phases swapped) breaks it
        bad = O.fdz(c, o, gd.pbc_c, phase[:, ::-1], psi[None])[0]
        assert np.abs(bad - lam * psi).max() > 1e-3 * np.abs(lam)


def test_complex_lfc_adjoint_with_phases():
    """<add(c), a> = (1/dv) <c, integrate(a)> with the k-point phases: the
    add carries conj(phase) exactly where integrate carries the phase; a
    plain phase on both sides fails it."""
    import types
    import oracle as O
    from gpawdft_amd.lfc import LFC
    from gpawdft_amd.synthetic import make_grid, projector_splines
    gd = make_grid(20, 0.2)
    pt = LFC(gd, [projector_splines(2)] * 2,
             kd=types.SimpleNamespace(gamma=False,
                                      ibzk_qc=np.array([[0.3, -0.2, 0.1]])),
             dtype=complex)
    pt.set_positions(np.array([[0.1, 0.2, 0.3], [0.6, 0.9, 0.4]]))
    rng = np.random.default_rng(3)
    a = rng.standard_normal((1, 20, 20, 20)) + 1j * rng.standard_normal(
        (1, 20, 20, 20))
    c = rng.standard_normal((1, pt.Mmax)) + 1j * rng.standard_normal(
        (1, pt.Mmax))
    ph = pt.phase_qW[0]
    assert np.abs(ph - 1).max() > 0.1          # images carry real phases
    ci = O.lfc_integrate_z(pt.A_Wgm, pt.M_W, pt.G_B, pt.W_B, gd.dv, a,
                           pt.Mmax, ph)
    b = O.lfc_add_z(pt.A_Wgm, pt.M_W, pt.G_B, pt.W_B, c,
                    np.zeros_like(a), ph)
    lhs = np.vdot(b, a)
    rhs = np.vdot(c, ci) / gd.dv
    assert abs(lhs - rhs) < 1e-12 * abs(rhs)
    bad = O.lfc_add_z(pt.A_Wgm, pt.M_W, pt.G_B, pt.W_B, c,
                      np.zeros_like(a), ph.conj())
    assert abs(np.vdot(bad, a) - rhs) > 1e-3 * abs(rhs)
