"""CPU oracle for the grid Hamiltonian path -- TEST INFRASTRUCTURE ONLY.

numpy front end of liboracle.so (oracle.c), the FP64 C restatement of the
reference's serial algorithms.  Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg use it, as the checker; the product never
imports it.

Each function names the reference code it restates; see oracle.c's header
for how each one is pinned (reference build, golden fixtures, closed forms).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'liboracle.so')

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)
_ip = ctypes.POINTER(ctypes.c_int)

_lib = None


def build():
    subprocess.run(['make', '-s', '-C', HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        i, l, d, v = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p
        sig = {
            'orc_halo_fill': (None, [_dp, _lp, _lp, _lp, _ip, i, _dp]),
            'orc_fd': (None, [i, _dp, _lp, _lp, _ip, l, _dp, _dp]),
            'orc_fdz': (None, [i, _dp, _lp, _lp, _ip, _dp, l, _dp, _dp]),
            'orc_halo_fill_z': (None, [_dp, _lp, _lp, _lp, _ip, i, _dp, _dp]),
            'orc_relax': (None, [i, i, _dp, _lp, _lp, _ip, l, _dp, _dp, i,
                                 d]),
            'orc_transform': (None, [i, i, _lp, _lp, _lp, _ip, _ip, _lp, l,
                                     _dp, _dp]),
            'orc_transform_z': (None, [i, i, _lp, _lp, _lp, _ip, _ip, _lp,
                                       _dp, l, _dp, _dp]),
            'orc_spline_build': (None, [d, i, _dp, _dp]),
            'orc_spline_value': (d, [_dp, i, d, d]),
            'orc_spline_to_grid': (l, [_dp, i, d, i, _lp, _lp, _dp, _dp, _lp,
                                       _lp, _dp, _ip, _lp]),
            'orc_lfc_integrate': (None, [i, _dp, _lp, _ip, _ip, l, _ip, _ip,
                                         d, l, l, _dp, l, _dp]),
            'orc_lfc_add': (None, [i, _dp, _lp, _ip, _ip, l, _ip, _ip, l, l,
                                   _dp, l, _dp]),
            'orc_lfc_integrate_z': (None, [i, _dp, _lp, _ip, _ip, l, _ip, _ip,
                                           d, l, l, _dp, l, _dp, _dp]),
            'orc_lfc_add_z': (None, [i, _dp, _lp, _ip, _ip, l, _ip, _ip, l,
                                     l, _dp, l, _dp, _dp]),
            'orc_add_local_potential': (None, [l, l, _dp, _dp, _dp]),
            'orc_projector_coefs': (None, [l, l, i, _ip, _ip, _lp, _dp, _dp,
                                           _dp]),
            'orc_precondition': (None, [v, v, v, v, v, v, v, _ip, d, l, _dp,
                                        _dp]),
            'orc_num_threads': (i, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _l(a):
    return a.ctypes.data_as(_lp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.intc)


def num_threads():
    return lib().orc_num_threads()


def fd(coefs, offsets_c, pbc_c, in_xg):
    """FD operator apply (c/operators.c:109-157, c/bmgs/fd.c:16-87).
    offsets_c: (ncoefs, 3) integer offsets."""
    in_xg = _f64(in_xg)
    n = _i64(in_xg.shape[-3:])
    nb = int(np.prod(in_xg.shape[:-3], dtype=np.int64))
    out = np.empty_like(in_xg)
    coefs = _f64(coefs)
    off = _i64(np.asarray(offsets_c).reshape(-1, 3))
    pbc = _i32(np.asarray(pbc_c, bool).astype(int))
    lib().orc_fd(len(coefs), _d(coefs), _l(off), _l(n), _i(pbc), nb,
                 _d(in_xg), _d(out))
    return out


def fdz(coefs, offsets_c, pbc_c, phase_cd, in_xg):
    """Complex FD operator apply with Bloch phases (c/operators.c:109-157
    with phases, c/bc.c:227-247, c/bmgs/fd.c:16-87 for double_complex).
    phase_cd: (3, 2) complex, [axis][0] for the image above, [axis][1]
    below."""
    in_xg = np.ascontiguousarray(in_xg, np.complex128)
    n = _i64(in_xg.shape[-3:])
    nb = int(np.prod(in_xg.shape[:-3], dtype=np.int64))
    out = np.empty_like(in_xg)
    coefs = _f64(coefs)
    off = _i64(np.asarray(offsets_c).reshape(-1, 3))
    pbc = _i32(np.asarray(pbc_c, bool).astype(int))
    ph = np.ascontiguousarray(np.asarray(phase_cd, np.complex128)
                              .reshape(3, 2)).view(np.float64)
    lib().orc_fdz(len(coefs), _d(coefs), _l(off), _l(n), _i(pbc), _d(ph), nb,
                  _d(in_xg.view(np.float64)), _d(out.view(np.float64)))
    return out


def relax(method, coefs, offsets_c, pbc_c, fun_xg, src_xg, nrelax, w=1.0):
    """Operator.relax (c/operators.c:49-91, c/bmgs/relax.c): returns the
    relaxed copy of fun_xg."""
    fun = _f64(fun_xg).copy()
    src = _f64(src_xg)
    assert src.shape == fun.shape
    n = _i64(fun.shape[-3:])
    nb = int(np.prod(fun.shape[:-3], dtype=np.int64))
    offs = _i64(np.asarray(offsets_c).reshape(-1))
    pbc = _i32(np.asarray(pbc_c, bool).astype(int))
    lib().orc_relax(int(method), len(coefs), _d(_f64(coefs)), _l(offs),
                    _l(n), _i(pbc), nb, _d(fun), _d(src), int(nrelax),
                    float(w))
    return fun


def transform(k, interpolate, pad_cd, skip_cd, nout_c, pbc_c, in_xg):
    """Transformer apply (c/transformers.c:54-123, c/bmgs/restrict.c,
    c/bmgs/interpolate.c).  pad_cd/skip_cd: (3, 2)."""
    in_xg = _f64(in_xg)
    n = _i64(in_xg.shape[-3:])
    pad = np.asarray(pad_cd, int)
    lo = _i64(pad[:, 0])
    hi = _i64(pad[:, 1])
    skip = _i32(np.asarray(skip_cd, int).reshape(-1))
    nout = _i64(nout_c)
    nb = int(np.prod(in_xg.shape[:-3], dtype=np.int64))
    out = np.empty(in_xg.shape[:-3] + tuple(int(x) for x in nout))
    pbc = _i32(np.asarray(pbc_c, bool).astype(int))
    lib().orc_transform(k, int(interpolate), _l(n), _l(lo), _l(hi), _i(pbc),
                        _i(skip), _l(nout), nb, _d(in_xg), _d(out))
    return out


def transform_z(k, interpolate, pad_cd, skip_cd, nout_c, pbc_c, phase_cd,
                in_xg):
    """Complex Transformer apply with Bloch phases (c/transformers.c:54-123
    with bmgs_restrictz / bmgs_interpolatez)."""
    in_xg = np.ascontiguousarray(in_xg, np.complex128)
    n = _i64(in_xg.shape[-3:])
    pad = np.asarray(pad_cd, int)
    skip = _i32(np.asarray(skip_cd, int).reshape(-1))
    nout = _i64(nout_c)
    nb = int(np.prod(in_xg.shape[:-3], dtype=np.int64))
    out = np.empty(in_xg.shape[:-3] + tuple(int(x) for x in nout), complex)
    pbc = _i32(np.asarray(pbc_c, bool).astype(int))
    ph = np.ascontiguousarray(np.asarray(phase_cd, np.complex128)
                              .reshape(3, 2)).view(np.float64)
    lib().orc_transform_z(k, int(interpolate), _l(n), _l(_i64(pad[:, 0])),
                          _l(_i64(pad[:, 1])), _i(pbc), _i(skip), _l(nout),
                          _d(ph), nb, _d(in_xg.view(np.float64)),
                          _d(out.view(np.float64)))
    return out


def halo_fill(a, lo_c, hi_c, pbc_c, corners):
    """Padded copy with periodic images (c/bc.c:138-250)."""
    a = _f64(a)
    n = _i64(a.shape)
    lo = _i64(lo_c)
    hi = _i64(hi_c)
    out = np.empty(tuple(int(x) for x in n + lo + hi))
    pbc = _i32(np.asarray(pbc_c, bool).astype(int))
    lib().orc_halo_fill(_d(a), _l(n), _l(lo), _l(hi), _i(pbc), int(corners),
                        _d(out))
    return out


class Spline:
    """Cubic radial spline (c/bmgs/spline.c:11-58)."""

    def __init__(self, l, rcut, f_g):
        f = _f64(f_g).copy()
        f[-1] = 0.0
        self.l = l
        self.nbins = len(f) - 1
        self.dr = rcut / self.nbins
        self.data = np.empty(4 * (self.nbins + 1))
        lib().orc_spline_build(ctypes.c_double(rcut), len(f), _d(f),
                               _d(self.data))

    def __call__(self, r):
        return lib().orc_spline_value(_d(self.data), self.nbins,
                                      ctypes.c_double(self.dr),
                                      ctypes.c_double(r))

    def get_cutoff(self):
        return self.dr * self.nbins

    def to_grid(self, beg_c, end_c, pos_v, h_cv, n_c, corner_c):
        """spline_to_grid (c/lfc.c:785-875) -> (A_gm, G_B)."""
        beg = _i64(beg_c)
        end = _i64(end_c)
        nbox = int(np.prod(end - beg))
        nm = 2 * self.l + 1
        A = np.empty((max(nbox, 1), nm))
        G = np.empty(2 * max(int((end[0] - beg[0]) * (end[1] - beg[1])), 1),
                     np.intc)
        nB = ctypes.c_long()
        ng = lib().orc_spline_to_grid(
            _d(self.data), self.nbins, ctypes.c_double(self.dr), self.l,
            _l(beg), _l(end), _d(_f64(pos_v)), _d(_f64(h_cv)), _l(_i64(n_c)),
            _l(_i64(corner_c)), _d(A), _i(G), ctypes.byref(nB))
        return A[:ng].copy(), G[:nB.value].copy()


def _lfc_tables(A_Wgm, M_W):
    ng_W = [A.shape[0] for A in A_Wgm]
    nm_W = _i32([A.shape[1] for A in A_Wgm])
    A_off = _i64(np.concatenate([[0], np.cumsum([A.size for A in A_Wgm])]))
    A = (_f64(np.concatenate([A.ravel() for A in A_Wgm])) if A_Wgm
         else np.zeros(1))
    return A, A_off, nm_W, _i32(M_W), ng_W


def lfc_integrate(A_Wgm, M_W, G_B, W_B, dv, a_xG, Mmax, c_xM=None):
    """LFC integrate (c/lfc.c:361-425): c_xM += dv sum_G a A."""
    a_xG = _f64(a_xG)
    nx = int(np.prod(a_xG.shape[:-3], dtype=np.int64))
    nG = int(np.prod(a_xG.shape[-3:]))
    if c_xM is None:
        c_xM = np.zeros((nx, Mmax))
    c_xM = _f64(c_xM).copy()
    A, A_off, nm_W, M_W, _ = _lfc_tables(A_Wgm, M_W)
    G_B = _i32(G_B)
    W_B = _i32(W_B)
    lib().orc_lfc_integrate(len(A_Wgm), _d(A), _l(A_off), _i(nm_W), _i(M_W),
                            len(G_B), _i(G_B), _i(W_B), ctypes.c_double(dv),
                            nx, nG, _d(a_xG), Mmax, _d(c_xM))
    return c_xM


def lfc_add(A_Wgm, M_W, G_B, W_B, c_xM, a_xG):
    """LFC add (c/lfc.c:724-783): returns a_xG + sum_M A c."""
    a_xG = _f64(a_xG).copy()
    nx = int(np.prod(a_xG.shape[:-3], dtype=np.int64))
    nG = int(np.prod(a_xG.shape[-3:]))
    c_xM = _f64(c_xM).reshape(nx, -1)
    A, A_off, nm_W, M_W, _ = _lfc_tables(A_Wgm, M_W)
    G_B = _i32(G_B)
    W_B = _i32(W_B)
    lib().orc_lfc_add(len(A_Wgm), _d(A), _l(A_off), _i(nm_W), _i(M_W),
                      len(G_B), _i(G_B), _i(W_B), nx, nG, _d(c_xM),
                      c_xM.shape[1], _d(a_xG))
    return a_xG


def _c128(a):
    return np.ascontiguousarray(a, np.complex128)


def lfc_integrate_z(A_Wgm, M_W, G_B, W_B, dv, a_xG, Mmax, phase_W,
                    c_xM=None):
    """Complex LFC integrate with per-volume phases (c/lfc.c:400-423):
    c_xM += sum_G a (phase dv) A."""
    a_xG = _c128(a_xG)
    nx = int(np.prod(a_xG.shape[:-3], dtype=np.int64))
    nG = int(np.prod(a_xG.shape[-3:]))
    c_xM = (np.zeros((nx, Mmax), complex) if c_xM is None
            else _c128(c_xM).copy())
    A, A_off, nm_W, M_W, _ = _lfc_tables(A_Wgm, M_W)
    ph = _c128(phase_W)
    lib().orc_lfc_integrate_z(
        len(A_Wgm), _d(A), _l(A_off), _i(nm_W), _i(M_W), len(G_B),
        _i(_i32(G_B)), _i(_i32(W_B)), ctypes.c_double(dv), nx, nG,
        _d(a_xG.view(np.float64)), Mmax, _d(c_xM.view(np.float64)),
        _d(ph.view(np.float64)))
    return c_xM


def lfc_add_z(A_Wgm, M_W, G_B, W_B, c_xM, a_xG, phase_W):
    """Complex LFC add with per-volume phases (c/lfc.c:758-781): returns
    a_xG + sum_M A c conj(phase)."""
    a_xG = _c128(a_xG).copy()
    nx = int(np.prod(a_xG.shape[:-3], dtype=np.int64))
    nG = int(np.prod(a_xG.shape[-3:]))
    c_xM = _c128(c_xM).reshape(nx, -1)
    A, A_off, nm_W, M_W, _ = _lfc_tables(A_Wgm, M_W)
    ph = _c128(phase_W)
    lib().orc_lfc_add_z(
        len(A_Wgm), _d(A), _l(A_off), _i(nm_W), _i(M_W), len(G_B),
        _i(_i32(G_B)), _i(_i32(W_B)), nx, nG, _d(c_xM.view(np.float64)),
        c_xM.shape[1], _d(a_xG.view(np.float64)), _d(ph.view(np.float64)))
    return a_xG


def add_local_potential(psi_xG, v_G, hpsi_xG):
    """Htpsit += psit * vt per band (gpaw/hamiltonian.py:397-424)."""
    psi = _f64(psi_xG)
    h = _f64(hpsi_xG).copy()
    nG = int(np.prod(psi.shape[-3:]))
    nb = int(np.prod(psi.shape[:-3], dtype=np.int64))
    lib().orc_add_local_potential(nb, nG, _d(psi), _d(_f64(v_G)), _d(h))
    return h


def projector_coefs(P_xM, atoms):
    """c = P dH per atom; atoms = [(M_a, dH_ii), ...], sums in i order."""
    P = _f64(P_xM)
    nb, nM = P.shape
    M_a = _i32([a[0] for a in atoms])
    ni_a = _i32([len(a[1]) for a in atoms])
    blocks = [_f64(a[1]).ravel() for a in atoms]
    off = _i64(np.concatenate([[0], np.cumsum([b.size for b in blocks])])[:-1])
    dH = _f64(np.concatenate(blocks))
    c = np.zeros_like(P)
    lib().orc_projector_coefs(nb, nM, len(atoms), _i(M_a), _i(ni_a), _l(off),
                              _d(dH), _d(P), _d(c))
    return c


class _Stencil(ctypes.Structure):
    _fields_ = [('ncoefs', ctypes.c_int), ('coefs', _dp), ('off_c', _lp),
                ('n', ctypes.c_long * 3)]


class _Transformer(ctypes.Structure):
    _fields_ = [('k', ctypes.c_int), ('interpolate', ctypes.c_int),
                ('n', ctypes.c_long * 3), ('lo', ctypes.c_long * 3),
                ('hi', ctypes.c_long * 3), ('skip', ctypes.c_int * 6),
                ('nout', ctypes.c_long * 3)]


def precondition(kin0, kin1, kin2, rst0, rst1, ip2, ip1, pbc_c, step, R_xG):
    """Multigrid V-cycle (gpaw/preconditioner.py:34-60).

    kin*: (coefs, offsets_c, n_c); rst*/ip*: (k, pad_cd, skip_cd, n_in,
    n_out).  Returns d0 for the residuals R_xG."""
    keep = []

    def stencil(s):
        coefs = _f64(s[0])
        off = _i64(np.asarray(s[1]).reshape(-1, 3))
        keep.extend([coefs, off])
        st = _Stencil()
        st.ncoefs = len(coefs)
        st.coefs = _d(coefs)
        st.off_c = _l(off)
        for c in range(3):
            st.n[c] = int(s[2][c])
        return st

    def transformer(t, interp):
        k, pad_cd, skip_cd, n_in, n_out = t
        tr = _Transformer()
        tr.k = k
        tr.interpolate = int(interp)
        pad = np.asarray(pad_cd, int)
        sk = np.asarray(skip_cd, int).reshape(-1)
        for c in range(3):
            tr.n[c] = int(n_in[c])
            tr.lo[c] = int(pad[c, 0])
            tr.hi[c] = int(pad[c, 1])
            tr.nout[c] = int(n_out[c])
        for i in range(6):
            tr.skip[i] = int(sk[i])
        return tr

    R = _f64(R_xG)
    nb = int(np.prod(R.shape[:-3], dtype=np.int64))
    out = np.empty_like(R)
    k0, k1, k2 = stencil(kin0), stencil(kin1), stencil(kin2)
    t0, t1 = transformer(rst0, False), transformer(rst1, False)
    t2, t3 = transformer(ip2, True), transformer(ip1, True)
    pbc = _i32(np.asarray(pbc_c, bool).astype(int))
    byref = ctypes.byref
    lib().orc_precondition(byref(k0), byref(k1), byref(k2), byref(t0),
                           byref(t1), byref(t2), byref(t3), _i(pbc),
                           ctypes.c_double(step), nb, _d(R), _d(out))
    return out
