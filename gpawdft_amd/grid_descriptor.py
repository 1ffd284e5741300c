"""Uniform real-space grids (serial).

Mirrors gpaw/grid_descriptor.py:71-412 and gpaw/domain.py:15-139 for one
domain: C-ordered grids with the last axis contiguous, `beg_c`/`end_c`
with the reference's convention that a non-periodic axis excludes the
boundary point (grid_descriptor.py:126-141), `neighbor_cd` (own rank on a
periodic axis, -1 otherwise, domain.py:111-139), `get_boxes` for
periodic images of spheres, coarsen/refine.  Units are Bohr.

Arrays returned by zeros()/empty() are device tensors of the same shape
as the reference's numpy arrays.
"""
import numbers

import numpy as np

from gpawdft_amd import device as _dev


class GridBoundsError(ValueError):
    pass


class SerialCommunicator:
    """A one-rank communicator (gpaw/mpi/__init__.py:576-660)."""
    size = 1
    rank = 0

    def sum(self, a, root=-1):
        if isinstance(a, (numbers.Number, np.number)):
            return a
        return None

    def max(self, a, root=-1):
        return a

    def barrier(self):
        pass

    def broadcast(self, a, root):
        pass

    def all_gather(self, a, b):
        b[:] = a

    def get_c_object(self):
        return None


serial_comm = SerialCommunicator()


class GridDescriptor:
    ndim = 3

    def __init__(self, N_c, cell_cv=(1, 1, 1), pbc_c=True, comm=None,
                 parsize_c=None):
        if isinstance(pbc_c, (int, bool, np.bool_)):
            pbc_c = (bool(pbc_c),) * 3
        if comm is None:
            comm = serial_comm
        if comm.size != 1:
            raise NotImplementedError(
                'domain decomposition is not part of this build (one GPU)')
        self.comm = comm
        self.N_c = np.array(N_c, int)
        if (self.N_c != N_c).any():
            raise ValueError('Non-int number of grid points %s' % N_c)

        cell_c = np.array(cell_cv, float)
        if cell_c.ndim == 1:
            self.cell_cv = np.diag(cell_c)
        else:
            self.cell_cv = cell_c
            cell_c = (self.cell_cv**2).sum(1)**0.5
        self.icell_cv = np.linalg.inv(self.cell_cv).T
        self.xxxucell_cv = np.array([self.cell_cv[x] / cell_c[x]
                                     for x in range(3)])
        self.xxxiucell_cv = np.linalg.inv(self.xxxucell_cv.T)
        self.pbc_c = np.asarray(pbc_c, bool)

        self.parsize_c = np.array([1, 1, 1])
        self.stride_c = np.array([1, 1, 1])
        self.parpos_c = np.array([0, 0, 0])
        self.rank = 0
        # own rank on periodic axes, no neighbour on the others
        self.neighbor_cd = np.zeros((3, 2), int)
        self.sdisp_cd = np.zeros((3, 2), int)
        for c in range(3):
            for d in range(2):
                if self.pbc_c[c]:
                    self.sdisp_cd[c, d] = -(2 * d - 1)
                else:
                    self.neighbor_cd[c, d] = -1

        self.beg_c = np.empty(3, int)
        self.end_c = np.empty(3, int)
        self.n_cp = []
        for c in range(3):
            n_p = np.array([0, self.N_c[c]])
            if not self.pbc_c[c]:
                n_p[0] = 1
            if not np.all(n_p[1:] - n_p[:-1] > 0):
                raise ValueError('Grid too small!')
            self.beg_c[c] = n_p[0]
            self.end_c[c] = n_p[1]
            self.n_cp.append(n_p)
        self.n_c = self.end_c - self.beg_c

        self.h_cv = self.cell_cv / self.N_c[:, np.newaxis]
        self.volume = abs(np.linalg.det(self.cell_cv))
        self.dv = self.volume / self.N_c.prod()
        self.orthogonal = not (self.cell_cv -
                               np.diag(self.cell_cv.diagonal())).any()
        h_c = self.get_grid_spacings()
        if max(h_c) / min(h_c) > 1.3:
            raise ValueError('Very anisotropic grid spacings: %s' % h_c)

    def __repr__(self):
        return 'GridDescriptor(%s, pbc_c=%s)' % (
            self.N_c.tolist(), self.pbc_c.astype(int).tolist())

    def new_descriptor(self, N_c=None, cell_cv=None, pbc_c=None, comm=None,
                       parsize_c=None):
        return self.__class__(self.N_c if N_c is None else N_c,
                              self.cell_cv if cell_cv is None else cell_cv,
                              self.pbc_c if pbc_c is None else pbc_c,
                              self.comm if comm is None else comm)

    def get_grid_spacings(self):
        L_c = (np.linalg.inv(self.cell_cv)**2).sum(0)**-0.5
        return L_c / self.N_c

    def get_size_of_global_array(self, pad=False):
        if pad:
            return self.N_c
        return self.N_c - 1 + self.pbc_c

    def flat_index(self, G_c):
        g1, g2, g3 = G_c - self.beg_c
        return g3 + self.n_c[2] * (g2 + g1 * self.n_c[1])

    def bytecount(self, dtype=float):
        return np.dtype(dtype).itemsize * self.n_c.prod()

    def _shape(self, n):
        if isinstance(n, numbers.Integral):
            n = (n,)
        return tuple(n) + tuple(int(x) for x in self.n_c)

    def zeros(self, n=(), dtype=float, global_array=False, pad=False,
              device=None):
        assert not global_array and not pad
        return _dev.zeros(self._shape(n), dtype, device)

    def empty(self, n=(), dtype=float, global_array=False, pad=False,
              device=None):
        assert not global_array and not pad
        return _dev.empty(self._shape(n), dtype, device)

    def coarsen(self):
        if (self.N_c % 2).any():
            raise ValueError('Grid %s not divisible by 2!' % self.N_c)
        return self.new_descriptor(self.N_c // 2)

    def refine(self):
        return self.new_descriptor(self.N_c * 2)

    def get_rank_from_position(self, spos_c):
        rnk_c = np.floor(np.asarray(spos_c) * self.parsize_c).astype(int)
        if (rnk_c < 0).any() or (rnk_c >= self.parsize_c).any():
            raise ValueError('Some atom is too close to the zero-boundary!')
        return 0

    def integrate(self, a_xg, b_yg=None, global_integral=True,
                  hermitian=False, _transposed_result=None):
        """Integral of a_xg (b_yg=None) or matrix of dv * <a_x|b_y>.

        The matrix form is provided by gpawdft_amd.hs_operators (MFMA
        kernels); the plain integral is a device reduction.
        """
        if b_yg is None:
            xshape = a_xg.shape[:-3]
            t = a_xg.reshape(tuple(xshape) + (-1,)).sum(-1) * self.dv
            return t.item() if t.ndim == 0 else t
        from gpawdft_amd.hs_operators import grid_inner_products
        return grid_inner_products(self, a_xg, b_yg, hermitian,
                                   _transposed_result)

    def get_boxes(self, spos_c, rcut, cut=True):
        """Boxes enclosing the sphere and its periodic images
        (gpaw/grid_descriptor.py:345-412)."""
        N_c = self.N_c
        ncut = rcut * (self.icell_cv**2).sum(axis=1)**0.5 * self.N_c
        npos_c = spos_c * N_c
        beg_c = np.ceil(npos_c - ncut).astype(int)
        end_c = np.ceil(npos_c + ncut).astype(int)

        if cut:
            for c in range(3):
                if not self.pbc_c[c]:
                    beg_c[c] = max(beg_c[c], 0)
                    end_c[c] = min(end_c[c], N_c[c])
        else:
            for c in range(3):
                if (not self.pbc_c[c] and
                        (beg_c[c] < 0 or end_c[c] > N_c[c])):
                    raise GridBoundsError(
                        'Box at %.3f %.3f %.3f crosses boundary.  Beg. of '
                        'box %s, end of box %s, max box size %s'
                        % (tuple(spos_c) + (beg_c, end_c, self.N_c)))

        # split each axis range at the cell boundaries
        range_c = ([], [], [])
        for c in range(3):
            b = beg_c[c]
            e = b
            while e < end_c[c]:
                b0 = b % N_c[c]
                e = min(end_c[c], b + N_c[c] - b0)
                b1 = b + self.beg_c[c] - b0 if b0 < self.beg_c[c] else b
                e0 = b0 - b + e
                e1 = e - (e0 - self.end_c[c]) if e0 > self.end_c[c] else e
                if e1 > b1:
                    range_c[c].append((b1, e1))
                b = e

        boxes = []
        for b0, e0 in range_c[0]:
            for b1, e1 in range_c[1]:
                for b2, e2 in range_c[2]:
                    b = np.array((b0, b1, b2))
                    e = np.array((e0, e1, e2))
                    beg = np.array((b0 % N_c[0], b1 % N_c[1], b2 % N_c[2]))
                    end = beg + e - b
                    disp = (b - beg) / N_c
                    beg = np.maximum(beg, self.beg_c)
                    end = np.minimum(end, self.end_c)
                    if (beg < end).all():
                        boxes.append((beg, end, disp))
        return boxes

