"""Ragged batches: many independent grid systems in one graph launch.

No counterpart in the reference (it runs one system per process).  A
batch holds systems of different grid sizes, band counts and atom counts;
one "step" applies, to every band of every system, H psi (projections,
P dH, stencil + v psi, projector add) and the multigrid preconditioner.

* `run_sequential()` launches system after system from Python (the
  per-system baseline).
* `capture()` records the same launches of all systems into one hipGraph
  (torch's graph object is only the recorder of the streams; the kernels
  are this library's), and `replay()` runs the whole batch with one launch.
  Systems are spread over `nstreams` capture streams (largest first, each
  to the least-loaded stream), which become parallel branches of the
  graph, so the kernels of small systems run side by side.
* `autotune()` picks, per distinct (grid, band count), the stencil's
  workgroup target (axis-0 chunk depth) from a few timed candidates.

Systems are FP32 and built with the synthetic generators.
"""
import time

import numpy as np
import torch

from gpawdft_amd.hamiltonian import Hamiltonian
from gpawdft_amd.preconditioner import Preconditioner
from gpawdft_amd.synthetic import System, smooth_bands_device

WG_TARGETS = (2048, 4096, 8192, 16384)


class _Member:
    def __init__(self, N, nb, natoms, seed, block):
        self.sysm = System(N, 0.18, natoms, seed=seed)
        gd = self.sysm.gd
        self.nb = nb
        self.G = int(np.prod(gd.n_c))
        self.kin, self.pt = self.sysm.operators()
        self.ham = Hamiltonian(gd, [self.sysm.vt_G], self.sysm.dH_asp)
        self.pc = Preconditioner(gd, self.kin, block=block)
        self.block = block
        self.psi = smooth_bands_device(nb, gd.n_c, seed + 7,
                                       dtype=torch.float32)
        self.out = torch.empty_like(self.psi)
        self.d = torch.empty_like(self.psi)
        self.P = torch.zeros((nb, self.pt.Mmax), dtype=torch.float32,
                             device='cuda')
        self.c = torch.empty_like(self.P)

    def step(self):
        self.ham.apply_fused(self.psi, self.out, self.kin, self.pt,
                             P_xM=self.P, c_xM=self.c)
        for b0 in range(0, self.nb, self.block):
            b1 = min(b0 + self.block, self.nb)
            self.pc(self.psi[b0:b1], None, out=self.d[b0:b1])


def random_specs(nsys=48, seed=0, sizes=(32, 48, 64, 80), bands=(16, 256),
                 atoms=(4, 48)):
    """(N, nb, natoms, seed) of each system: grids drawn from `sizes`,
    band counts U[bands], atom counts U[atoms]."""
    rng = np.random.default_rng(seed)
    return [(int(rng.choice(sizes)), int(rng.integers(bands[0], bands[1] + 1)),
             int(rng.integers(atoms[0], atoms[1] + 1)), 100 + i)
            for i in range(nsys)]


class RaggedBatch:
    def __init__(self, specs, block=16):
        self.members = [_Member(N, nb, na, sd, block)
                        for N, nb, na, sd in specs]
        self.graph = None
        self.tuning = {}

    @property
    def band_points(self):
        return sum(m.nb * m.G for m in self.members)

    def run_sequential(self):
        for m in self.members:
            m.step()

    def outputs(self):
        return [(m.out.clone(), m.d.clone()) for m in self.members]

    def autotune(self, reps=3):
        """Per distinct (grid, nb): time the FP32 H psi of one member with
        each workgroup target and keep the fastest for all such members."""
        groups = {}
        for m in self.members:
            groups.setdefault((m.G, m.nb), []).append(m)
        for key, ms in groups.items():
            m = ms[0]
            best = None
            for target in WG_TARGETS:
                m.kin.set_launch(target)
                m.step()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    m.ham.apply_fused(m.psi, m.out, m.kin, m.pt,
                                      P_xM=m.P, c_xM=m.c)
                torch.cuda.synchronize()
                t = time.perf_counter() - t0
                if best is None or t < best[0]:
                    best = (t, target)
            for mm in ms:
                mm.kin.set_launch(best[1])
            self.tuning[key] = best[1]
        return self.tuning

    def capture(self, nstreams=4):
        """Record one step of every member into a single graph, members
        spread over nstreams parallel branches.  The members run once
        first (lazy device tables, scratch)."""
        self.run_sequential()
        torch.cuda.synchronize()
        order = sorted(range(len(self.members)),
                       key=lambda i: -self.members[i].nb * self.members[i].G)
        load = [0] * nstreams
        branch = [[] for _ in range(nstreams)]
        for i in order:
            k = int(np.argmin(load))
            branch[k].append(self.members[i])
            load[k] += self.members[i].nb * self.members[i].G
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        side = [torch.cuda.Stream() for _ in range(nstreams)]
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                for st, ms in zip(side, branch):
                    st.wait_stream(s)
                    with torch.cuda.stream(st):
                        for m in ms:
                            m.step()
                for st in side:
                    s.wait_stream(st)
        torch.cuda.current_stream().wait_stream(s)
        self.graph = g
        self.nstreams = nstreams

    def replay(self):
        self.graph.replay()
