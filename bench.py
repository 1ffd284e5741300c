"""Benchmark: batched H psi on one MI355X.

Workload (the headline configuration): periodic 96^3 grid, h = 0.18 A,
256 bands of smoothed Gaussian noise, local potential of 32 Gaussian wells,
32 atoms with l = 0..2 Gaussian projectors (9 per atom) and random
symmetric dH.  One step = H psi for all 256 bands:
projections P = <p|psi>, c = P dH, then T psi + v psi + sum p c in one
stencil launch.  Metric: band * gridpoints / s (higher is better).

Prints one JSON line.  Also reported: the FP64 path, the multigrid
preconditioner on the same grid, the stencil kernel's roofline (HBM bytes
per launch / measured launch time) and the CPU oracle on a bounded sample
of bands (all host cores, FP64).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=20)
    p.add_argument('--warmup', type=int, default=3)
    p.add_argument('--N', type=int, default=96)
    p.add_argument('--bands', type=int, default=256)
    p.add_argument('--atoms', type=int, default=32)
    p.add_argument('--no-cpu', action='store_true',
                   help='skip the CPU oracle baseline')
    p.add_argument('--cpu-seconds', type=float, default=12.0,
                   help='target length of the CPU baseline sample')
    p.add_argument('--no-ttt', action='store_true',
                   help='skip the time-to-1e-8 comparison of the multigrid '
                        'and learned preconditioners')
    p.add_argument('--ttt-max-iter', type=int, default=400)
    p.add_argument('--no-ragged', action='store_true',
                   help='skip the ragged-batch graph measurement')
    p.add_argument('--ragged-systems', type=int, default=48)
    p.add_argument('--no-kpoint', action='store_true',
                   help='skip the complex (k-point) H psi timing')
    p.add_argument('--no-poisson', action='store_true',
                   help='skip the multigrid Poisson solver timing')
    p.add_argument('--copy-calibration', action='store_true',
                   help='also run a plain copy of one band block (for '
                        'counter calibration under rocprofv3 --pmc)')
    return p.parse_args()


def time_events(fn, steps, stream_sync):
    """Run fn() `steps` times; return (wall seconds, per-call ms list of the
    events fn records)."""
    import torch
    stream_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    stream_sync()
    return time.perf_counter() - t0


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py: this build runs on one GPU; --gpus %d is not '
              'supported (bands shard across GPUs with no communication for '
              'H psi, see the design notes)' % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from gpawdft_amd.hamiltonian import Hamiltonian
    from gpawdft_amd.preconditioner import Preconditioner
    from gpawdft_amd.synthetic import System, smooth_bands_device

    assert torch.cuda.is_available(), 'bench.py needs a HIP device'
    dev = torch.device('cuda')
    N, nb = args.N, args.bands
    sysm = System(N, 0.18, args.atoms, seed=2024)
    gd = sysm.gd
    G = int(np.prod(gd.n_c))
    ham = Hamiltonian(gd, [sysm.vt_G], sysm.dH_asp)
    sync = torch.cuda.synchronize

    if args.copy_calibration:
        copy_calibration(4 * nb * G)

    results = {}
    for dt, name in ((torch.float32, 'f32'), (torch.float64, 'f64')):
        kin, pt = sysm.operators()
        psi = smooth_bands_device(nb, gd.n_c, 7, dtype=dt)
        out = torch.empty_like(psi)
        P = torch.zeros((nb, pt.Mmax), dtype=dt, device=dev)
        c = torch.empty_like(P)
        ham.apply_fused(psi, out, kin, pt, P_xM=P, c_xM=c)   # tables + warm
        for _ in range(args.warmup):
            ham.apply_fused(psi, out, kin, pt, P_xM=P, c_xM=c)
        # whole steps, wall clock between two synchronisations
        sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ham.apply_fused(psi, out, kin, pt, P_xM=P, c_xM=c)
        sync()
        wall = time.perf_counter() - t0
        # per-part timing with events on the launch stream (torch's current
        # stream, which stream_ptr() hands to every launch): projections,
        # the stencil kernel with the potential epilogue alone (the
        # roofline kernel), then the projector add
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        t_int = t_stencil = t_add = 0.0
        from gpawdft_amd._lib import check, lib
        from gpawdft_amd.device import stream_ptr, DeviceView
        vt = ham._vt(0, dt, dev)
        code = DeviceView(psi).code
        for _ in range(args.steps):
            ev[0].record()
            P.zero_()
            pt.integrate_xM(psi, P)
            ham.projector_coefs(pt, P, 0, c)
            ev[1].record()
            check(lib().gpd_hamiltonian_apply(
                kin.handle, None, code, nb, psi.data_ptr(), vt.data_ptr(),
                None, out.data_ptr(), stream_ptr()), 'hamiltonian_apply')
            ev[2].record()
            pt.add_xM(c, out)
            ev[3].record()
            ev[3].synchronize()
            t_int += ev[0].elapsed_time(ev[1])
            t_stencil += ev[1].elapsed_time(ev[2])
            t_add += ev[2].elapsed_time(ev[3])
        ms_step = wall / args.steps * 1e3
        ms_stencil = t_stencil / args.steps
        ms_proj = t_int / args.steps
        ms_add = t_add / args.steps
        esize = 4 if dt == torch.float32 else 8
        # algorithmic bytes of the stencil launch: read psi, write H psi,
        # read v once
        alg_bytes = 2 * esize * nb * G + esize * G
        results[name] = dict(
            ms_per_step=ms_step,
            value=nb * G / (ms_step * 1e-3),
            ms_stencil=ms_stencil,
            ms_projections=ms_proj,
            ms_projector_add=ms_add,
            stencil_gbs=alg_bytes / (ms_stencil * 1e-3) / 1e9,
            alg_bytes=alg_bytes)
        # multigrid preconditioner on the same grid: all bands in one call
        # (the throughput figure) and in blocks of 16 (eigensolver-sized)
        for blk, key in ((nb, 'precond'), (16, 'precond_blocks16')):
            pc = Preconditioner(gd, kin, block=blk)
            d = torch.empty_like(psi)
            for _ in range(2):
                pc(psi[:blk], None, out=d[:blk])
            sync()
            t0 = time.perf_counter()
            nrep = max(1, args.steps // 2)
            for _ in range(nrep):
                for b0 in range(0, nb, blk):
                    pc(psi[b0:b0 + blk], None, out=d[b0:b0 + blk])
            sync()
            tpc = (time.perf_counter() - t0) / nrep
            results[name][key + '_ms'] = tpc * 1e3
            results[name][key + '_value'] = nb * G / tpc
            del pc
        del psi, out, P, c, d
        torch.cuda.empty_cache()

    kpoint = None
    if not args.no_kpoint:
        kpoint = complex_hpsi(sysm, nb, args.steps)

    ragged = None
    if not args.no_ragged:
        ragged = ragged_batch(args.ragged_systems, args.steps)

    poisson = None
    if not args.no_poisson:
        poisson = poisson_solve(2 * N, args.no_cpu)

    ttt = None
    if not args.no_ttt:
        ttt = time_to_1e8(args.ttt_max_iter)

    # CPU baseline: the FP64 oracle (all host cores) on a bounded sample
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sysm, args.cpu_seconds)

    f32 = results['f32']
    traffic = measured_traffic()
    line = {
        'metric': 'band*gridpts/s for H*psi at 96^3x256',
        'value': f32['value'],
        'unit': 'band*gridpts/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': f32['ms_per_step'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic (seeded Gaussian-well potential, Gaussian '
                'projectors, smoothed N(0,1) bands)',
        'config': {'workload': 'H*psi, 96^3 periodic grid, h=0.18 A, %d '
                               'bands, %d atoms x 9 projectors, nn=3 stencil'
                               % (nb, args.atoms),
                   'grid': [N, N, N], 'bands': nb, 'atoms': args.atoms},
        'roofline': {'bound': 'hbm',
                     'kernel': 'fd_axis_kernel<float,3,EPI_POT,false> '
                               '(T psi + v psi, 256 bands)',
                     'achieved': f32['stencil_gbs'],
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': f32['stencil_gbs'] / HBM_PEAK_GBS,
                     'traffic': traffic,
                     'alg_bytes_per_launch': f32['alg_bytes'],
                     'ms_per_launch': f32['ms_stencil']},
        'cpu_baseline': cpu,
        'time_to_1e-8': ttt,
        'ragged_batch': ragged,
        'poisson': poisson,
        'kpoint_c128': kpoint,
        'f32_detail': {k: f32[k] for k in ('ms_stencil', 'ms_projections',
                                           'ms_projector_add', 'precond_ms',
                                           'precond_value',
                                           'precond_blocks16_ms')},
        'f64': {k: results['f64'][k] for k in
                ('value', 'ms_per_step', 'ms_stencil', 'ms_projections',
                 'ms_projector_add', 'stencil_gbs', 'precond_ms',
                 'precond_value', 'precond_blocks16_ms')},
    }
    for name in ('f32', 'f64'):
        r = results[name]
        r['hpsi_plus_precond_value'] = nb * G / (
            (r['ms_per_step'] + r['precond_ms']) * 1e-3)
    line['hpsi_plus_precond_value'] = f32['hpsi_plus_precond_value']
    if cpu:
        line['speedup_vs_cpu_f32'] = f32['value'] / cpu['value']
        line['speedup_vs_cpu_f64'] = results['f64']['value'] / cpu['value']
        line['hpsi_plus_precond_speedup_vs_cpu_f32'] = (
            f32['hpsi_plus_precond_value'] / cpu['hpsi_plus_precond_value'])
    print(json.dumps(line))


def complex_hpsi(sysm, nb, steps):
    """H psi for complex bands at a k-point (Bloch stencil, local
    potential, phased projector integrate / P dH / add) on the same 96^3
    system: the reference's operator sequence (gpaw/hamiltonian.py:
    426-460) in complex128."""
    import types
    import numpy as np
    import torch
    from gpawdft_amd.hamiltonian import FDWaveFunctions, Hamiltonian, KPoint
    from gpawdft_amd.lfc import LFC
    gd = sysm.gd
    k_c = np.array([0.25, 0.0, -0.375])
    phase = np.exp(2j * np.pi * np.array([[1, -1]] * 3) * k_c[:, None])
    pt = LFC(gd, [sysm.splines] * sysm.natoms,
             kd=types.SimpleNamespace(gamma=False, ibzk_qc=k_c[None]),
             dtype=complex)
    pt.set_positions(sysm.spos_ac)
    wfs = FDWaveFunctions(gd, pt, sysm.nn, complex)
    ham = Hamiltonian(gd, [sysm.vt_G], sysm.dH_asp)
    kpt = KPoint(q=0, phase_cd=phase)
    G = int(np.prod(gd.n_c))
    g = torch.Generator(device='cuda')
    g.manual_seed(3)
    psi = torch.randn((nb,) + tuple(int(n) for n in gd.n_c),
                      dtype=torch.complex128, device='cuda', generator=g)
    out = torch.empty_like(psi)
    ham.apply(psi, out, wfs, kpt)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ham.apply(psi, out, wfs, kpt)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    ev[0].record()
    wfs.kin.apply(psi, out, phase)
    ev[1].record()
    ev[1].synchronize()
    ms_kin = ev[0].elapsed_time(ev[1])
    res = {'workload': '96^3 periodic grid, %d complex128 bands, k = '
                       '(0.25, 0, -0.375), %d atoms x 9 projectors'
                       % (nb, sysm.natoms),
           'ms_per_step': ms, 'value': nb * G / (ms * 1e-3),
           'unit': 'band*gridpts/s', 'ms_stencil': ms_kin,
           'stencil_gbs': 2 * 16 * nb * G / (ms_kin * 1e-3) / 1e9}
    del psi, out
    torch.cuda.empty_cache()
    return res


def poisson_solve(N, no_cpu):
    """FP64 multigrid Poisson solve (gpaw/poisson.py, Jacobi smoothing,
    nn = 3) of a charged periodic density on the N^3 fine grid of the H psi
    system (twice its points per axis, as GPAW's finegd), tolerance 2e-10;
    the CPU oracle's V-cycle timed on one cycle for comparison."""
    import numpy as np
    import torch
    from gpawdft_amd.grid_descriptor import GridDescriptor
    from gpawdft_amd.poisson import PoissonSolver
    from gpawdft_amd.synthetic import Bohr
    n_c = np.array([N, N, N])
    gd = GridDescriptor(n_c, n_c * 0.09 / Bohr, True)
    x = np.arange(N) * 2 * np.pi / N
    rho = (np.cos(x)[:, None, None] * np.sin(2 * x)[None, :, None] *
           np.cos(3 * x)[None, None, :] + 0.1)
    ps = PoissonSolver()
    ps.set_grid_descriptor(gd)
    ps.initialize()
    drho = torch.from_numpy(rho).cuda()
    phi = torch.zeros_like(drho)
    ps.solve(phi, drho)                      # warm up (tables, scratch)
    phi.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    niter = ps.solve(phi, drho)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    out = {'workload': 'FP64 %d^3 periodic, nn=3, Jacobi, eps 2e-10, '
                       '%d levels' % (N, ps.levels + 1),
           'v_cycles': niter, 'seconds': t, 'ms_per_v_cycle': t / niter * 1e3}
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, 'oracle'))
        import poisson as P
        ref = P.Poisson(gd)
        phis = [np.zeros(tuple(n_c))] + [None] * ref.levels
        rhos = [rho - rho.mean()] + [None] * ref.levels
        t0 = time.perf_counter()
        ref.iterate2(phis, rhos)
        tc = time.perf_counter() - t0
        out['cpu_ms_per_v_cycle'] = tc * 1e3
        out['cpu_sample'] = 'one V-cycle of the C oracle (1 thread)'
        out['speedup_per_v_cycle'] = tc / (t / niter)
    return out


def ragged_batch(nsys, steps):
    """Throughput of one step (H psi + multigrid preconditioner, FP32) over
    nsys independent systems (grids from {32,48,64,80}^3, 16-256 bands, 4-48
    atoms): per-system eager launches versus one replay of a graph holding
    the same launches in parallel branches, after the launch-shape
    autotuner."""
    import torch
    from gpawdft_amd.ragged import RaggedBatch, random_specs
    t0 = time.perf_counter()
    rb = RaggedBatch(random_specs(nsys, seed=0), block=16)
    setup = time.perf_counter() - t0
    tuning = rb.autotune(reps=2)
    rb.capture()
    res = {}
    for name, fn in (('sequential', rb.run_sequential), ('graph', rb.replay)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        res[name] = {'ms_per_step': dt * 1e3,
                     'band_gridpts_per_s': rb.band_points / dt}
    res['graph_speedup'] = (res['sequential']['ms_per_step'] /
                            res['graph']['ms_per_step'])
    res['graph_branches'] = rb.nstreams
    res['systems'] = nsys
    res['band_gridpts_per_step'] = rb.band_points
    res['setup_s'] = setup
    res['wg_targets'] = {'%d pts x %d bands' % k: v for k, v in tuning.items()}
    del rb
    torch.cuda.empty_cache()
    return res


def time_to_1e8(max_iter):
    """Wall-clock seconds for FP64 RMM-DIIS to bring the occupied-band
    residual error below 1e-8 with the multigrid preconditioner and with the
    learned one (multigrid fallback), on a held-out 32^3 system of the same
    generators (h = 0.18 A, 3 wells, unit-norm projectors, 16 bands, 12
    occupied), both from the same Gaussian-orbital start.  Both runs must
    end on the same eigenvalues for the times to compare like with like:
    eps_max_diff is the largest difference of their occupied eigenvalues."""
    import numpy as np
    from gpawdft_amd.learned import UNet
    from gpawdft_amd.precond_training import time_to_tolerance
    from gpawdft_amd.synthetic import System
    net = UNet.load(os.path.join(ROOT, 'gpawdft_amd', 'data',
                                 'unet_precond.npz'))
    sysm = System(32, 0.18, 3, seed=4242, unit_projectors=True)
    out = {'workload': '32^3 periodic grid, h=0.18 A, 3 wells, unit-norm '
                       'projectors, 16 FP64 bands (12 occupied) from '
                       'Gaussian orbitals on the wells, blocksize 8, '
                       'tolerance 1e-8, at most %d iterations' % max_iter}
    for name, n in (('multigrid', None), ('learned', net)):
        out[name] = time_to_tolerance(sysm, n, nbands=16, nocc=12,
                                      max_iter=max_iter)
    e_m = np.array(out['multigrid']['eps_n'][:12])
    e_l = np.array(out['learned']['eps_n'][:12])
    out['eps_max_diff'] = float(np.abs(e_m - e_l).max())
    return out


def measured_traffic():
    """HBM bytes per launch of the roofline kernel from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, scaled by a plain
    copy's counters), or None when absent."""
    import re
    path = os.path.join(ROOT, 'profiles', 'fd_axis_f32_96x256_traffic.md')
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        m = re.search(r'^hbm_bytes_per_launch:\s*([0-9.eE+]+)', fh.read(), re.M)
    return float(m.group(1)) if m else None


def copy_calibration(nbytes, reps=3):
    """A plain device-to-device copy of nbytes: its FETCH_SIZE / WRITE_SIZE
    counters calibrate the stencil's (run under rocprofv3 --pmc)."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()


def cpu_baseline(sysm, target_s):
    """FP64 oracle H psi (stencil, potential, projectors) on all host
    cores, over as many bands as fit in ~target_s seconds."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle as O
    from gpawdft_amd.hamiltonian import unpack
    from gpawdft_amd.synthetic import smooth_bands_numpy
    gd = sysm.gd
    kin_c, kin_o = None, None
    from gpawdft_amd.fd_operators import guc_laplace_stencil
    coefs, offs, _ = guc_laplace_stencil(gd, -0.5, sysm.nn)
    from gpawdft_amd.lfc import LFC
    pt = LFC(gd, [sysm.splines] * sysm.natoms)
    pt.set_positions(sysm.spos_ac)
    atoms = [(pt.M_a[a], unpack(sysm.dH_asp[a][0])) for a in pt.atom_indices]
    threads = O.num_threads()

    def run(psi):
        t = O.fd(coefs, offs, gd.pbc_c, psi)
        t = O.add_local_potential(psi, sysm.vt_G, t)
        P = O.lfc_integrate(pt.A_Wgm, pt.M_W, pt.G_B, pt.W_B, gd.dv, psi,
                            pt.Mmax)
        c = O.projector_coefs(P, atoms)
        return O.lfc_add(pt.A_Wgm, pt.M_W, pt.G_B, pt.W_B, c, t)

    # the two-level preconditioner (gpaw/preconditioner.py) on the same grid
    from gpawdft_amd.transformers import transformer_paddings
    gd1 = gd.coarsen()
    gd2 = gd1.coarsen()
    c1, o1, _ = guc_laplace_stencil(gd1, -0.5, 1)
    c2, o2, _ = guc_laplace_stencil(gd2, -0.5, 1)

    def targs(a, b):
        pad, npad, skip, interp = transformer_paddings(a, b, 1)
        return (2, pad, skip, a.n_c, b.n_c)

    pargs = ((coefs, offs, gd.n_c), (c1, o1, gd1.n_c), (c2, o2, gd2.n_c),
             targs(gd, gd1), targs(gd1, gd2), targs(gd2, gd1), targs(gd1, gd),
             gd.pbc_c, 0.66666666 / coefs[0])

    nb = max(threads, 1)
    psi = smooth_bands_numpy(nb, gd.n_c, 11)
    t0 = time.perf_counter()
    run(psi)
    t1 = time.perf_counter() - t0
    reps = max(1, int(0.75 * target_s / max(t1, 1e-3)))
    t0 = time.perf_counter()
    for _ in range(reps):
        run(psi)
    t = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    O.precondition(*pargs, psi)
    tp1 = time.perf_counter() - t0
    preps = max(1, int(0.25 * target_s / max(tp1, 1e-3)))
    t0 = time.perf_counter()
    for _ in range(preps):
        O.precondition(*pargs, psi)
    tp = (time.perf_counter() - t0) / preps
    G = int(np.prod(gd.n_c))
    return {'value': nb * G / t, 'unit': 'band*gridpts/s', 'cores': threads,
            'kind': 'port',
            'precond_value': nb * G / tp,
            'hpsi_plus_precond_value': nb * G / (t + tp),
            'sample': '%d bands x %d repeats of FP64 H*psi (oracle: stencil '
                      '+ potential + projector integrate/dH/add) and %d '
                      'repeats of the multigrid preconditioner on the same '
                      '96^3 system' % (nb, reps, preps)}


if __name__ == '__main__':
    main()
