"""Ragged batches: one graph launch over systems of different sizes gives
the same bits as launching them one by one, on every replay; the stencil's
launch shape (autotuned) does not change results."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.array_equal(a.cpu().numpy().view(np.int32),
                          b.cpu().numpy().view(np.int32))


def test_launch_shape_does_not_change_results(cuda):
    import torch
    from gpawdft_amd.synthetic import System, smooth_bands_device
    sysm = System(48, 0.18, 4, seed=3)
    kin, pt = sysm.operators()
    psi = smooth_bands_device(24, sysm.gd.n_c, 1, dtype=torch.float32)
    ref = torch.empty_like(psi)
    kin.apply(psi, ref)
    for target in (512, 2048, 16384):
        kin.set_launch(target)
        out = torch.empty_like(psi)
        kin.apply(psi, out)
        assert _same(out, ref), target


def test_graph_replay_matches_sequential_launches(cuda):
    import torch
    from gpawdft_amd.ragged import RaggedBatch, random_specs
    specs = random_specs(6, seed=1, sizes=(32, 48), bands=(16, 40),
                         atoms=(4, 8))
    rb = RaggedBatch(specs, block=16)
    tuning = rb.autotune(reps=1)
    assert len(tuning) >= 1
    rb.run_sequential()
    torch.cuda.synchronize()
    ref = rb.outputs()
    rb.capture()
    for rep in range(3):
        for m in rb.members:
            m.out.fill_(float('nan'))
            m.d.fill_(float('nan'))
        torch.cuda.synchronize()
        rb.replay()
        torch.cuda.synchronize()
        for (o, d), m in zip(ref, rb.members):
            assert _same(m.out, o) and _same(m.d, d), rep
        # eager work between replays
        rb.members[0].step()
        torch.cuda.synchronize()
