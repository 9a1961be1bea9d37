"""Several frames of uneven shape in one launch of the tiled DAS kernel
(mista_das_batch_device): frames with the same storage, coherence factor,
transmit block and workgroup size share a launch, each taking its own tiles
from the launch's per-XCD work counters; the others run alone.  Every frame's
images must be bit-identical to the same frame's own launch (which is pinned
to the oracle by the parity tests)."""
import numpy as np
import pytest

from us_lab4a_amd import workloads
from us_lab4a_amd.sta import CoherenceFactor, HipSTAProcessor, das_batch_device

from parity import say

pytestmark = pytest.mark.gpu

MM = 1e-3

# elements, samples, (n_x, n_z), coherence factor, storage, tile shape (None: tuned)
FRAMES = [
    (32, 2048, (128, 128), "phase_coherence_factor", "f16", 1),
    (24, 1536, (96, 160), "phase_coherence_factor", "f16", 1),
    (32, 2048, (64, 256), "phase_coherence_factor", "f16", 1),
    (16, 1024, (64, 64), "sign_coherence_factor", "f16", 1),
    (32, 2048, (128, 96), "phase_coherence_factor", "f32", 1),
    (48, 2048, (96, 96), "none", "f16", None),
    (24, 1536, (80, 128), "phase_coherence_factor", "f16", 1),
]


def frame_wl(k):
    N, ns, (nx, nz), cf, _, _ = FRAMES[k]
    half = N * 0.3 * MM / 2
    return workloads.Workload2D(name=f"batch{k}", num_elements=N, num_samples=ns, upsampling=4, cf=cf,
                                grid_x=(-half, half), grid_z=(5 * MM, 30 * MM), n_x=nx, n_z=nz, points=8,
                                point_x=(-half, half), point_z=(6 * MM, 28 * MM), seed_points=60 + k, seed_noise=70 + k)


def test_batched_frames_equal_their_own_launches():
    import torch
    procs, outs, refs = [], [], []
    for k in range(len(FRAMES)):
        wl = frame_wl(k)
        storage, shape = FRAMES[k][4], FRAMES[k][5]
        p = HipSTAProcessor(wl.sta_config(), None, wl.upsampling, CoherenceFactor(wl.cf, wl.cf_param), 0.0,
                            wl.rx_apodization(), storage=storage)
        p.set_grid(wl.grid())
        if shape is not None:
            p.set_kernel(2, shape, 0)
        d = torch.empty((wl.num_tx, wl.num_elements, wl.num_samples), dtype=torch.float32, device="cuda")
        wl.echoes_gpu(d)
        n = wl.n_x * wl.n_z
        v, f = torch.empty(n, device="cuda"), torch.empty(n, device="cuda")
        p.process_device(d, wl.num_samples, v, f)
        torch.cuda.synchronize()
        assert p.last_kernel()[0] >= 0, (k, "the tiled DAS ran")
        refs.append((v.cpu().numpy(), f.cpu().numpy()))
        procs.append(p)
        outs.append((torch.full((n,), -1.0, device="cuda"), torch.full((n,), -1.0, device="cuda")))
    say("batch of", len(procs), "frames")
    for rep in range(2):
        das_batch_device(procs, [o[0] for o in outs], [o[1] for o in outs])
        torch.cuda.synchronize()
        for k, (o, r) in enumerate(zip(outs, refs)):
            assert np.array_equal(o[0].cpu().numpy(), r[0]), (rep, k)
            assert np.array_equal(o[1].cpu().numpy(), r[1]), (rep, k)
        for o in outs:
            o[0].fill_(-1.0)
            o[1].fill_(-1.0)
    # an untuned handle in the batch runs alone (and tunes)
    wl = frame_wl(5)
    q = HipSTAProcessor(wl.sta_config(), None, wl.upsampling, CoherenceFactor(wl.cf, wl.cf_param), 0.0,
                        wl.rx_apodization(), storage="f16")
    q.set_grid(wl.grid())
    d = torch.empty((wl.num_tx, wl.num_elements, wl.num_samples), dtype=torch.float32, device="cuda")
    wl.echoes_gpu(d)
    q.prepare_signals_device(d, wl.num_samples)
    torch.cuda.synchronize()   # q prepared on its own stream; the batch runs on procs[0]'s
    n = wl.n_x * wl.n_z
    vq, fq = torch.empty(n, device="cuda"), torch.empty(n, device="cuda")
    das_batch_device([procs[0], q], [outs[0][0], vq], [outs[0][1], fq])
    torch.cuda.synchronize()
    assert np.array_equal(vq.cpu().numpy(), refs[5][0])
    assert np.array_equal(outs[0][0].cpu().numpy(), refs[0][0])
    for p in procs + [q]:
        p.close()
