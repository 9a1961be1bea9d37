"""Frame graphs (mista_set_graph): steady device frames recorded once as a
hipGraph and replayed.  Every replay must give the eager frame's images bit
for bit, with other GPU work and synchronisations between replays; new
output buffers or a changed setting must be picked up (re-capture), never a
stale graph."""
import sys

import numpy as np
import pytest

from us_lab4a_amd import imaging, workloads
from us_lab4a_amd.sta import CoherenceFactor, HipSTAProcessor

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def make_proc(wl):
    p = HipSTAProcessor(wl.sta_config(), imaging.ArrayAcquisition(np.zeros((1, 1, 1), np.float32)), wl.upsampling,
                        CoherenceFactor(wl.cf, wl.cf_param), wl.peak_offset, wl.rx_apodization(),
                        tx_apod=wl.tx_apodization())
    p.set_grid(wl.grid())
    return p


def test_graph_replays_match_eager_frames():
    import torch
    wl = workloads.sta2d_32_small()
    wl.cf = "phase_coherence_factor"
    n = wl.n_x * wl.n_z
    d_echoes = torch.empty((wl.num_tx, wl.num_elements, wl.num_samples), dtype=torch.float32, device="cuda")
    wl.echoes_gpu(d_echoes)
    torch.cuda.synchronize()

    def frame(p, v, f):
        p.process_device(d_echoes, wl.num_samples, v, f)
        torch.cuda.synchronize()
        return v.cpu().numpy().copy(), f.cpu().numpy().copy()

    eager = make_proc(wl)
    ev, ef = torch.empty(n, device="cuda"), torch.empty(n, device="cuda")
    ref_v, ref_f = frame(eager, ev, ef)

    g = make_proc(wl)
    g.set_graph(True)
    v, f = torch.empty(n, device="cuda"), torch.empty(n, device="cuda")
    for k in range(6):
        v.fill_(-1.0)
        f.fill_(-1.0)
        gv, gf = frame(g, v, f)
        say("graph frame", k, g.graph_stats())
        assert np.array_equal(gv, ref_v) and np.array_equal(gf, ref_f)
        frame(eager, ev, ef)   # other work on the device between replays
    caps, launches = g.graph_stats()
    assert caps == 1 and launches == 4   # frame 0 tunes, frame 1 is recorded, 2-5 replay
    assert g.last_timing() == (-1.0, -1.0)   # one launch per replayed frame: no stage split

    # new output buffers: recorded again, same images
    v2, f2 = torch.empty(n, device="cuda"), torch.empty(n, device="cuda")
    for k in range(3):
        gv, gf = frame(g, v2, f2)
        assert np.array_equal(gv, ref_v) and np.array_equal(gf, ref_f)
    caps2, launches2 = g.graph_stats()
    assert caps2 == 2 and launches2 == launches + 2

    # a changed setting (rx apodization) invalidates the graph
    apod = np.linspace(0.2, 1.0, wl.num_elements).astype(np.float32)
    g.set_rx_apod(apod)
    eager.set_rx_apod(apod)
    ref2_v, ref2_f = frame(eager, ev, ef)
    assert not np.array_equal(ref2_v, ref_v)
    for k in range(3):
        gv, gf = frame(g, v2, f2)
        assert np.array_equal(gv, ref2_v) and np.array_equal(gf, ref2_f)
    say("after the apodization change", g.graph_stats())
    assert g.graph_stats()[1] > launches2
    g.close()
    eager.close()
