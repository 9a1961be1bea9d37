"""GPU parity on the reference's own edge cases, through the C ABI, on the
tiled and the gather DAS kernel, with FP32 and FP16 storage:

- a 10 mm dead zone, as the reference project's configuration sets it
  (project/sta/exp-0.5mhz_64elem/config-sta.txt:45): removeDC averages over
  [dead zone, L) only (VectorialSTAProcessor.h:138, :183-187, Util.h:1009-1018);
- a single transmit element, first = last = 15 of 32
  (project/sta/exp-0.5mhz_64elem/config-sta-cyl.txt:13-14), and a middle
  sub-range 8..23: tx t reads delay row first + t and tensor row t
  (VectorialSTAProcessor.h:251-284);
- a peak offset of one period, (fs U) peakOffset / fc samples added to every
  position (VectorialSTAProcessor.h:147);
- U = 8 with a dead zone, which runs the FIR's 4-output-per-store path.

Preparation: within 2e-6 of the FP64 oracle's peak.  Images: see
tests/parity.py (value within 1e-4 of the peak against the
reference-equivalent oracle; value and CF image within 1e-4 at every pixel
against the oracle fed the GPU's delay table).  Tiled equals gather bit for
bit with a coherence factor."""
import numpy as np
import pytest

import oracle
from us_lab4a_amd import imaging, workloads
from us_lab4a_amd.sta import CoherenceFactor, HipSTAProcessor

from parity import TOL, check_reference_oracle, check_table_oracle, say

pytestmark = pytest.mark.gpu

MM = 1e-3


def _case(name):
    if name == "dead_zone_10mm_pcf":
        wl = workloads.sta2d_32_small()
        wl.dead_zone_m = 10 * MM
        wl.cf = "phase_coherence_factor"
    elif name == "single_tx15_pcf":
        wl = workloads.sta2d_32_small()
        wl.first_tx, wl.last_tx = 15, 15
        wl.tx_apod = None
        wl.cf = "phase_coherence_factor"
    elif name == "tx8_23_scf":
        wl = workloads.sta2d_32_small()
        wl.first_tx, wl.last_tx = 8, 23
        wl.cf = "sign_coherence_factor"
    elif name == "peak_offset1_pcf":
        wl = workloads.sta2d_32_small()
        wl.peak_offset = 1.0
        wl.cf = "phase_coherence_factor"
    elif name == "u8_dead_zone_3mm_none":
        wl = workloads.tiny(num_elements=16, num_samples=1024, upsampling=8, n_x=32, n_z=48)
        wl.grid_z = (2 * MM, 12 * MM)
        wl.point_z = (3 * MM, 11 * MM)
        wl.dead_zone_m = 3 * MM
        wl.tx_apod = "hanning"
    else:
        raise KeyError(name)
    wl.name = name
    return wl


CASES = ["dead_zone_10mm_pcf", "single_tx15_pcf", "tx8_23_scf", "peak_offset1_pcf", "u8_dead_zone_3mm_none"]
_cache = {}


def case_data(name):
    """Workload, echoes, the FP64 oracle tensor and the oracle's FP32 tensor."""
    if name not in _cache:
        wl = _case(name)
        echoes = wl.echoes_cpu()
        cfg = oracle.cfg_for(wl)
        _cache[name] = (wl, echoes, oracle.prepare(cfg, echoes, np.float64), oracle.prepare(cfg, echoes, np.float32))
    return _cache[name]


def make_proc(wl, storage):
    p = HipSTAProcessor(wl.sta_config(), None, wl.upsampling, CoherenceFactor(wl.cf, wl.cf_param), wl.peak_offset,
                        wl.rx_apodization(), tx_apod=wl.tx_apodization(), storage=storage)
    p.set_grid(wl.grid())
    return p


def run_das(p, wl, variant):
    import torch
    n = wl.n_x * wl.n_z
    p.set_kernel(variant, 0, 0)
    dv = torch.empty(n, dtype=torch.float32, device="cuda")
    df = torch.empty(n, dtype=torch.float32, device="cuda")
    p.das_device(dv, df)
    torch.cuda.synchronize()
    assert (p.last_kernel()[0] >= 0) == (variant == 2), "the requested DAS kernel did not run"
    return dv.cpu().numpy().reshape(wl.n_x, wl.n_z), df.cpu().numpy().reshape(wl.n_x, wl.n_z)


@pytest.mark.parametrize("method", [1, 2, 3], ids=["library_fft", "fused_lds", "four_step"])
@pytest.mark.parametrize("name", CASES)
def test_edge_prepare_matches_fp64_oracle(name, method):
    wl, echoes, a64, _ = case_data(name)
    say("edge prepare", name, "method", method)
    p = make_proc(wl, "f32")
    p.set_prep(method)
    p.process_echoes(echoes)
    a = p.download_analytic()
    p.close()
    assert a.shape == a64.shape
    err = float(np.abs(a - a64).max() / np.abs(a64).max())
    say(name, "prepare error / peak", f"{err:.2e}")
    assert err <= 2e-6


@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("name", CASES)
def test_edge_das_matches_oracles(name, storage):
    wl, echoes, a64, a32 = case_data(name)
    say("edge das", name, storage)
    cfg = oracle.cfg_for(wl)
    xz = imaging.grid_xz(wl.grid())
    rx, tx = wl.rx_apodization(), wl.tx_apodization()
    p = make_proc(wl, storage)
    p.process_echoes(echoes)
    a_gpu = p.download_analytic()
    d_gpu = p.download_delays()
    out = {variant: run_das(p, wl, variant) for variant in (2, 1)}
    p.close()
    for variant, (v, f) in out.items():
        tag = f"{name} {storage} {'tiled' if variant == 2 else 'gather'}"
        check_table_oracle(v, f, cfg, a_gpu, d_gpu, rx, tx, tag, branch_rel=1e-6 if storage != "f32" else 0.0)
        if storage == "f32":
            # the reference-equivalent oracle on its own tensor and distances
            check_reference_oracle(v, f, cfg, a32, xz, d_gpu, rx, tx, tag)
        else:
            v64, _ = oracle.das(cfg, a64, xz, rx, tx)
            err = workloads.db_error(v, v64)
            say(tag, "envelope error vs FP64 (dB):", round(err, 1))
            assert err <= -60.0
    if wl.cf != "none":
        assert np.array_equal(out[1][0], out[2][0]) and np.array_equal(out[1][1], out[2][1])
    else:
        assert float(np.abs(out[1][0] - out[2][0]).max()) <= TOL * float(np.abs(out[1][0]).max())


def test_single_transmit_reads_its_own_rows():
    """first = last = 15: the image must come from transmit element 15 alone.
    Zeroing every other transmit's echoes of a full 32-transmit acquisition
    and imaging it with all 32 transmits gives 32 x the single-transmit value
    (the normalisation is 1 / (Ntx Nrx)) wherever the full image is formed."""
    wl = _case("single_tx15_pcf")
    full = workloads.sta2d_32_small()
    full.tx_apod = None
    full.cf = "none"
    e_full = full.echoes_cpu()
    e_one = e_full[15:16].copy()
    say("single transmit rows")
    wl.cf = "none"
    p1 = make_proc(wl, "f32")
    v1, _ = p1.process_echoes(e_one)
    p1.close()
    e_masked = np.zeros_like(e_full)
    e_masked[15] = e_full[15]
    p2 = make_proc(full, "f32")
    v2, _ = p2.process_echoes(e_masked)
    p2.close()
    assert np.abs(v1 - 32.0 * v2).max() <= 1e-4 * np.abs(v1).max()
