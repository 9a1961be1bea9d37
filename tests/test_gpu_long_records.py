"""Long records on the hand-written four-step FFT (fft4.hip), through the C ABI.

The reference's analytic signal takes records of up to 2^20 samples
(HilbertEnvelope.h:53, :142-251); the fused on-chip preparation stops at FFT
sizes of 36 864, so longer records (8192 samples at U = 8: FFT size 67 500)
run the four-step transform: FIR x U, a two-pass forward FFT, the spectrum
rule, a two-pass backward FFT, the storage conversion.

- automatic preparation equals the explicitly selected four-step path bit
  for bit where the fused kernel does not fit (so no library FFT runs);
- FP32 storage within 2e-6 of the FP64 oracle's peak, at several record
  lengths and upsampling factors whose FFT sizes split differently;
- images over a long record against the table oracle (FP32 and FP16) and
  the reference-equivalent oracle (FP32)."""
import numpy as np
import pytest

import oracle
from us_lab4a_amd import workloads
from us_lab4a_amd.imaging import grid_xz
from us_lab4a_amd.sta import CoherenceFactor, HipSTAProcessor

from parity import check_reference_oracle, check_table_oracle, say

pytestmark = pytest.mark.gpu

MM = 1e-3


def long_wl(num_samples, upsampling, num_elements=16, ntx=4, cf="phase_coherence_factor", n_x=48, n_z=768):
    zmax = min(120 * MM, num_samples / 40e6 * 1540 / 2 * 0.9)
    half = num_elements * 0.3 * MM / 2
    wl = workloads.Workload2D(name=f"long_{num_samples}x{upsampling}", num_elements=num_elements,
                              num_samples=num_samples, upsampling=upsampling, cf=cf, tx_apod="hanning",
                              grid_x=(-half, half), grid_z=(5 * MM, zmax), n_x=n_x, n_z=n_z, points=12,
                              point_x=(-half, half), point_z=(6 * MM, zmax), seed_points=31, seed_noise=32)
    wl.last_tx = ntx - 1
    return wl


def make_proc(wl, storage, method=0):
    p = HipSTAProcessor(wl.sta_config(), None, wl.upsampling, CoherenceFactor(wl.cf, wl.cf_param), wl.peak_offset,
                        wl.rx_apodization(), tx_apod=wl.tx_apodization(), storage=storage)
    p.set_grid(wl.grid())
    p.set_prep(method)
    return p


# (samples, U): FFT sizes 67 500 = 2^2 3^3 5^4, 34 560 = 2^8 3^3 5,
# 17 496 = 2^3 3^7 and 4 050 = 2 3^4 5^2 (the last three fit the fused kernel
# and take the four-step path only when it is selected).
SIZES = [(8192, 8), (4096, 8), (2048, 8), (1000, 3)]


@pytest.mark.parametrize("ns,up", SIZES, ids=[f"{n}x{u}" for n, u in SIZES])
def test_four_step_prepare_matches_fp64_oracle(ns, up):
    wl = long_wl(ns, up)
    echoes = wl.echoes_cpu()
    a64 = oracle.prepare(oracle.cfg_for(wl), echoes, np.float64)
    say("four-step prepare", ns, "samples x", up)
    p = make_proc(wl, "f32", method=3)
    p.process_echoes(echoes)
    L, nf, _ = p.query()
    a3 = p.download_analytic()
    p.close()
    err = float(np.abs(a3 - a64).max() / np.abs(a64).max())
    say("FFT size", nf, "four-step error / peak", f"{err:.2e}")
    assert a3.shape == a64.shape
    assert err <= 2e-6
    if nf > 36864:
        # the automatic choice for a record this long is the four-step path
        p = make_proc(wl, "f32", method=0)
        p.process_echoes(echoes)
        a0 = p.download_analytic()
        p.close()
        assert np.array_equal(a0, a3)
        with pytest.raises(Exception, match="fused"):
            q = make_proc(wl, "f32", method=2)
            try:
                q.process_echoes(echoes)
            finally:
                q.close()


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_long_record_image_matches_oracles(storage):
    wl = long_wl(8192, 8, num_elements=24, ntx=24)
    echoes = wl.echoes_cpu()
    cfg = oracle.cfg_for(wl)
    say("long-record image", storage)
    p = make_proc(wl, storage)
    v, f = p.process_echoes(echoes)
    assert p.last_kernel()[0] >= 0, "the tiled DAS ran"
    a = p.download_analytic()
    d = p.download_delays()
    p.close()
    rx, tx = wl.rx_apodization(), wl.tx_apodization()
    check_table_oracle(v, f, cfg, a, d, rx, tx, f"long {storage}", branch_rel=(1e-6 if storage != "f32" else 0.0))
    if storage == "f32":
        check_reference_oracle(v, f, cfg, a, grid_xz(wl.grid()), d, rx, tx, "long f32")
