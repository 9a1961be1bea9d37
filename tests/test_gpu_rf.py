"""GPU parity of the RF path (DefaultSTAProcessor, DefaultSTAProcessor.h:
136-226, with the envelope STAMethod takes afterwards, STAMethod.cpp:63-72)
through the C ABI against the oracle: preparation (Interpolator4X + DC
removal), real-sample DAS with and without SCF, the row envelope and the
coherence-factor image envelope(value) * factor."""
import sys

import numpy as np
import pytest

import oracle
from us_lab4a_amd import imaging, workloads
from us_lab4a_amd.sta import CoherenceFactor, HipSTARFProcessor

pytestmark = pytest.mark.gpu

TOL = 1e-4


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def frame():
    wl = workloads.sta2d_32_small()
    wl.upsampling = 4
    return wl, wl.echoes_cpu()


def make_proc(wl, cf):
    return HipSTARFProcessor(wl.sta_config(), None, CoherenceFactor(cf, 1.0), wl.peak_offset)


def oracle_cfg(wl, cf):
    cfg = oracle.cfg_for(wl)
    cfg.upsampling_factor = 4
    cfg.cf_kind = oracle.CF_SCF if cf == "sign_coherence_factor" else oracle.CF_NONE
    return cfg


def test_rf_prepare_matches_oracle(frame):
    wl, echoes = frame
    say("rf prepare")
    p = make_proc(wl, "none")
    p.set_grid(wl.grid())
    p.process_echoes(echoes)
    t = p.download_analytic()
    p.close()
    t64 = oracle.prepare_rf(oracle_cfg(wl, "none"), echoes, np.float64)
    err = float(np.abs(t - t64).max() / np.abs(t64).max())
    say("rf prepare error / peak", err)
    assert err <= 2e-6


@pytest.mark.parametrize("cf", ["none", "sign_coherence_factor"])
def test_rf_frame_matches_oracle(frame, cf):
    wl, echoes = frame
    say("rf frame", cf)
    p = make_proc(wl, cf)
    p.set_grid(wl.grid())
    v, f = p.process_echoes(echoes)
    cfg = oracle_cfg(wl, cf)
    t32 = oracle.prepare_rf(cfg, echoes, np.float32)
    xz = imaging.grid_xz(wl.grid())
    v_o, f_o, sums = oracle.das_rf_rxsums(cfg, t32, xz)
    peak = float(np.abs(v_o).max())
    err = float(np.abs(v.astype(np.float64) - v_o).max()) / peak
    say("rf value error / peak", err)
    assert err <= TOL
    # envelope along z and the CF image, as STAMethod::useCoherenceFactor
    env = p.envelope(v)
    env_o = oracle.envelope_dim2(v_o.astype(np.float64))
    epk = float(np.abs(env_o).max())
    eerr = float(np.abs(env - env_o).max()) / epk
    say("rf envelope error / peak", eerr)
    assert eerr <= TOL
    if cf == "none":
        assert np.all(f == 1.0)
    else:
        # per-rx sums over the GPU's own delay table and tensor, in the CUDA
        # kernels' arithmetic (real samples: imaginary parts zero, no
        # apodization): the SCF image envelope * factor at every pixel
        t = p.download_analytic()
        d = p.download_delays()
        tc = np.zeros(t.shape + (2,), np.float32)
        tc[..., 0] = t
        _, f_t, sums = oracle.das_table(cfg, tc, d, np.ones(wl.num_elements, np.float32), None,
                                        form=oracle.FORM_CUDA, rxsums=True)
        v_t = sums.real.astype(np.float64).sum(axis=1) / (wl.num_tx * wl.num_elements)
        assert float(np.abs(v.reshape(-1) - v_t).max()) <= TOL * peak
        bad = np.abs(env.astype(np.float64) * f - env_o * f_t.reshape(env.shape)) > TOL * epk
        say("rf SCF image over the bar (table oracle)", int(bad.sum()), "of", bad.size)
        assert not bad.any()
    p.close()


def test_rf_tiles_match_gather(frame):
    wl, echoes = frame
    say("rf tiles vs gather")
    p = make_proc(wl, "sign_coherence_factor")
    p.set_grid(wl.grid())
    p.set_kernel(2)
    v_t, f_t = p.process_echoes(echoes)
    p.set_kernel(1)
    v_g, f_g = p.process_echoes(echoes)
    p.close()
    np.testing.assert_array_equal(v_t, v_g)
    np.testing.assert_array_equal(f_t, f_g)


def test_envelope_rows_match_oracle():
    rng = np.random.default_rng(9)
    data = rng.standard_normal((37, 700)).astype(np.float32)
    wl = workloads.sta2d_32_small()
    p = make_proc(wl, "none")
    e = p.envelope(data)
    p.close()
    e_o = oracle.envelope_dim2(data.astype(np.float64))
    assert float(np.abs(e - e_o).max() / np.abs(e_o).max()) <= 2e-6


def test_rf_rejects_pcf_and_other_upsampling(frame):
    from us_lab4a_amd import _lib
    wl, _ = frame
    with pytest.raises(_lib.InvalidParameterError):
        HipSTARFProcessor(wl.sta_config(), None, CoherenceFactor("phase_coherence_factor", 1.0), 0.0)
