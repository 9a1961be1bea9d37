"""CPU tests of the host side: the C ABI library loads and exports every
declared entry point (no compute calls), configuration validation at the
boundary, and the host helpers against the oracle / closed forms."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from us_lab4a_amd import _lib, imaging, synth, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "mista.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mista_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(_lib.LIB_PATH):
        from us_lab4a_amd import build
        build.build_library(verbose=False)
    return _lib.lib()


def test_library_exports_every_declared_symbol(L):
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    # the Python binding declares a signature for every header function
    assert sorted(_lib.SIGNATURES) == names


def test_abi_version(L):
    assert L.mista_abi_version() == 1


def _cfg(**kw):
    c = _lib.StaConfig()
    c.num_elements, c.first_tx_elem, c.last_tx_elem, c.upsampling_factor = 32, 0, 31, 4
    c.pitch, c.sampling_frequency, c.propagation_speed, c.center_frequency = 0.3e-3, 40e6, 1540.0, 5e6
    c.dead_zone_m, c.peak_offset, c.cf_kind, c.cf_param, c.storage, c.device = 0.0, 0.0, 0, 1.0, 0, 0
    for k, v in kw.items():
        setattr(c, k, v)
    return c


@pytest.mark.parametrize("kw,code", [
    (dict(num_elements=0), _lib.MISTA_E_INVALID_PARAM),
    (dict(last_tx_elem=32), _lib.MISTA_E_INVALID_PARAM),
    (dict(first_tx_elem=5, last_tx_elem=4), _lib.MISTA_E_INVALID_PARAM),
    (dict(upsampling_factor=65), _lib.MISTA_E_INVALID_VALUE),
    (dict(cf_kind=7), _lib.MISTA_E_INVALID_PARAM),
    (dict(storage=9), _lib.MISTA_E_INVALID_PARAM),
    (dict(sampling_frequency=0.0), _lib.MISTA_E_INVALID_PARAM),
    (dict(peak_offset=-1.0), _lib.MISTA_E_INVALID_PARAM),
])
def test_create_rejects_invalid_configuration(L, kw, code):
    h = ctypes.c_void_p()
    rc = L.mista_create(ctypes.byref(_cfg(**kw)), ctypes.byref(h))
    assert rc == code
    assert h.value is None
    assert L.mista_last_error(None)


def _cfg3d(**kw):
    c = _lib.Sta3dConfig()
    base = dict(num_tx=4, num_rx=16, upsampling_factor=4, sampling_frequency=40e6, propagation_speed=1540.0,
                center_frequency=3e6, dead_zone_m=0.0, peak_offset=0.0, cf_kind=_lib.CF_SCF, cf_param=1.0,
                storage=_lib.STORAGE_F32, device=0)
    base.update(kw)
    for k, v in base.items():
        setattr(c, k, v)
    return c


@pytest.mark.parametrize("kw,code", [
    (dict(num_tx=0), _lib.MISTA_E_INVALID_VALUE),           # SA3DConfiguration: empty active list
    (dict(num_rx=0), _lib.MISTA_E_INVALID_VALUE),
    (dict(upsampling_factor=0), _lib.MISTA_E_INVALID_VALUE),
    (dict(sampling_frequency=0.0), _lib.MISTA_E_INVALID_PARAM),
    (dict(cf_kind=7), _lib.MISTA_E_INVALID_PARAM),
    (dict(storage=9), _lib.MISTA_E_INVALID_PARAM),
])
def test_create_3d_rejects_invalid_configuration(L, kw, code):
    h = ctypes.c_void_p()
    xy = np.zeros((64, 2), np.float32)
    rc = L.mista_create_3d(ctypes.byref(_cfg3d(**kw)), xy.ctypes.data, xy.ctypes.data, ctypes.byref(h))
    assert rc == code
    assert h.value is None
    assert L.mista_last_error(None)


def test_engine_rejects_invalid_configuration(L):
    h = ctypes.c_void_p()
    for kw in (dict(max_inflight=0, max_queued=4), dict(max_inflight=65, max_queued=4), dict(max_inflight=2, max_queued=0)):
        c = _lib.EngineConfig()
        c.device = 0
        for k, v in kw.items():
            setattr(c, k, v)
        assert L.mista_engine_create(ctypes.byref(c), ctypes.byref(h)) == _lib.MISTA_E_INVALID_PARAM
        assert h.value is None
    assert L.mista_engine_create(None, ctypes.byref(h)) == _lib.MISTA_E_INVALID_PARAM
    assert L.mista_engine_destroy(None) == _lib.MISTA_OK
    assert L.mista_engine_submit(None, None, None) == _lib.MISTA_E_INVALID_PARAM
    assert L.mista_engine_wait(None, 1, None) == _lib.MISTA_E_INVALID_PARAM


def test_null_handle_calls_fail_cleanly(L):
    assert L.mista_destroy(None) == 0
    assert L.mista_set_grid(None, None, 1, 1) == _lib.MISTA_E_INVALID_PARAM
    assert L.mista_process(None, None, 1, None, None) == _lib.MISTA_E_INVALID_PARAM
    assert L.mista_set_grid_xyz(None, None, 1, 1, 1) == _lib.MISTA_E_INVALID_PARAM
    assert L.mista_set_elements_3d(None, None, None) == _lib.MISTA_E_INVALID_PARAM
    assert L.mista_create_3d(None, None, None, None) == _lib.MISTA_E_INVALID_PARAM


def test_windows_match_oracle():
    for name in ("rectangular", "hanning", "hanning2", "hamming", "blackman", "blackman2"):
        np.testing.assert_allclose(imaging.window(name, 33), oracle.window(name, 33).astype(np.float32), atol=0)
    with pytest.raises(ValueError):
        imaging.window("nope", 4)


def test_element_positions():
    x = imaging.element_x_centered_2d(4, 0.3e-3)
    np.testing.assert_allclose(x, [-0.45e-3, -0.15e-3, 0.15e-3, 0.45e-3], atol=1e-9)
    p = imaging.element_positions_2d_array(0.3e-3, 2, 0.0, 0.3e-3, 2, 1e-3)
    np.testing.assert_allclose(p, [[-0.15e-3, 0.85e-3], [0.15e-3, 0.85e-3], [-0.15e-3, 1.15e-3], [0.15e-3, 1.15e-3]],
                               atol=1e-9)


def test_rectangular_grid_layout():
    g = imaging.rectangular_grid(-1e-3, 1e-3, 3, 5e-3, 7e-3, 5)
    assert g.shape == (3, 5)
    assert np.all(g["x"][:, 0] == np.float32([-1e-3, 0.0, 1e-3]))
    np.testing.assert_allclose(g["z"][0], np.linspace(5e-3, 7e-3, 5).astype(np.float32))
    assert np.all(g["factor"] == 1.0)
    xz = imaging.grid_xz(g)
    assert xz.shape == (3, 5, 2) and xz.flags.c_contiguous
    g2 = imaging.rectangular_grid_lambda(-1e-3, 1e-3, 5e-3, 6e-3, 0.3e-3, 3.0)
    assert np.diff(g2["x"][:, 0]).max() <= 0.1e-3 + 1e-9


def test_counter_rng_is_deterministic_and_uniform():
    a = synth.uniform01(7, np.arange(100000))
    b = synth.uniform01(7, np.arange(100000))
    assert np.array_equal(a, b)
    assert 0.0 < a.min() and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.005
    g = synth.gaussian(3, 200000)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1.0) < 0.01
    assert not np.array_equal(synth.uniform01(8, np.arange(10)), a[:10])


def test_echo_model_peaks_at_round_trip_time():
    tx = np.zeros((1, 3), np.float32)
    rx = np.array([[0, 0, 0], [1e-3, 0, 0]], np.float32)
    scat = np.array([[0.0, 0.0, 10e-3, 1.0]], np.float32)
    e = synth.echoes_cpu(tx, rx, scat, 1024, 40e6, 1540.0, 5e6, 0.6)
    for r in range(2):
        d = 10e-3 + np.linalg.norm(scat[0, :3] - rx[r])
        assert abs(int(np.argmax(e[0, r])) - d / 1540.0 * 40e6) <= 1.0
    assert abs(e.max() - 1.0) < 0.02


def test_workload_sizes():
    w1 = workloads.sta2d_32_small()
    assert w1.pixel_pairs == 256 * 256 * 32 * 32
    w2 = workloads.sta2d_128_pcf()
    assert w2.pixel_pairs == 512 * 1024 * 128 * 128
    assert w2.num_samples * w2.upsampling == 32768
    assert oracle.next_fast_even(32768 + 1024) == 34560
    sp = synth.speckle_scatterers(3, 5000.0, w2.grid_x, w2.grid_z)
    assert len(sp) == 114000


H5DUMP = "/opt/conda/bin/h5dump"
CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "us_lab4a_amd", "mista_sta_cli")


@pytest.mark.skipif(not (os.path.exists(H5DUMP) and os.path.exists(CLI)), reason="h5dump or the host driver absent")
def test_cli_writes_reference_acquisition_layout(tmp_path):
    """--write-acq stores one HDF5 file per transmit in the reference's
    saved-acquisition layout (FileUtil.cpp txElemSignalsPath, dataset
    "signal" [rx][samples], IEEE_F64LE as HDF5Util::save writes)."""
    import subprocess
    rng = np.random.default_rng(2)
    echoes = rng.standard_normal((2, 3, 5)).astype(np.float32)
    raw = tmp_path / "e.f32"
    echoes.tofile(raw)
    out = tmp_path / "acq"
    out.mkdir()
    r = subprocess.run([CLI, "--echoes", str(raw), "--elements", "3", "--samples", "5", "--write-acq", str(out),
                        "--base", "7"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    for t in range(2):
        f = out / f"signals-base0007-tx{t:04d}.h5"
        assert f.exists()
        head = subprocess.run([H5DUMP, "-H", str(f)], capture_output=True, text=True, timeout=60).stdout
        assert 'DATASET "signal"' in head and "H5T_IEEE_F64LE" in head and "( 3, 5 )" in head
        b = tmp_path / f"t{t}.bin"
        subprocess.run([H5DUMP, "-d", "signal", "-b", "LE", "-o", str(b), str(f)], capture_output=True, timeout=60)
        np.testing.assert_array_equal(np.fromfile(b, np.float64).reshape(3, 5), echoes[t].astype(np.float64))
