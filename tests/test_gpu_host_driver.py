"""The C++ host driver (us_lab4a_amd/host/sta_cli.cpp, built on the header-only
HipSTAProcessor adapter) reproduces the oracle's image from a raw echo file,
as a reference caller of the processor interface would."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from us_lab4a_amd import imaging, workloads

from parity import check_reference_oracle, fp64_delays_2d

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "us_lab4a_amd", "mista_sta_cli")


@pytest.mark.parametrize("cf", ["none", "pcf"])
def test_cli_matches_oracle(tmp_path, cf):
    if not os.path.exists(CLI):
        pytest.fail("host driver not built (python -m us_lab4a_amd.build)")
    wl = workloads.sta2d_32_small()
    wl.tx_apod = None          # the driver applies rx apodization only, like the reference
    wl.cf = {"none": "none", "pcf": "phase_coherence_factor"}[cf]
    echoes = wl.echoes_cpu()
    path = tmp_path / "echoes.f32"
    echoes.astype(np.float32).tofile(path)
    out = tmp_path / "img"
    cmd = [CLI, "--echoes", str(path), "--elements", str(wl.num_elements), "--samples", str(wl.num_samples),
           "--up", str(wl.upsampling), "--pitch", repr(wl.pitch), "--fs", repr(wl.fs), "--c", repr(wl.c),
           "--fc", repr(wl.fc), "--grid", str(wl.n_x), str(wl.n_z), repr(wl.grid_x[0]), repr(wl.grid_x[1]),
           repr(wl.grid_z[0]), repr(wl.grid_z[1]), "--cf", cf, "--out", str(out)]
    print("[gpu-test] cli", cf, flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stderr
    v = np.fromfile(str(out) + "_value.f32", np.float32).reshape(wl.n_x, wl.n_z)
    f = np.fromfile(str(out) + "_factor.f32", np.float32).reshape(wl.n_x, wl.n_z)
    # the driver spaces grid points in double like Util::fillSequenceFromStartToEndWithSize
    g = imaging.rectangular_grid(wl.grid_x[0], wl.grid_x[1], wl.n_x, wl.grid_z[0], wl.grid_z[1], wl.n_z)
    v_o, f_o, _, _ = oracle.process(oracle.cfg_for(wl), echoes, imaging.grid_xz(g), wl.rx_apodization())
    peak = float(np.abs(v_o).max())
    assert float(np.abs(v - v_o).max()) / peak <= 1e-4
    if cf != "none":
        cfimg = np.fromfile(str(out) + "_cf.f32", np.float32).reshape(wl.n_x, wl.n_z)
        np.testing.assert_array_equal(cfimg, v * f)
        assert float(np.abs(v * f - v_o * f_o).max()) / peak <= 1e-4


H5DUMP = "/opt/conda/bin/h5dump"


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump absent")
def test_cli_hdf5_acquisition_and_images(tmp_path):
    """Saved acquisition in the reference's HDF5 layout in, the reference's
    image files out (image_value/factor/cf/x/z.h5): same images as the raw
    path, within 1e-4 of the oracle."""
    wl = workloads.sta2d_32_small()
    wl.tx_apod = None
    echoes = wl.echoes_cpu()
    raw = tmp_path / "e.f32"
    echoes.astype(np.float32).tofile(raw)
    acq = tmp_path / "acq"
    acq.mkdir()
    r = subprocess.run([CLI, "--echoes", str(raw), "--elements", str(wl.num_elements), "--samples",
                        str(wl.num_samples), "--write-acq", str(acq)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    common = ["--elements", str(wl.num_elements), "--up", str(wl.upsampling), "--pitch", repr(wl.pitch),
              "--fs", repr(wl.fs), "--c", repr(wl.c), "--fc", repr(wl.fc), "--grid", str(wl.n_x), str(wl.n_z),
              repr(wl.grid_x[0]), repr(wl.grid_x[1]), repr(wl.grid_z[0]), repr(wl.grid_z[1]), "--cf", "scf"]
    h5o = tmp_path / "img"
    h5o.mkdir()
    print("[gpu-test] cli hdf5", flush=True)
    r = subprocess.run([CLI, "--acq-dir", str(acq)] + common + ["--out", str(tmp_path / "a"), "--h5-out", str(h5o)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([CLI, "--echoes", str(raw), "--samples", str(wl.num_samples)] + common +
                       ["--out", str(tmp_path / "b")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr

    def h5(name, ds):
        b = tmp_path / f"{name}.bin"
        subprocess.run([H5DUMP, "-d", ds, "-b", "LE", "-o", str(b), str(h5o / f"{name}.h5")], capture_output=True,
                       timeout=60)
        return np.fromfile(b, np.float64).reshape(wl.n_x, wl.n_z)

    v_raw = np.fromfile(str(tmp_path / "b") + "_value.f32", np.float32).reshape(wl.n_x, wl.n_z)
    f_raw = np.fromfile(str(tmp_path / "b") + "_factor.f32", np.float32).reshape(wl.n_x, wl.n_z)
    np.testing.assert_array_equal(h5("image_value", "value"), v_raw.astype(np.float64))
    np.testing.assert_array_equal(h5("image_factor", "factor"), f_raw.astype(np.float64))
    np.testing.assert_array_equal(h5("image_cf", "cf"), (v_raw * f_raw).astype(np.float64))
    g = imaging.rectangular_grid(wl.grid_x[0], wl.grid_x[1], wl.n_x, wl.grid_z[0], wl.grid_z[1], wl.n_z)
    np.testing.assert_array_equal(h5("image_x", "x"), g["x"].astype(np.float64))
    np.testing.assert_array_equal(h5("image_z", "z"), g["z"].astype(np.float64))
    cfg = oracle.cfg_for(wl)
    cfg.cf_kind = oracle.CF_SCF
    v_o, _, _, _ = oracle.process(cfg, echoes, imaging.grid_xz(g), wl.rx_apodization())
    assert float(np.abs(v_raw - v_o).max()) / float(np.abs(v_o).max()) <= 1e-4


def test_cli_project_config_matches_oracle(tmp_path):
    """The driver configured by the reference's key=value project files
    (STA, imaging, coherence factor, rectangular grid) images the raw echo
    file like the oracle on the grid those files describe."""
    if not os.path.exists(CLI):
        pytest.fail("host driver not built (python -m us_lab4a_amd.build)")
    wl = workloads.sta2d_32_small()
    wl.tx_apod = None
    wl.cf = "phase_coherence_factor"
    echoes = wl.echoes_cpu()
    path = tmp_path / "echoes.f32"
    echoes.astype(np.float32).tofile(path)
    n = wl.num_elements
    (tmp_path / "sta.txt").write_text(
        f"num_elements = {n}\nnum_elements_mux = {n}\nfirst_tx_elem = 0\nlast_tx_elem = {n - 1}\nnum_pulses = 1\n"
        f"pitch = {wl.pitch!r}\ncenter_frequency = {wl.fc!r}\nmax_frequency = 10.0e6\nacquisition_time = 1.0e-4\n"
        f"min_gain = 0.0\nmax_gain = 0.0\npropagation_speed_1 = {wl.c!r}\nacquisition_delay = 0.0\n"
        f"sampling_frequency = {wl.fs!r}\ndead_zone_m = 0.0\nvalue_scale = 0.0\n")
    (tmp_path / "img.txt").write_text(f"peak_offset = 0.0\nupsampling_factor = {wl.upsampling}\nrx_apodization = hanning\n")
    (tmp_path / "cf.txt").write_text("analytic_signal_coherence_factor_method = phase_coherence_factor\n"
                                     "phase_coherence_factor_gamma = 1.0\n")
    (tmp_path / "grid.txt").write_text(
        "image_grid_type = rectangular\nrectangular_min_x = -5.0e-3\nrectangular_max_x = 5.0e-3\n"
        "rectangular_min_z = 10.0e-3\nrectangular_max_z = 30.0e-3\nrectangular_step_div = 0.5\n"
        "rectangular_origin_x = 0.0\nrectangular_origin_z = 0.0\n")
    cfg = ["--sta-config", str(tmp_path / "sta.txt"), "--imag-config", str(tmp_path / "img.txt"),
           "--cf-config", str(tmp_path / "cf.txt"), "--grid-config", str(tmp_path / "grid.txt")]
    dump = tmp_path / "dump"
    dump.mkdir()
    r = subprocess.run([CLI] + cfg + ["--dump-config", str(dump)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    n1, n2 = map(int, dict(l.split(" = ", 1) for l in open(dump / "config.txt").read().splitlines())["grid"].split())
    out = tmp_path / "img"
    print("[gpu-test] cli with project files", n1, n2, flush=True)
    r = subprocess.run([CLI] + cfg + ["--echoes", str(path), "--samples", str(wl.num_samples), "--out", str(out)],
                       capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stderr
    v = np.fromfile(str(out) + "_value.f32", np.float32).reshape(n1, n2)
    f = np.fromfile(str(out) + "_factor.f32", np.float32).reshape(n1, n2)
    xz = np.stack([np.fromfile(dump / "grid_x.f64").astype(np.float32),
                   np.fromfile(dump / "grid_z.f64").astype(np.float32)], axis=-1).reshape(n1, n2, 2)
    cfg_o = oracle.cfg_for(wl)
    # value within 1e-4 of the peak; CF image to the same bar except where the
    # oracle itself moves it when given the GPU's delays (FP64 distances
    # rounded once) and the CUDA kernels' arithmetic
    check_reference_oracle(v, f, cfg_o, oracle.prepare(cfg_o, echoes, np.float32), xz, fp64_delays_2d(wl, xz),
                           wl.rx_apodization(), None, "cli with project files")
