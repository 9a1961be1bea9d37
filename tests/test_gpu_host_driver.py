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
    v_o, f_o, _, _ = oracle.process(wl.oracle_cfg(), echoes, imaging.grid_xz(g), wl.rx_apodization())
    peak = float(np.abs(v_o).max())
    assert float(np.abs(v - v_o).max()) / peak <= 1e-4
    if cf != "none":
        cfimg = np.fromfile(str(out) + "_cf.f32", np.float32).reshape(wl.n_x, wl.n_z)
        np.testing.assert_array_equal(cfimg, v * f)
        assert float(np.abs(v * f - v_o * f_o).max()) / peak <= 1e-4
