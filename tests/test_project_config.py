"""The reference's key=value project files through the CLI (no GPU use:
--dump-config resolves the configuration and exits).  Pinned by closed
forms: KeyValueFileReader.cpp:30-73 (format and errors), ParameterMap.h:
84-118 (ranges), STAConfiguration.h:59-74 (keys), ImageGrid.h:120-186 with
Util.h:266-292 (fewest evenly spaced points whose step does not exceed
lambda / step_div, lambda = c / (2 max_frequency)), WindowFunction.h:
237-289 (numpy's hanning / hamming / blackman / bartlett are the same
formulas)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "us_lab4a_amd", "mista_sta_cli")
PROJ = os.path.join(ROOT, "tests", "golden", "project")


def run(tmp_path, grid="config-grid-rectangular.txt", imaging="config-imaging.txt", sta="config-sta.txt",
        cf="config-coherence_factor.txt"):
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    args = [CLI, "--sta-config", os.path.join(PROJ, sta) if os.sep not in sta else sta,
            "--dump-config", str(tmp_path)]
    if grid:
        args += ["--grid-config", grid if os.sep in grid else os.path.join(PROJ, grid)]
    if imaging:
        args += ["--imag-config", imaging if os.sep in imaging else os.path.join(PROJ, imaging)]
    if cf:
        args += ["--cf-config", cf if os.sep in cf else os.path.join(PROJ, cf)]
    return subprocess.run(args, capture_output=True, text=True, timeout=60)


def dumped(tmp_path):
    kv = dict(l.split(" = ", 1) for l in open(tmp_path / "config.txt").read().splitlines())
    n1, n2 = map(int, kv["grid"].split())
    x = np.fromfile(tmp_path / "grid_x.f64").reshape(n1, n2)
    z = np.fromfile(tmp_path / "grid_z.f64").reshape(n1, n2)
    return kv, x, z, np.fromfile(tmp_path / "apod.f64")


def max_step_points(a, b, step):
    n = 1 + int(np.ceil((b - a) / step))
    return a + np.arange(n) * ((b - a) / (n - 1))


def test_rectangular_grid_and_settings(tmp_path):
    r = run(tmp_path)
    assert r.returncode == 0, r.stderr
    kv, x, z, apod = dumped(tmp_path)
    assert kv["num_elements"] == "32" and kv["upsampling_factor"] == "4" and kv["coherence"] == "pcf"
    lam = 1540.0 / (2 * 10.0e6)
    step = lam / 0.5
    xs = max_step_points(-10e-3, 10e-3, step)
    zs = max_step_points(5e-3, 55e-3, step)
    assert x.shape == (xs.size, zs.size) == (131, 326)
    np.testing.assert_allclose(x, np.broadcast_to(xs[:, None], x.shape), rtol=0, atol=1e-15)
    np.testing.assert_allclose(z, np.broadcast_to(zs[None, :] + 1e-3, z.shape), rtol=0, atol=1e-15)
    # the spacing never exceeds the maximum step, and one point fewer would
    assert np.diff(xs).max() <= step and (xs[-1] - xs[0]) / (xs.size - 2) > step
    np.testing.assert_allclose(apod, np.hanning(32), atol=1e-15)


def test_sectorial_grid(tmp_path):
    r = run(tmp_path, grid="config-grid-sectorial.txt")
    assert r.returncode == 0, r.stderr
    _, x, z, _ = dumped(tmp_path)
    lam = 1540.0 / (2 * 10.0e6)
    rs = max_step_points(13e-3, 30e-3, lam / 1.0)
    ang = max_step_points(np.deg2rad(-45.0), np.deg2rad(45.0), np.deg2rad(0.5))
    assert x.shape == (ang.size, rs.size) and ang.size == 181
    np.testing.assert_allclose(x, rs[None, :] * np.sin(ang)[:, None] + 2e-3, atol=1e-15)
    np.testing.assert_allclose(z, rs[None, :] * np.cos(ang)[:, None], atol=1e-15)


@pytest.mark.parametrize("desc,n,expect", [
    ("rectangular", 7, lambda n: np.ones(n)),
    ("hanning", 9, np.hanning),
    ("hanning2", 9, lambda n: np.hanning(n + 2)[1:-1]),
    ("hamming", 8, np.hamming),
    ("blackman", 10, np.blackman),
    ("blackman2", 10, lambda n: np.blackman(n + 2)[1:-1]),
    ("triangular", 7, np.bartlett),
    ("triangular", 8, np.bartlett),
    ("triangular2", 7, lambda n: np.bartlett(n + 2)[1:-1]),
    ("triangular2", 8, lambda n: np.bartlett(n + 2)[1:-1]),
    ("trapezoidal 2 4 3", 9, lambda n: np.array([0, 0.5, 1, 1, 1, 1, 2 / 3, 1 / 3, 0])),
    ("trapezoidal2 2 4 3", 9, lambda n: np.array([1 / 3, 2 / 3, 1, 1, 1, 1, 0.75, 0.5, 0.25])),
])
def test_window_functions(tmp_path, desc, n, expect):
    sta = tmp_path / "sta.txt"
    sta.write_text(open(os.path.join(PROJ, "config-sta.txt")).read().replace("num_elements = 32", f"num_elements = {n}")
                   .replace("last_tx_elem = 31", f"last_tx_elem = {n - 1}"))
    img = tmp_path / "img.txt"
    img.write_text(f"peak_offset = 0.0\nupsampling_factor = 4\nrx_apodization = {desc}\n")
    out = tmp_path / "out"
    out.mkdir()
    r = run(out, grid="config-grid-rectangular.txt", imaging=str(img), sta=str(sta))
    assert r.returncode == 0, r.stderr
    np.testing.assert_allclose(dumped(out)[3], expect(n), atol=1e-15)


@pytest.mark.parametrize("text,message", [
    ("num_elements 32\n", "Missing '=' separator at line 1"),
    ("= 3\n", "Missing key at line 1"),
    ("pitch =\n", "Missing value at line 1"),
    ("pi tch = 1\n", "Key has space at line 1"),
    ("# c\n\npitch = 1\npitch = 2\n", "Duplicate key 'pitch' at line 4"),
])
def test_key_value_format_errors(tmp_path, text, message):
    bad = tmp_path / "bad.txt"
    bad.write_text(text)
    r = run(tmp_path, grid=None, imaging=None, cf=None, sta=str(bad))
    assert r.returncode == 1 and message in r.stderr, r.stderr


@pytest.mark.parametrize("edit,message", [
    (("num_elements = 32", "num_elements = 300"), "'num_elements' must be <= 128"),
    (("pitch = 0.3e-3", "pitch = 0.3mm"), "Invalid value for key 'pitch'"),
    (("dead_zone_m = 0.0\n", ""), "Key 'dead_zone_m' not found"),
])
def test_sta_config_ranges(tmp_path, edit, message):
    sta = tmp_path / "sta.txt"
    sta.write_text(open(os.path.join(PROJ, "config-sta.txt")).read().replace(*edit))
    r = run(tmp_path, grid=None, imaging=None, cf=None, sta=str(sta))
    assert r.returncode == 1 and message in r.stderr, r.stderr


def test_grid_and_coherence_errors(tmp_path):
    g = tmp_path / "g.txt"
    g.write_text("image_grid_type = polar\n")
    r = run(tmp_path, grid=str(g))
    assert r.returncode == 1 and "Invalid grid type: polar" in r.stderr
    c = tmp_path / "c.txt"
    c.write_text("analytic_signal_coherence_factor_method = other\n")
    r = run(tmp_path, cf=str(c))
    assert r.returncode == 1 and "Invalid coherence factor method: other" in r.stderr
    w = tmp_path / "w.txt"
    w.write_text("peak_offset = 0.0\nupsampling_factor = 4\nrx_apodization = kaiser\n")
    r = run(tmp_path, imaging=str(w))
    assert r.returncode == 1 and "Invalid window function name: kaiser" in r.stderr
