"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs.  Bar (FP32 storage): max |value_gpu - value_oracle| <=
1e-4 x image peak; the CF image value*factor to the same bar.  FP16/BF16
storage: envelope error in dB against the FP64 oracle, stated per test."""
import numpy as np
import pytest

import oracle
from us_lab4a_amd import imaging, workloads
from us_lab4a_amd.sta import CoherenceFactor, HipSTAProcessor, STAConfiguration
from us_lab4a_amd import _lib

pytestmark = pytest.mark.gpu

from parity import TOL, check_reference_oracle, check_table_oracle, peak_err, say

MM = 1e-3


def torch_cuda():
    import torch
    return torch


def make_proc(wl, storage="f32", cf=None):
    cfgs = wl.sta_config()
    p = HipSTAProcessor(cfgs, imaging.ArrayAcquisition(np.zeros((1, 1, 1), np.float32)), wl.upsampling,
                        CoherenceFactor(cf or wl.cf, wl.cf_param), wl.peak_offset, wl.rx_apodization(),
                        tx_apod=wl.tx_apodization(), storage=storage)
    p.set_grid(wl.grid())
    return p


@pytest.fixture(scope="module")
def small():
    wl = workloads.sta2d_32_small()
    echoes = wl.echoes_cpu()
    return wl, echoes


def test_delay_table_fp64_rounded_once():
    wl = workloads.sta2d_128_pcf()
    say("delay table", wl.name)
    p = make_proc(wl)
    d = p.download_delays()
    g = wl.grid()
    x = g["x"].reshape(-1).astype(np.float64)
    z = g["z"].reshape(-1).astype(np.float64)
    xe = np.arange(wl.num_elements) * np.float64(np.float32(wl.pitch)) - (wl.num_elements - 1) * np.float64(np.float32(wl.pitch)) / 2
    inv = np.float64(np.float32(wl.fs)) * wl.upsampling / np.float64(np.float32(wl.c))
    ref = (np.sqrt((x[None, :] - xe[:, None]) ** 2 + z[None, :] ** 2) * inv)
    # one rounding to FP32: within half an ulp of the FP64 value (+ FP64 noise)
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(d.astype(np.float64) - ref) <= 0.5 * ulp + 1e-9)
    p.close()


@pytest.mark.parametrize("method", [1, 2], ids=["library_fft", "fused_lds"])
def test_prepare_matches_oracle(small, method):
    """Both preparation paths (hipFFT and the fused on-chip kernel) against
    the FP64 oracle: interpolation, DC removal and analytic signal."""
    wl, echoes = small
    say("prepare", wl.name, "method", method)
    p = make_proc(wl)
    p.set_prep(method)
    p.process_echoes(echoes)
    a_gpu = p.download_analytic()
    a32 = oracle.prepare(oracle.cfg_for(wl), echoes, np.float32)
    a64 = oracle.prepare(oracle.cfg_for(wl), echoes, np.float64)
    m = np.abs(a64).max()
    assert np.abs(a_gpu - a64).max() / m <= 2e-6
    assert np.abs(a32 - a64).max() / m <= 2e-6
    p.close()


@pytest.mark.parametrize("cf", ["none", "phase_coherence_factor", "sign_coherence_factor"])
def test_das_matches_oracle_on_oracle_tensor(small, cf):
    """The oracle's FP32 tensor uploaded as is: value within 1e-4 of the peak
    of the reference-equivalent oracle, and value and CF image within 1e-4
    at every pixel of the oracle fed the GPU's delay table."""
    wl, echoes = small
    say("das", wl.name, cf)
    a32 = oracle.prepare(oracle.cfg_for(wl), echoes, np.float32)
    wl_cf = workloads.sta2d_32_small()
    wl_cf.cf = cf
    ocfg = oracle.cfg_for(wl_cf)
    xz = imaging.grid_xz(wl_cf.grid())
    p = make_proc(wl_cf)
    p.upload_analytic(a32)
    d_gpu = p.download_delays()
    import torch
    n1, n2 = wl_cf.n_x, wl_cf.n_z
    dv = torch.empty(n1 * n2, dtype=torch.float32, device="cuda")
    df = torch.empty(n1 * n2, dtype=torch.float32, device="cuda")
    p.das_device(dv, df)
    torch.cuda.synchronize()
    v = dv.cpu().numpy().reshape(n1, n2)
    f = df.cpu().numpy().reshape(n1, n2)
    p.close()
    rx, tx = wl_cf.rx_apodization(), wl_cf.tx_apodization()
    check_table_oracle(v, f, ocfg, a32, d_gpu, rx, tx, f"cfg1 {cf}")
    check_reference_oracle(v, f, ocfg, a32, xz, d_gpu, rx, tx, f"cfg1 {cf}")


@pytest.mark.parametrize("cf", ["none", "phase_coherence_factor"])
def test_full_frame_matches_oracle(small, cf):
    wl, echoes = small
    wl.cf = cf
    say("frame", wl.name, cf)
    v_o, f_o, _, _ = oracle.process(oracle.cfg_for(wl), echoes, imaging.grid_xz(wl.grid()), wl.rx_apodization(),
                                     wl.tx_apodization())
    cfgs = wl.sta_config()
    grid = wl.grid()
    acq = imaging.ArrayAcquisition(echoes)
    p = HipSTAProcessor(cfgs, acq, wl.upsampling, CoherenceFactor(cf, 1.0), 0.0, wl.rx_apodization(),
                        tx_apod=wl.tx_apodization())
    p.prepare(0)
    p.process(grid)
    assert peak_err(grid["value"], v_o) <= TOL
    if cf != "none":
        assert peak_err(grid["value"] * grid["factor"], v_o * f_o) <= TOL
    # determinism: a second frame is bit-identical
    g2 = wl.grid()
    p.process(g2)
    assert np.array_equal(g2["value"], grid["value"]) and np.array_equal(g2["factor"], grid["factor"])
    p.close()
    wl.cf = "none"


@pytest.mark.parametrize("storage,limit_db", [("f16", -60.0), ("bf16", -40.0)])
def test_reduced_storage_db_error_vs_fp64(small, storage, limit_db):
    wl, echoes = small
    say("storage", storage)
    a64 = oracle.prepare(oracle.cfg_for(wl), echoes, np.float64)
    xz = imaging.grid_xz(wl.grid())
    v64, _ = oracle.das(oracle.cfg_for(wl), a64, xz, wl.rx_apodization(), wl.tx_apodization())
    p = make_proc(wl, storage=storage)
    v, _ = p.process_echoes(echoes)
    err = workloads.db_error(v, v64)
    say(storage, "envelope error vs FP64 (dB):", round(err, 1))
    assert err <= limit_db
    p32 = make_proc(wl)
    v32, _ = p32.process_echoes(echoes)
    err32 = workloads.db_error(v32, v64)
    say("f32 envelope error vs FP64 (dB):", round(err32, 1))
    assert err32 <= -80.0
    p.close()
    p32.close()


def test_graded_shape_columns_match_oracle():
    """128 elements, 128x128x4096 echoes (GPU-synthesised), U = 8, PCF, grid
    512x1024: prep on a channel subset and DAS on a column subset against the
    oracle, with the full frame run on the GPU."""
    import torch
    wl = workloads.sta2d_128_pcf()
    say("graded shape", wl.name)
    d_echoes = torch.empty((wl.num_tx, wl.num_elements, wl.num_samples), dtype=torch.float32, device="cuda")
    wl.echoes_gpu(d_echoes)
    torch.cuda.synchronize()
    p = make_proc(wl)
    npix = wl.n_x * wl.n_z
    dv = torch.empty(npix, dtype=torch.float32, device="cuda")
    df = torch.empty(npix, dtype=torch.float32, device="cuda")
    p.process_device(d_echoes, wl.num_samples, dv, df)
    torch.cuda.synchronize()
    say("graded shape: frame done")
    v = dv.cpu().numpy().reshape(wl.n_x, wl.n_z)
    f = df.cpu().numpy().reshape(wl.n_x, wl.n_z)
    assert np.isfinite(v).all() and (f >= 0).all() and (f <= 1).all()
    echoes = d_echoes.cpu().numpy()
    a_gpu = p.download_analytic()
    say("graded shape: tensor downloaded")
    # prep parity on a channel subset (oracle FP64)
    sub = echoes[[0, 63, 127]][:, [0, 64, 127]]
    cfg_sub = oracle.cfg_for(wl)
    cfg_sub.num_elements = 3
    cfg_sub.last_tx_elem = 2
    a_sub = oracle.prepare(cfg_sub, sub, np.float64)
    ref = a_gpu[[0, 63, 127]][:, [0, 64, 127]]
    assert np.abs(ref - a_sub).max() / np.abs(a_sub).max() <= 2e-6
    # DAS parity on a column subset: the oracle over the GPU's tensor, with
    # the GPU's delay table (value and CF image at every pixel) and with its
    # own FP32 distances (value; CF image except where the distances alone
    # move it)
    cols = [0, 137, 255, 256, 400, 511]
    xz = np.ascontiguousarray(imaging.grid_xz(wl.grid())[cols])
    pix = (np.asarray(cols)[:, None] * wl.n_z + np.arange(wl.n_z)[None, :]).reshape(-1)
    d_sub = p.download_delays_at(pix)
    peak = float(np.abs(v).max())
    cfg = oracle.cfg_for(wl)
    check_table_oracle(v[cols], f[cols], cfg, a_gpu, d_sub, wl.rx_apodization(), None, "graded columns", peak=peak)
    check_reference_oracle(v[cols], f[cols], cfg, a_gpu, xz, d_sub, wl.rx_apodization(), None, "graded columns",
                           peak=peak)
    p.close()


@pytest.mark.parametrize("tz,storage,chunking", [(0, "f32", 4), (2, "f32", 4), (3, "f32", 4), (0, "f16", 4),
                                                 (0, "bf16", 4), (5, "bf16", 8 | (150 << 16)),
                                                 (6, "f16", 8 | (150 << 16))])
def test_tiled_kernel_equals_gather_kernel(small, tz, storage, chunking):
    """The LDS-tiled kernel (256-lane tiles 16x16, 8x32, 32x8 and 512-lane
    tiles 16x32, 32x16) uses the gather kernel's arithmetic and, with a
    coherence factor, its summation order: results are bit-identical."""
    import torch
    wl, echoes = small
    wl.cf = "phase_coherence_factor"
    say("tiled vs gather, tile", tz, storage)
    p = make_proc(wl, storage=storage)
    p.process_echoes(echoes)
    n = wl.n_x * wl.n_z
    out = {}
    for variant, p0 in ((1, 0), (2, tz)):
        p.set_kernel(variant, p0, chunking)
        dv = torch.empty(n, dtype=torch.float32, device="cuda")
        df = torch.empty(n, dtype=torch.float32, device="cuda")
        p.das_device(dv, df)
        torch.cuda.synchronize()
        out[variant] = (dv.cpu().numpy(), df.cpu().numpy())
        assert (p.last_kernel()[0] >= 0) == (variant == 2)   # the tiled kernel ran, not the fallback
    assert np.array_equal(out[1][0], out[2][0]) and np.array_equal(out[1][1], out[2][1])
    p.close()
    wl.cf = "none"


def test_autotuned_kernel_equals_gather_kernel(small):
    """The automatic variant times every tile shape and LDS budget on its
    first frame and keeps the fastest: that frame and the tuned frames after
    it are bit-identical to the gather kernel's."""
    import torch
    wl, echoes = small
    wl.cf = "phase_coherence_factor"
    p = make_proc(wl)
    p.process_echoes(echoes)
    n = wl.n_x * wl.n_z
    outs = []
    for variant in (0, 0, 1):
        p.set_kernel(variant)
        dv = torch.empty(n, dtype=torch.float32, device="cuda")
        df = torch.empty(n, dtype=torch.float32, device="cuda")
        p.das_device(dv, df)
        torch.cuda.synchronize()
        outs.append((dv.cpu().numpy(), df.cpu().numpy()))
        if variant == 0:
            tuned = p.tuned_kernel()
            say("autotuner chose", tuned)
            # a shape with a budget of 52, 76 or 150 KiB and 4 or 8 pairs per
            # chunk, or (0) the smallest budget in which the windows fit
            assert tuned[0] in (1, 2, 3, 5, 6) and (tuned[1] == 0 or (
                (tuned[1] >> 16) in (52, 76, 150) and (tuned[1] & 0xFF) in (4, 8)))
    for v, f in outs[:2]:
        assert np.array_equal(v, outs[2][0]) and np.array_equal(f, outs[2][1])
    p.close()
    wl.cf = "none"


def test_error_paths_raise_reference_style_errors():
    wl = workloads.tiny()
    say("error paths")
    cfgs = wl.sta_config()
    p = HipSTAProcessor(cfgs, None, wl.upsampling, CoherenceFactor(), 0.0, wl.rx_apodization())
    with pytest.raises(_lib.InvalidStateError):
        p.process_echoes(np.zeros((wl.num_tx, wl.num_elements, wl.num_samples), np.float32))
    with pytest.raises(_lib.InvalidValueError):
        p.set_rx_apod(np.ones(wl.num_elements + 1, np.float32))
    p.close()
    cfgs.dead_zone_m = 1.0   # 1 m dead zone >= record length
    p = HipSTAProcessor(cfgs, None, wl.upsampling, CoherenceFactor(), 0.0, wl.rx_apodization())
    p.set_grid(wl.grid())
    with pytest.raises(_lib.InvalidValueError, match="dead zone"):
        p.process_echoes(np.zeros((wl.num_tx, wl.num_elements, wl.num_samples), np.float32))
    p.close()


@pytest.mark.parametrize("shape", [(16, 301, 1), (8, 1000, 3), (4, 4096, 8)], ids=["U1_odd", "U3", "U8_4096"])
def test_fused_prepare_shapes(shape):
    """Fused preparation at other FFT sizes (odd record, radix-3/5 heavy
    sizes, the 4096-sample U = 8 channel of the 128-element frame) against
    the FP64 oracle."""
    N, ns, up = shape
    wl = workloads.tiny(num_elements=N, num_samples=ns, upsampling=up)
    echoes = wl.echoes_cpu()
    say("fused prepare", shape)
    p = make_proc(wl)
    p.set_prep(2)
    p.process_echoes(echoes)
    a_gpu = p.download_analytic()
    p.close()
    a64 = oracle.prepare(oracle.cfg_for(wl), echoes, np.float64)
    m = np.abs(a64).max()
    err = np.abs(a_gpu - a64).max() / m
    say("fused prepare error / peak", err)
    assert err <= 2e-6


def _bf16_round(x):
    """float32 -> bfloat16 (round to nearest even) -> float32, in numpy."""
    b = x.astype(np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
    return (b.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("shape", [(16, 301, 1), (4, 4096, 8)], ids=["U1_odd", "U8_4096"])
@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_fused_reduced_storage_is_scaled_fp32(shape, storage):
    """The fused kernel writes FP16/BF16 directly: each channel must hold
    round(z / s) with s = max(|re|, |im|) of the channel's FP32 analytic
    signal, exactly as converting the FP32 tensor afterwards would."""
    N, ns, up = shape
    wl = workloads.tiny(num_elements=N, num_samples=ns, upsampling=up)
    echoes = wl.echoes_cpu()
    say("fused reduced storage", storage, shape)
    out = {}
    for st in ("f32", storage):
        p = make_proc(wl, storage=st)
        p.set_prep(2)
        p.process_echoes(echoes)
        out[st] = p.download_analytic()
        p.close()
    t = out["f32"]
    s = np.abs(t).max(axis=(2, 3), keepdims=True).astype(np.float32)
    s = np.where(s > 0, s, np.float32(1.0)).astype(np.float32)
    inv = (np.float32(1.0) / s).astype(np.float32)
    q = (t * inv).astype(np.float32)
    r = q.astype(np.float16).astype(np.float32) if storage == "f16" else _bf16_round(q)
    expect = (r * s).astype(np.float32)
    assert np.array_equal(out[storage], expect)


@pytest.mark.parametrize("elements,cf", [(192, "none"), (192, "phase_coherence_factor"), (192, "sign_coherence_factor"),
                                         (200, "phase_coherence_factor"), (136, "sign_coherence_factor"),
                                         (300, "phase_coherence_factor")])
def test_tiled_kernel_over_128_transmits(elements, cf):
    """More than 128 transmits: the tiled kernel runs them in equal blocks of
    whole 8-pair chunks (192: two of 96; 200: two of 104, the second with 8
    padding transmits; 136: two of 72; 300: three of 104), block after block
    over all receive elements, each receive element's partial sum kept in
    HBM between blocks.  With a coherence factor the sum per receive element
    runs over the transmits in order, the gather kernel's fma chain, so the
    two are bit-identical (without one the gather kernel sums
    transmit-outer, like the reference's no-CF loop, and the two agree to
    the tolerance).  Both 512-lane tile shapes give the same image; a
    256-lane tile is refused and the rx-parallel form runs (per-rx sums a
    lane per (pixel, rx), then the statistics in rx order), equal to the
    tiles bit for bit; the frame matches the oracle."""
    import torch
    wl = workloads.tiny(num_elements=elements, num_samples=2048, upsampling=2, cf=cf, n_x=24, n_z=32)
    echoes = wl.echoes_cpu()
    say("over 128 transmits", elements, cf)
    p = make_proc(wl)
    p.process_echoes(echoes)
    n = wl.n_x * wl.n_z
    out = {}
    for variant, p0 in ((1, 0), (2, 5), (2, 6), (2, 1)):
        p.set_kernel(variant, p0, 0)
        dv = torch.empty(n, dtype=torch.float32, device="cuda")
        df = torch.empty(n, dtype=torch.float32, device="cuda")
        p.das_device(dv, df)
        torch.cuda.synchronize()
        out[p0] = (dv.cpu().numpy().reshape(wl.n_x, wl.n_z), df.cpu().numpy().reshape(wl.n_x, wl.n_z))
        # 512-lane shapes run tiled; the 256-lane one falls back to the rx-parallel form
        expect = p0 if p0 in (5, 6) else (-1 if variant == 1 else -2)
        assert p.last_kernel()[0] == expect, (p0, p.last_kernel())
    p.set_kernel(0)
    v_a, f_a = p.process_echoes(echoes)
    tuned = p.tuned_kernel()
    a_gpu = p.download_analytic()
    d_gpu = p.download_delays()
    p.close()
    say("autotuner chose", tuned)
    assert tuned[0] in (5, 6)
    assert np.array_equal(out[5][0], out[6][0]) and np.array_equal(out[5][1], out[6][1])
    assert np.array_equal(out[1][0], out[5][0]) and np.array_equal(out[1][1], out[5][1])
    assert np.array_equal(v_a.reshape(wl.n_x, wl.n_z), out[5][0])
    if cf == "none":
        assert peak_err(out[5][0], out[0][0]) <= TOL
    else:
        assert np.array_equal(out[0][0], out[5][0]) and np.array_equal(out[0][1], out[5][1])
    ocfg = oracle.cfg_for(wl)
    a32 = oracle.prepare(ocfg, echoes, np.float32)
    xz = imaging.grid_xz(wl.grid())
    check_table_oracle(out[5][0], out[5][1], ocfg, a_gpu, d_gpu, wl.rx_apodization(), wl.tx_apodization(),
                       f"over 128 transmits {elements} {cf}")
    check_reference_oracle(out[5][0], out[5][1], ocfg, a32, xz, d_gpu, wl.rx_apodization(), wl.tx_apodization(),
                           f"over 128 transmits {elements} {cf}")


@pytest.mark.parametrize("cf,storage", [("phase_coherence_factor", "f32"), ("sign_coherence_factor", "f16"),
                                        ("none", "f32")])
def test_rx_parallel_where_no_tile_fits(cf, storage):
    """A grid so coarse (24 x 24 points over 20 x 55 mm) that a tile's
    window passes 64 KiB: no tile shape fits, and the automatic variant runs
    the rx-parallel form (per-rx sums a lane per (pixel, rx), the statistics
    a lane per pixel in rx order).  Its per-rx sums are the tiled kernel's fma
    chain, so with a coherence factor it equals the gather kernel bit for
    bit (without one the gather kernel sums transmit-outer: within the
    tolerance); the frame matches the oracle on the GPU's own tensor and
    delay table."""
    import torch
    wl = workloads.Workload2D(name="coarse", num_elements=64, num_samples=4096, upsampling=8, cf=cf,
                              grid_x=(-10 * MM, 10 * MM), grid_z=(5 * MM, 60 * MM), n_x=24, n_z=24, points=6,
                              point_x=(-8 * MM, 8 * MM), point_z=(10 * MM, 50 * MM), seed_points=77, seed_noise=78)
    say("no tile fits", cf, storage)
    p = make_proc(wl, storage=storage)
    p.process_echoes(wl.echoes_cpu())
    n = wl.n_x * wl.n_z
    out = {}
    for variant in (0, 1):
        p.set_kernel(variant)
        dv = torch.empty(n, dtype=torch.float32, device="cuda")
        df = torch.empty(n, dtype=torch.float32, device="cuda")
        p.das_device(dv, df)
        torch.cuda.synchronize()
        out[variant] = (dv.cpu().numpy(), df.cpu().numpy())
        assert p.last_kernel()[0] == (-2 if variant == 0 else -1), (variant, p.last_kernel())
    a_gpu = p.download_analytic()
    d_gpu = p.download_delays()
    p.close()
    if cf == "none":
        assert peak_err(out[0][0], out[1][0]) <= TOL
    else:
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    check_table_oracle(out[0][0], out[0][1], oracle.cfg_for(wl), a_gpu, d_gpu, wl.rx_apodization(),
                       wl.tx_apodization(), f"rx-parallel {cf} {storage}",
                       branch_rel=1e-6 if storage != "f32" else 0.0)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_tiled_sixteen_pair_chunks(storage):
    """512-lane 16x32 tiles over a 128-transmit block with 16 pairs per chunk
    (two windows per wave per chunk, half the barriers of 8): the same
    per-rx fma chains, so the images equal the 8-pair chunks' and the gather
    kernel's bit for bit (PCF).  Pixel pitch as the 128-element frame's
    (0.074 x 0.059 mm), so the windows fit two 16-pair buffers."""
    import torch
    wl = workloads.Workload2D(name="tc16", num_elements=128, num_samples=2048, upsampling=8,
                              cf="phase_coherence_factor", grid_x=(-4.75 * MM, 4.75 * MM), grid_z=(5 * MM, 20 * MM),
                              n_x=128, n_z=256, points=12, point_x=(-4 * MM, 4 * MM), point_z=(6 * MM, 19 * MM),
                              seed_points=91, seed_noise=92)
    say("16 pairs per chunk", storage)
    p = make_proc(wl, storage=storage)
    d = torch.empty((wl.num_tx, wl.num_elements, wl.num_samples), dtype=torch.float32, device="cuda")
    wl.echoes_gpu(d)
    torch.cuda.synchronize()
    p.prepare_signals_device(d, wl.num_samples)
    torch.cuda.synchronize()
    n = wl.n_x * wl.n_z
    out = {}
    for name, variant, p0, p1 in (("gather", 1, 0, 0), ("8 pairs", 2, 5, 8 | (150 << 16)),
                                  ("16 pairs", 2, 5, 16 | (150 << 16))):
        p.set_kernel(variant, p0, p1)
        dv = torch.empty(n, dtype=torch.float32, device="cuda")
        df = torch.empty(n, dtype=torch.float32, device="cuda")
        p.das_device(dv, df)
        torch.cuda.synchronize()
        out[name] = (dv.cpu().numpy(), df.cpu().numpy())
        ran = p.last_kernel()
        if variant == 2:
            assert ran[0] == 5 and (ran[1] & 0xFF) == (p1 & 0xFF), (name, ran)
    p.close()
    for name in ("8 pairs", "16 pairs"):
        assert np.array_equal(out[name][0], out["gather"][0]), name
        assert np.array_equal(out[name][1], out["gather"][1]), name


@pytest.fixture(scope="module")
def graded_fp64():
    """The 128-element PCF frame's echoes (GPU-synthesised) and the FP64
    oracle's value and factor on a column subset of the 512x1024 grid."""
    import torch
    wl = workloads.sta2d_128_pcf()
    d_echoes = torch.empty((wl.num_tx, wl.num_elements, wl.num_samples), dtype=torch.float32, device="cuda")
    wl.echoes_gpu(d_echoes)
    torch.cuda.synchronize()
    echoes = d_echoes.cpu().numpy()
    say("graded FP64 oracle: preparing 16384 channels in FP64")
    a64 = oracle.prepare(oracle.cfg_for(wl), echoes, np.float64)
    cols = [0, 137, 255, 256, 400, 511]
    xz = np.ascontiguousarray(imaging.grid_xz(wl.grid())[cols])
    v64, f64 = oracle.das(oracle.cfg_for(wl), a64, xz, wl.rx_apodization())
    del a64
    say("graded FP64 oracle: done")
    return wl, d_echoes, cols, v64, f64


@pytest.mark.parametrize("storage,limit_db,cf_limit_db", [("f32", -80.0, -55.0), ("f16", -60.0, -40.0),
                                                          ("bf16", -40.0, -28.0)])
def test_graded_shape_db_error_vs_fp64(graded_fp64, storage, limit_db, cf_limit_db):
    """Precision variant of the 128-element PCF frame: FP32/FP16/BF16 storage
    (per-channel scale), FP32 accumulation; envelope (value) and PCF image
    (value * factor) error in dB against the FP64 oracle on a column subset
    (peaks taken over the subset).  The envelope meets the storage's bar
    (measured -103.7 / -96.9 / -80.3 dB); the PCF image has its own, looser
    bar (measured -60.3 / -43.5 / -31.9 dB): the factor depends on the
    phases of the per-rx sums, so a small error in a weak per-rx sum, or a
    phase near the atan2 cut, moves the factor of a pixel by a step."""
    import torch
    wl, d_echoes, cols, v64, f64 = graded_fp64
    p = make_proc(wl, storage=storage)
    npix = wl.n_x * wl.n_z
    dv = torch.empty(npix, dtype=torch.float32, device="cuda")
    df = torch.empty(npix, dtype=torch.float32, device="cuda")
    p.process_device(d_echoes, wl.num_samples, dv, df)
    torch.cuda.synchronize()
    v = dv.cpu().numpy().reshape(wl.n_x, wl.n_z)[cols]
    f = df.cpu().numpy().reshape(wl.n_x, wl.n_z)[cols]
    p.close()
    err_v = workloads.db_error(v, v64)
    err_cf = workloads.db_error(v.astype(np.float64) * f, v64 * f64)
    say("graded", storage, "envelope error vs FP64 (dB):", round(err_v, 1), "; PCF image:", round(err_cf, 1))
    assert err_v <= limit_db
    assert err_cf <= cf_limit_db
