#!/usr/bin/env python3
"""STA/TFM imaging benchmark on one MI355X.

Workload (default): 128 elements, pitch 0.3 mm, 5 MHz pulse, fs 40 MHz,
128 tx x 128 rx x 4096 samples from 200 point scatterers + speckle
(5000/cm^2, Rayleigh) + AWGN 30 dB, generated on the GPU from fixed seeds;
U = 8 (L = 32768), PCF gamma = 1, Hann rx apodization, grid 512 x 1024 over
x [-19, 19] mm, z [5, 65] mm.

A step is one frame: raw echoes resident in HBM -> upsampling, DC removal,
analytic signal -> delay-and-sum with PCF -> value/factor images in HBM
(the delay table is built once per geometry and kept resident, as the
serving path does).  `value` = pixel*pairs per frame / frame time.

Prints one JSON line.  --gpus N > 1 is refused (single-GPU build).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="sta2d_128_pcf")
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--kernel", type=int, default=0, help="DAS variant (0 auto, 1 gather, 2 tiled)")
    ap.add_argument("--graph", action="store_true",
                    help="replay steady frames as one hipGraph (stage times then cover the whole frame)")
    ap.add_argument("--tile", type=int, nargs=2, metavar=("SHAPE", "CHUNKING"), default=None,
                    help="fixed tiled-DAS parameters instead of the autotuner (profiling runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU oracle leg (baseline + FP64 errors)")
    ap.add_argument("--no-headline", action="store_true", help="skip the FP16 batched comparison")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-reps", type=int, default=10, help="timed repetitions of the CPU sample (after 1 warm-up)")
    ap.add_argument("--mode", default="frame", choices=["frame", "batch", "serve"],
                    help="frame: the headline line; batch: FP16 frames through the serving engine vs plain "
                         "FP32 frames; serve: Poisson arrivals of a mixed-shape job stream")
    ap.add_argument("--inflight", type=int, default=4)
    ap.add_argument("--rates", default="5,20,50,200,1000,2000", help="serve mode: arrival rates (frames/s)")
    ap.add_argument("--jobs", type=int, default=160, help="serve mode: most jobs per rate (at most 8 s of arrivals)")
    return ap.parse_args()


def algorithmic_bytes(wl, storage):
    """Compulsory HBM bytes of one DAS launch: analytic tensor read once,
    FP32 delay table read once, value+factor written, apodization read."""
    npix = wl.n_x * wl.n_z
    bc = 8 if storage == "f32" else 4
    cf = 0 if wl.cf == "none" else 1
    return (wl.num_tx * wl.num_elements * wl.num_samples * wl.upsampling * bc
            + wl.num_elements * npix * 4 + npix * 4 * (1 + cf) + 4 * wl.num_elements)


def measured_traffic(wl, storage):
    """HBM-side bytes per DAS launch from the committed rocprofv3 counter
    summary for this workload and storage (FETCH_SIZE + WRITE_SIZE, each
    scaled by a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"das_traffic_{wl.name}_{storage}.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def cpu_oracle_leg(wl, threads, reps, d_echoes, cols, images):
    """The CPU oracle (FP32 restatement of the reference TBB processor, its own
    FFT instead of FFTW) on this host: a bounded sample of the frame timed as
    1 warm-up + `reps` repetitions (prep of a tx subset, DAS of a column
    subset over all transmits), scaled to one frame; and, as the checker of
    the GPU images, the FP64 oracle's envelope on the column subset `cols`,
    from which the envelope error in dB of every GPU image in `images`
    ({name: value image [n_x][n_z]}) is stated."""
    import platform
    import oracle
    from us_lab4a_amd import imaging, workloads
    ntx_s, ncol_s = 16, 32
    echoes = d_echoes.cpu().numpy()
    cfg = oracle.cfg_for(wl)
    cfg_s = oracle.cfg_for(wl)
    cfg_s.last_tx_elem = cfg_s.first_tx_elem + ntx_s - 1
    xz_all = imaging.grid_xz(wl.grid())
    scols = np.linspace(0, wl.n_x - 1, ncol_s).astype(int)
    xz_s = np.ascontiguousarray(xz_all[scols])
    t_prep, t_das = [], []
    a_full = None
    for k in range(reps + 1):
        t0 = time.perf_counter()
        oracle.prepare(cfg_s, echoes[:ntx_s], np.float32, nthreads=threads)
        t1 = time.perf_counter()
        if a_full is None:
            a_full = oracle.prepare(cfg, echoes, np.float32, nthreads=threads)   # the DAS sample reads every tx
        t2 = time.perf_counter()
        oracle.das(cfg, a_full, xz_s, wl.rx_apodization(), wl.tx_apodization(), nthreads=threads)
        t3 = time.perf_counter()
        if k > 0:   # the first repetition is the warm-up
            t_prep.append(t1 - t0)
            t_das.append(t3 - t2)
    del a_full
    prep_scale = wl.num_tx / ntx_s
    das_scale = wl.n_x / ncol_s
    frame = np.array(t_prep) * prep_scale + np.array(t_das) * das_scale
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        model = platform.processor()
    base = {
        "value": wl.pixel_pairs / float(frame.mean()) / 1e9,
        "unit": "Gpixel*pair/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle FP32 restatement of VectorialSTAProcessor (own mixed-radix FFT and direct-sum FIR, not "
                   f"FFTW): prep of {ntx_s}/{wl.num_tx} tx and DAS+{wl.cf} of {ncol_s}/{wl.n_x} columns x all "
                   f"{wl.num_tx} tx, 1 warm-up + {reps} repetitions, each scaled to one frame"),
        "frame_s": {"mean": float(frame.mean()), "min": float(frame.min()), "std": float(frame.std())},
        "prep_s_per_frame": {"mean": float(np.mean(t_prep) * prep_scale), "min": float(np.min(t_prep) * prep_scale),
                             "std": float(np.std(t_prep) * prep_scale)},
        "das_s_per_frame": {"mean": float(np.mean(t_das) * das_scale), "min": float(np.min(t_das) * das_scale),
                            "std": float(np.std(t_das) * das_scale)},
        "nproc": os.cpu_count(),
        "cpu_model": model,
    }
    print(f"[bench] cpu oracle frame {frame.mean():.2f} s; FP64 reference on {len(cols)} columns", file=sys.stderr,
          flush=True)
    a64 = oracle.prepare(cfg, echoes, np.float64, nthreads=threads)
    v64, _ = oracle.das(cfg, a64, np.ascontiguousarray(xz_all[cols]), wl.rx_apodization(), wl.tx_apodization(),
                        nthreads=threads)
    del a64
    errs = {name: round(workloads.db_error(img[cols], v64), 2) for name, img in images.items()}
    return base, errs


def fp16_batched(wl, d_echoes, steps, warmup, inflight):
    """FP16 storage (per-channel scale, FP32 accumulation) with frames
    batched through the serving engine, `inflight` frames on the device at
    once: frames/s over `steps` frames, engine latency p50/p99, and the last
    value image (host)."""
    import torch
    from us_lab4a_amd.serve import Engine, Job2D
    from us_lab4a_amd.sta import CoherenceFactor
    npix = wl.n_x * wl.n_z
    outs = [(torch.empty(npix, device="cuda"), torch.empty(npix, device="cuda")) for _ in range(inflight)]

    def job(k):
        v, f = outs[k % len(outs)]
        return Job2D(config=wl.sta_config(), upsampling=wl.upsampling, coherence=CoherenceFactor(wl.cf, wl.cf_param),
                     grid=wl.grid(), rx_apod=wl.rx_apodization(), tx_apod=wl.tx_apodization(), echoes=d_echoes,
                     value=v, factor=f, storage="f16", peak_offset=wl.peak_offset)
    with Engine(max_inflight=inflight, max_queued=4 * steps + 8) as eng:
        for t in [eng.submit(job(k)) for k in range(max(warmup, inflight))]:
            assert eng.wait(t)["ok"]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tickets = [eng.submit(job(k)) for k in range(steps)]
        lat = [eng.wait(t)["latency_ms"] for t in tickets]
        torch.cuda.synchronize()
        fps = steps / (time.perf_counter() - t0)
    v = outs[(steps - 1) % len(outs)][0].cpu().numpy().reshape(wl.n_x, wl.n_z)
    return {"frames_per_s": fps, "p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))}, v


def achievable_hbm_gbs():
    """Streaming-read rate of a 4 GiB device buffer (HIP kernel, events)."""
    import ctypes
    import torch
    from us_lab4a_amd._lib import check, lib
    buf = torch.empty(4 << 30, dtype=torch.uint8, device="cuda")
    buf.fill_(1)
    out = ctypes.c_double(0.0)
    check(lib().mista_read_bandwidth_device(ctypes.c_void_p(buf.data_ptr()), buf.numel(), 10, ctypes.byref(out), None))
    del buf
    return out.value


def batch_mode(args):
    """FP16 storage + engine batching (several frames in flight, overlapping
    prep and DAS) against the plain FP32 path one frame at a time, same
    frame, same echoes."""
    import torch
    from us_lab4a_amd import workloads
    from us_lab4a_amd.serve import Engine, Job2D
    from us_lab4a_amd.sta import CoherenceFactor, HipSTAProcessor
    wl = workloads.WORKLOADS[args.workload]()
    npix = wl.n_x * wl.n_z
    dev = torch.device("cuda:0")
    d_echoes = torch.empty((wl.num_tx, wl.num_elements, wl.num_samples), dtype=torch.float32, device=dev)
    wl.echoes_gpu(d_echoes)
    torch.cuda.synchronize()
    # plain FP32, one frame at a time
    proc = HipSTAProcessor(wl.sta_config(), None, wl.upsampling, CoherenceFactor(wl.cf, wl.cf_param),
                           wl.peak_offset, wl.rx_apodization(), tx_apod=wl.tx_apodization(), storage="f32")
    proc.set_grid(wl.grid())
    v32 = torch.empty(npix, device=dev)
    f32 = torch.empty(npix, device=dev)
    for _ in range(args.warmup):
        proc.process_device(d_echoes, wl.num_samples, v32, f32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        proc.process_device(d_echoes, wl.num_samples, v32, f32)
    torch.cuda.synchronize()
    fps32 = args.steps / (time.perf_counter() - t0)
    proc.close()
    # FP16 storage, frames through the engine, args.inflight in flight
    outs = [(torch.empty(npix, device=dev), torch.empty(npix, device=dev)) for _ in range(args.inflight)]

    def job(k):
        v, f = outs[k % len(outs)]
        return Job2D(config=wl.sta_config(), upsampling=wl.upsampling, coherence=CoherenceFactor(wl.cf, wl.cf_param),
                     grid=wl.grid(), rx_apod=wl.rx_apodization(), tx_apod=wl.tx_apodization(), echoes=d_echoes,
                     value=v, factor=f, storage="f16", peak_offset=wl.peak_offset)
    with Engine(max_inflight=args.inflight, max_queued=4 * args.steps) as eng:
        for t in [eng.submit(job(k)) for k in range(max(args.warmup, args.inflight))]:
            assert eng.wait(t)["ok"]
        t0 = time.perf_counter()
        tickets = [eng.submit(job(k)) for k in range(args.steps)]
        lat = [eng.wait(t)["latency_ms"] for t in tickets]
        fps16 = args.steps / (time.perf_counter() - t0)
    torch.cuda.synchronize()
    v16 = outs[(args.steps - 1) % len(outs)][0]
    ref = v32.double()
    err_db = 20 * math.log10(float((v16.double() - ref).abs().max() / ref.abs().max()))
    out = {"metric": "frames/s: FP16 storage + batched engine vs plain FP32", "value": round(fps16, 3),
           "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "fp32_plain_frames_per_s": round(fps32, 3),
           "speedup_vs_fp32_plain": round(fps16 / fps32, 3), "inflight": args.inflight,
           "envelope_error_db_vs_fp32_gpu": round(err_db, 2),
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                          "p99": round(float(np.percentile(lat, 99)), 3)},
           "config": {"workload": wl.name, "storage": "f16"}, "data": "synthetic"}
    print(json.dumps(out), flush=True)


def serve_mode(args):
    """Poisson arrivals of 2-D jobs with seeded uneven shapes (elements in
    {32,64,128,192,256}, samples in {2048,4096,8192}, grids 128x128 ..
    512x1024, coherence none/PCF/SCF), FP16 storage, U = 8, through the
    serving engine.  Every shape is built, tuned and warmed before the
    measured windows; each window reports how many handles were built and
    tuners run inside it (both should be 0).  Per arrival rate: jobs served
    and refused, frames/s, Gpixel*pair/s, p50/p99 latency, and latency over
    the job's own device time alone; the capacity is the back-to-back rate
    of the same jobs (their summed device time alone)."""
    import torch
    from us_lab4a_amd import workloads, _lib
    from us_lab4a_amd.serve import Engine, Job2D
    from us_lab4a_amd.sta import CoherenceFactor
    MM = 1e-3
    rng = np.random.default_rng(5)
    cfs = ["none", "phase_coherence_factor", "sign_coherence_factor"]
    grids = [(128, 128), (256, 256), (512, 512), (512, 1024)]
    catalog = []
    for i in range(8):   # a seeded catalog of shapes; jobs draw from it
        N = int(rng.choice([32, 64, 128, 192, 256]))
        ns = int(rng.choice([2048, 4096, 8192]))
        gx, gz = grids[int(rng.integers(len(grids)))]
        half = N * 0.3 * MM / 2
        zmax = min(65 * MM, ns / 40e6 * 1540 / 2 * 0.9)
        wl = workloads.Workload2D(name=f"serve{i}", num_elements=N, num_samples=ns, upsampling=8, cf=cfs[i % 3],
                                  grid_x=(-half, half), grid_z=(5 * MM, zmax), n_x=gx, n_z=gz, points=50,
                                  point_x=(-half, half), point_z=(5 * MM, zmax), speckle_density=2000.0,
                                  seed_points=300 + i, seed_speckle=400 + i, seed_noise=500 + i)
        d = torch.empty((wl.num_tx, N, ns), dtype=torch.float32, device="cuda")
        wl.echoes_gpu(d)
        catalog.append((wl, d))
        print(f"[bench] shape {i}: {N} el, {ns} samples, grid {gx}x{gz}, {wl.cf}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    outs = {}

    def job(k, slot):
        wl, d = catalog[k]
        npix = wl.n_x * wl.n_z
        key = (k, slot % 256)   # distinct outputs for every job that can be queued or in flight
        if key not in outs:
            outs[key] = (torch.empty(npix, device="cuda"), torch.empty(npix, device="cuda"))
        v, f = outs[key]
        return wl, Job2D(config=wl.sta_config(), upsampling=8, coherence=CoherenceFactor(wl.cf, wl.cf_param),
                         grid=wl.grid(), rx_apod=wl.rx_apodization(), echoes=d, value=v, factor=f, storage="f16")

    results = []
    with Engine(max_inflight=args.inflight, max_queued=1024, handles_per_shape=2) as eng:
        # build, tune and warm every shape's handles, then time each job alone
        for k in range(len(catalog)):
            assert eng.wait(eng.submit(job(k, 0)[1]))["ok"]
        assert eng.idle()
        alone_ms, tiles = [], []
        for k in range(len(catalog)):
            ms = []
            for r in range(3):
                s = eng.wait(eng.submit(job(k, r)[1]))
                assert s["ok"]
                ms.append(s["device_ms"])
            alone_ms.append(float(np.median(ms)))
            tiles.append(s["das_tile"])
            print(f"[bench] shape {k} alone: {alone_ms[-1]:.2f} ms, DAS tile {s['das_tile']}", file=sys.stderr,
                  flush=True)
        warm = eng.counters()
        print(f"[bench] warm engine: {warm}", file=sys.stderr, flush=True)
        for rate in [float(r) for r in args.rates.split(",")]:
            n_jobs = max(16, min(args.jobs, int(rate * 8)))
            picks = rng.integers(len(catalog), size=n_jobs)
            gaps = rng.exponential(1.0 / rate, size=n_jobs)
            b2b_ms = float(sum(alone_ms[int(k)] for k in picks))
            capacity = n_jobs / (b2b_ms / 1e3)
            c0 = eng.counters()
            pending, pairs, busy = [], 0, 0
            t0 = time.perf_counter()
            t_next = t0
            for n, (k, g) in enumerate(zip(picks, gaps)):
                t_next += g
                while time.perf_counter() < t_next:
                    time.sleep(min(1e-3, max(0.0, t_next - time.perf_counter())))
                wl, j = job(int(k), n)
                try:
                    pending.append((int(k), wl, j, eng.submit(j)))
                except _lib.BusyError:
                    busy += 1
            lat, ratio = [], []
            for k, wl, j, t in pending:
                s = eng.wait(t)
                assert s["ok"], s
                lat.append(s["latency_ms"])
                ratio.append(s["latency_ms"] / alone_ms[k])
                pairs += wl.pixel_pairs
            wall = time.perf_counter() - t0
            c1 = eng.counters()
            fps = len(pending) / wall
            below = rate < 0.8 * capacity
            r = {"rate_fps": rate, "jobs": len(pending), "refused": busy,
                 "frames_per_s": round(fps, 2), "gpixel_pair_per_s": round(pairs / wall / 1e9, 2),
                 "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
                 "p50_latency_over_device_alone": round(float(np.percentile(ratio, 50)), 2),
                 "back_to_back_frames_per_s": round(capacity, 2),
                 "frac_of_back_to_back": round(fps / capacity, 3),
                 "regime": "below capacity" if below else "saturated",
                 "handles_built_in_window": c1["handles_created"] - c0["handles_created"],
                 "tunes_in_window": c1["tunes"] - c0["tunes"]}
            results.append(r)
            print(f"[bench] rate {rate}: {r}", file=sys.stderr, flush=True)
        final = eng.counters()
    print(json.dumps({"metric": "serving mix: Gpixel*pair/s and p50/p99 latency per arrival rate",
                      "unit": "Gpixel*pair/s", "n_gpus": 1, "inflight": args.inflight, "handles_per_shape": 2,
                      "storage": "f16", "data": "synthetic",
                      "shapes": [{"elements": wl.num_elements, "samples": wl.num_samples, "grid": [wl.n_x, wl.n_z],
                                  "coherence": wl.cf, "device_ms_alone": round(alone_ms[k], 3),
                                  "das_kernel": {-1: "gather", -2: "rx-parallel"}.get(tiles[k], f"tiled (shape {tiles[k]})")}
                                 for k, (wl, _) in enumerate(catalog)],
                      "engine_counters": final, "rates": results}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.mode == "batch":
        return batch_mode(args)
    if args.mode == "serve":
        return serve_mode(args)
    import torch
    from us_lab4a_amd import imaging, workloads
    from us_lab4a_amd.sta import CoherenceFactor, HipSTAProcessor

    wl = workloads.WORKLOADS[args.workload]()
    npix = wl.n_x * wl.n_z
    dev = torch.device("cuda:0")
    print(f"[bench] workload {wl.name}: {wl.num_tx}x{wl.num_elements}x{wl.num_samples}, U={wl.upsampling}, "
          f"grid {wl.n_x}x{wl.n_z}, cf={wl.cf}, storage={args.storage}", file=sys.stderr, flush=True)
    d_echoes = torch.empty((wl.num_tx, wl.num_elements, wl.num_samples), dtype=torch.float32, device=dev)
    wl.echoes_gpu(d_echoes)
    torch.cuda.synchronize()
    proc = HipSTAProcessor(wl.sta_config(), None, wl.upsampling, CoherenceFactor(wl.cf, wl.cf_param),
                           wl.peak_offset, wl.rx_apodization(), tx_apod=wl.tx_apodization(), storage=args.storage)
    proc.set_grid(wl.grid())
    if args.tile:
        proc.set_kernel(2, *args.tile)
    elif args.kernel:
        proc.set_kernel(args.kernel)
    if args.graph:
        proc.set_graph(True)
    d_value = torch.empty(npix, dtype=torch.float32, device=dev)
    d_factor = torch.empty(npix, dtype=torch.float32, device=dev)

    for _ in range(args.warmup):
        proc.process_device(d_echoes, wl.num_samples, d_value, d_factor)
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)

    # throughput: K frames back to back
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        proc.process_device(d_echoes, wl.num_samples, d_value, d_factor)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    ms_per_step = total / args.steps * 1e3
    print(f"[bench] {args.steps} frames in {total:.3f} s", file=sys.stderr, flush=True)

    # latency: frames one at a time, with the HIP-event stage times
    lat, prep_ms, das_ms = [], [], []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        proc.process_device(d_echoes, wl.num_samples, d_value, d_factor)
        p_ms, k_ms = proc.last_timing()   # synchronises on the DAS end event
        lat.append((time.perf_counter() - t1) * 1e3)
        prep_ms.append(p_ms)
        das_ms.append(k_ms)
    lat = np.array(lat)
    if args.graph:
        # replayed frame graphs have no stage split: take it from eager frames
        proc.set_graph(False)
        prep_ms, das_ms = [], []
        for _ in range(max(3, min(args.steps, 10))):
            proc.process_device(d_echoes, wl.num_samples, d_value, d_factor)
            p_ms, k_ms = proc.last_timing()
            prep_ms.append(p_ms)
            das_ms.append(k_ms)
    # host-buffer frames (mista_process): echoes in from pinned host memory,
    # images back; reported beside the device-resident rate, never as value
    h_echoes = torch.empty(d_echoes.shape, dtype=torch.float32, pin_memory=True)
    h_echoes.copy_(d_echoes)
    h_np = h_echoes.numpy()
    proc.process_echoes(h_np)
    host_ms = []
    for _ in range(3):
        t1 = time.perf_counter()
        proc.process_echoes(h_np)
        host_ms.append((time.perf_counter() - t1) * 1e3)
    host_frame_ms = float(np.median(host_ms))
    del h_echoes, h_np
    das_avg = float(np.mean(das_ms))
    prep_avg = float(np.mean(prep_ms))
    alg = algorithmic_bytes(wl, args.storage)
    achieved = alg / (das_avg * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(wl, args.storage)
    value = wl.pixel_pairs / (ms_per_step * 1e-3) / 1e9
    cols = [0, 137, 255, 256, 400, 511] if wl.n_x == 512 else list(np.linspace(0, wl.n_x - 1, 6).astype(int))
    images = {args.storage: d_value.cpu().numpy().reshape(wl.n_x, wl.n_z)}
    ran = proc.last_kernel()   # the configuration the timed frames ran: shape, chunk size, ring depth
    proc.close()
    headline = None
    if not args.no_headline:
        # FP16 storage + frames batched through the engine against this plain frame
        print("[bench] FP16 storage, batched frames", file=sys.stderr, flush=True)
        b16, v16 = fp16_batched(wl, d_echoes, args.steps, args.warmup, args.inflight)
        images["f16_batched"] = v16
        fps_plain = 1e3 / ms_per_step
        headline = {"plain_frames_per_s": round(fps_plain, 3), "plain_storage": args.storage,
                    "fp16_batched_frames_per_s": round(b16["frames_per_s"], 3),
                    "fp16_batched_gpixel_pair_per_s": round(wl.pixel_pairs * b16["frames_per_s"] / 1e9, 3),
                    "speedup": round(b16["frames_per_s"] / fps_plain, 3), "inflight": args.inflight,
                    "fp16_batched_latency_ms": {"p50": round(b16["p50"], 3), "p99": round(b16["p99"], 3)},
                    "target": "fp16 batched >= 1.8x the plain FP32 frames/s at <= -60 dB envelope error vs FP64"}
    hbm = achievable_hbm_gbs()
    print(f"[bench] achievable HBM read {hbm:.0f} GB/s", file=sys.stderr, flush=True)
    out = {
        "metric": "TFM Gpixel*pair/s (end-to-end frame: prep + DAS + PCF + envelope)",
        "value": round(value, 3),
        "unit": "Gpixel*pair/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "f32" else f"{args.storage}-storage/f32-accumulate",
        "data": "synthetic (GPU-generated seeded point + speckle scatterers, AWGN 30 dB)",
        "config": {"workload": wl.name, "elements": wl.num_elements, "tx": wl.num_tx, "samples": wl.num_samples,
                   "upsampling": wl.upsampling, "grid": [wl.n_x, wl.n_z], "coherence": wl.cf,
                   "storage": args.storage, "pixel_pairs_per_frame": wl.pixel_pairs,
                   "das_tile": tile_name(ran),
                   "graph": bool(args.graph)},
        "frame_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                             "p99": round(float(np.percentile(lat, 99)), 4)},
        "stage_ms": {"prep": round(prep_avg, 4), "das": round(das_avg, 4)},
        "host_buffer_frame": {"ms": round(host_frame_ms, 3),
                              "gpixel_pair_per_s": round(wl.pixel_pairs / (host_frame_ms * 1e-3) / 1e9, 3),
                              "note": "mista_process from pinned host echoes, images copied back (PCIe included)"},
        "das_gpixel_pair_per_s": round(wl.pixel_pairs / (das_avg * 1e-3) / 1e9, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     # measured HBM-side bytes per launch over this run's launch time
                     "traffic_gbs": round(traffic / (das_avg * 1e-3) / 1e9, 1) if traffic else None,
                     "achievable_gbs": round(hbm, 1), "frac_of_achievable": round(achieved / hbm, 4),
                     "kernel": "das (delay-and-sum + PCF + envelope)", "algorithmic_bytes": alg,
                     "bounds": instruction_bounds(wl, args.storage, das_avg)},
    }
    if headline:
        out["headline"] = headline
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        print(f"[bench] cpu oracle leg on {threads} threads", file=sys.stderr, flush=True)
        cb, errs = cpu_oracle_leg(wl, threads, args.cpu_reps, d_echoes, cols, images)
        out["cpu_baseline"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in cb.items()}
        out["speedup_vs_cpu"] = round(value / cb["value"], 1)
        out["envelope_error_db_vs_fp64"] = dict(errs, columns=[int(c) for c in cols])
        if headline:
            headline["meets_target"] = bool(headline["speedup"] >= 1.8 and errs.get("f16_batched", 0.0) <= -60.0)
    print(json.dumps(out), flush=True)


def instruction_bounds(wl, storage, das_ms):
    """Issue-rate bounds of the DAS launch from the committed SQ counter
    summary for this workload and storage (instructions per 64
    pixel-pairs): the time the VALU, the scalar unit and the LDS would need
    alone at one wave64 instruction per 4 cycles per SIMD (VALU, SALU) and
    the LDS's cycles (instructions plus bank-conflict cycles) per CU, at
    2.4 GHz over 256 CUs; each against the measured launch."""
    path = os.path.join(ROOT, "profiles", f"sq_counters_{wl.name}_{storage}.md")
    if not os.path.exists(path):
        return None
    per = {}
    for line in open(path):
        if line.startswith("per_wave_pair:"):
            for kv in line.split(":", 1)[1].split():
                k, v = kv.split("=")
                per[k] = float(v)
    if not per:
        return None
    wave_pairs = wl.pixel_pairs / 64.0
    simd_cycles = 256 * 4 * 2.4e9        # SIMD issue cycles per second
    cu_cycles = 256 * 2.4e9
    valu_ms = wave_pairs * per["valu"] * 4 / simd_cycles * 1e3
    salu_ms = wave_pairs * per["salu"] * 4 / simd_cycles * 1e3
    lds_ms = wave_pairs * (per["lds"] * per.get("lds_cycles_per_inst", 4.0) + per["lds_conflict"]) / cu_cycles * 1e3
    return {"valu_ms": round(valu_ms, 3), "salu_ms": round(salu_ms, 3), "lds_ms": round(lds_ms, 3),
            "launch_ms": round(das_ms, 3), "source": os.path.relpath(path, ROOT), "per_64_pixel_pairs": per}


def tile_name(ran):
    """Human-readable DAS configuration (mista_last_kernel: tile shape; chunk
    size, ring buffers and LDS budget)."""
    p0, p1 = ran
    if p0 < 0:
        return "gather kernel" if p0 == -1 else "rx-parallel (per-rx sums, then statistics; no tile fits)"
    shape = {2: "8x32", 3: "32x8", 4: "8x8x4", 5: "16x32", 6: "32x16", 7: "8x8x8"}.get(p0, "16x16")
    tc, nb, kb = p1 & 0xFF, (p1 >> 8) & 0xFF, (p1 >> 16) & 0xFF
    budget = f"{kb} KiB budget" if kb else "smallest fitting budget"
    return f"{shape} tile, {tc} pairs per chunk, {nb} ring buffers ({budget})"


if __name__ == "__main__":
    main()
