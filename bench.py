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
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def algorithmic_bytes(wl, storage):
    """Compulsory HBM bytes of one DAS launch: analytic tensor read once,
    FP32 delay table read once, value+factor written, apodization read."""
    npix = wl.n_x * wl.n_z
    bc = 8 if storage == "f32" else 4
    cf = 0 if wl.cf == "none" else 1
    return (wl.num_tx * wl.num_elements * wl.num_samples * wl.upsampling * bc
            + wl.num_elements * npix * 4 + npix * 4 * (1 + cf) + 4 * wl.num_elements)


def cpu_baseline(wl, threads):
    """Oracle (FP32 restatement of the reference TBB processor) on a bounded
    sample of the same frame: prep of a tx subset and DAS of a column subset,
    scaled to one frame."""
    import oracle
    import torch
    from us_lab4a_amd import imaging
    ntx_s = 64
    ncol_s = 256
    d = torch.empty((wl.num_tx, wl.num_elements, wl.num_samples), dtype=torch.float32, device="cuda")
    wl.echoes_gpu(d)
    echoes = d[:ntx_s].cpu().numpy()
    del d
    cfg = wl.oracle_cfg()
    cfg.last_tx_elem = cfg.first_tx_elem + ntx_s - 1
    t0 = time.perf_counter()
    a = oracle.prepare(cfg, echoes, np.float32, nthreads=threads)
    t_prep = time.perf_counter() - t0
    # DAS sample: the first ntx_s transmits over ncol_s columns spread over x
    cols = np.linspace(0, wl.n_x - 1, ncol_s).astype(int)
    xz = np.ascontiguousarray(imaging.grid_xz(wl.grid())[cols])
    t0 = time.perf_counter()
    oracle.das(cfg, a, xz, wl.rx_apodization(), None, nthreads=threads)
    t_das = time.perf_counter() - t0
    frac_prep = ntx_s / wl.num_tx
    pp_sample = ncol_s * wl.n_z * ntx_s * wl.num_elements
    t_frame = t_prep / frac_prep + t_das * (wl.pixel_pairs / pp_sample)
    return {
        "value": wl.pixel_pairs / t_frame / 1e9,
        "unit": "Gpixel*pair/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle FP32 restatement of VectorialSTAProcessor: prep of {ntx_s}/{wl.num_tx} tx "
                   f"({t_prep:.2f} s) + DAS/PCF of {ncol_s}/{wl.n_x} columns x {ntx_s} tx ({t_das:.2f} s), "
                   f"scaled to one frame = {t_frame:.1f} s"),
        "frame_s": t_frame,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
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
    if args.kernel:
        proc.set_kernel(args.kernel)
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
    das_avg = float(np.mean(das_ms))
    prep_avg = float(np.mean(prep_ms))
    alg = algorithmic_bytes(wl, args.storage)
    achieved = alg / (das_avg * 1e-3) / 1e9
    value = wl.pixel_pairs / (ms_per_step * 1e-3) / 1e9
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
                   "storage": args.storage, "pixel_pairs_per_frame": wl.pixel_pairs},
        "frame_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                             "p99": round(float(np.percentile(lat, 99)), 4)},
        "stage_ms": {"prep": round(prep_avg, 4), "das": round(das_avg, 4)},
        "das_gpixel_pair_per_s": round(wl.pixel_pairs / (das_avg * 1e-3) / 1e9, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "das (delay-and-sum + PCF + envelope)", "algorithmic_bytes": alg},
    }
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        print(f"[bench] cpu baseline on {threads} threads", file=sys.stderr, flush=True)
        cb = cpu_baseline(wl, threads)
        out["cpu_baseline"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in cb.items()}
        out["speedup_vs_cpu"] = round(value / cb["value"], 1)
    proc.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
