"""Shared parity checks of the GPU tests against the CPU oracle (test
infrastructure; imported by tests/test_gpu_*.py only).

Two oracles check a GPU image:

- the reference-equivalent FP32 restatement (`oracle.das`), which computes
  its own FP32 element-to-pixel distances like the reference CPU processor
  (VectorialSTAProcessor.h:265-267): the value must be within TOL of the
  image peak everywhere;
- the same restatement fed the GPU's own delay table and tensor, in the
  arithmetic of the reference CUDA kernels (`oracle.das_table`,
  VectorialSTACUDAProcessor.cu:93-103, :163-169): value and coherence-factor
  image (value * factor) within TOL of the peak at every pixel, no pixel
  exempt.  The GPU delays are FP64 distances rounded once to FP32; the CPU's
  are FP32 distances, up to a few ulps apart, which moves a per-rx sum near
  an atan2 cut (PCF) or a zero real part (SCF) across the discontinuity, and
  moves terms on and off at the record end.  Feeding the GPU's table removes
  exactly that difference, so the CF image can be held to the bar without
  exemptions.

The CF image is also compared with the reference-equivalent oracle on its
own FP32 distances, exempting only the pixels whose CF image the oracle
itself moves by more than TOL when it is given the GPU's delays and the
CUDA kernels' arithmetic instead of its own distances and arithmetic (a
difference that owes nothing to the GPU); at most `max_exempt` of the pixels
may be exempt.
"""
import sys

import numpy as np

import oracle

TOL = 1e-4


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def peak_err(a, b, peak=None):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / (np.abs(b).max() if peak is None else peak))


def check_table_oracle(v, f, cfg, a_gpu, d_gpu, rx_apod, tx_apod, tag, peak=None, branch_rel=0.0, tx_row0=None):
    """GPU value/factor (flat or [n1][n2]) against the oracle over the GPU's
    delay table d_gpu [rows][npix] and tensor a_gpu, CUDA-kernel arithmetic.
    branch_rel > 0 (reduced storage only: the GPU interpolates the stored
    halves and folds the channel scale into the weight, the oracle
    interpolates the rescaled FP32 values) excuses a pixel over the bar only
    if one of its per-rx sums lies within branch_rel of a coherence-factor
    discontinuity; at most 1 % of the pixels may need that excuse.  tx_row0: delay-table row of transmit 0 (default
    cfg.first_tx_elem; a 3-D table lists the active rx rows first).  Returns
    the oracle's (value, factor)."""
    v = np.asarray(v, np.float32).reshape(-1)
    f = np.asarray(f, np.float32).reshape(-1)
    cf = cfg.cf_kind
    if cf == oracle.CF_NONE:
        v_t, f_t = oracle.das_table(cfg, a_gpu, d_gpu, rx_apod, tx_apod, form=oracle.FORM_CUDA, tx_row0=tx_row0)
        sums = None
    else:
        v_t, f_t, sums = oracle.das_table(cfg, a_gpu, d_gpu, rx_apod, tx_apod, form=oracle.FORM_CUDA, rxsums=True,
                                          tx_row0=tx_row0)
    pk = float(np.abs(v_t).max()) if peak is None else peak
    ev = peak_err(v, v_t, pk)
    cf_img = np.abs(v.astype(np.float64) * f - v_t.astype(np.float64) * f_t)
    bad = cf_img > TOL * pk
    exempt = np.zeros_like(bad)
    if branch_rel > 0 and sums is not None:
        exempt = oracle.branch_sensitive(sums, cf, rel=branch_rel).reshape(-1)
    say(tag, "table oracle: value err", f"{ev:.2e}", "of peak; CF image max err", f"{cf_img.max() / pk:.2e}",
        "; over the bar", int(bad.sum()), "of", bad.size, "; near a discontinuity", int(exempt.sum()),
        "; excused", int((bad & exempt).sum()))
    assert ev <= TOL
    assert not np.any(bad & ~exempt)
    assert (bad & exempt).mean() <= 0.01
    if cf == oracle.CF_NONE:
        assert np.all(f == 1.0)
    return v_t, f_t


def check_reference_oracle(v, f, cfg, a, xz, d_gpu, rx_apod, tx_apod, tag, max_exempt=0.01, peak=None):
    """GPU value/factor [n1][n2] against the reference-equivalent FP32
    restatement with its own distances (value: everywhere; CF image: except
    where the oracle's own CF image moves by more than TOL when given the
    GPU's delays and the CUDA kernels' arithmetic).  a: the FP32 tensor both
    oracles read (the GPU's own, or the oracle's)."""
    n1, n2 = xz.shape[:2]
    v = np.asarray(v, np.float32).reshape(n1, n2)
    f = np.asarray(f, np.float32).reshape(n1, n2)
    v_o, f_o = oracle.das(cfg, a, xz, rx_apod, tx_apod)
    pk = float(np.abs(v_o).max()) if peak is None else peak
    ev = peak_err(v, v_o, pk)
    say(tag, "reference oracle (FP32 distances): value err", f"{ev:.2e}", "of peak")
    assert ev <= TOL
    if cfg.cf_kind == oracle.CF_NONE:
        return v_o, f_o
    v_g, f_g = oracle.das_table(cfg, a, d_gpu, rx_apod, tx_apod, form=oracle.FORM_CUDA)
    moved = np.abs(v_g.astype(np.float64) * f_g - (v_o.astype(np.float64) * f_o).reshape(-1)) > TOL * pk
    bad = (np.abs(v.astype(np.float64) * f - v_o.astype(np.float64) * f_o) > TOL * pk).reshape(-1)
    say(tag, "reference oracle: CF image over the bar at", int(bad.sum()), "of", bad.size,
        "pixels; moved by the delays and arithmetic form alone", int(moved.sum()), "; unexplained", int((bad & ~moved).sum()))
    assert not np.any(bad & ~moved)
    assert moved.mean() <= max_exempt
    return v_o, f_o


def fp64_delays_2d(wl, xz):
    """The 2-D delay table as the GPU builds it (FP64 distance times
    fs U / c, rounded once to FP32), computed in numpy: for callers that
    cannot download the GPU's own table (a separate process)."""
    x = xz[..., 0].reshape(-1).astype(np.float64)
    z = xz[..., 1].reshape(-1).astype(np.float64)
    p = np.float64(np.float32(wl.pitch))
    n = wl.num_elements
    xe = np.arange(n) * p - (n - 1) * p / 2
    inv = np.float64(np.float32(wl.fs)) * wl.upsampling / np.float64(np.float32(wl.c))
    return (np.sqrt((x[None] - xe[:, None]) ** 2 + z[None] ** 2) * inv).astype(np.float32)
