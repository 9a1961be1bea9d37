"""Seeded synthetic STA acquisitions (point scatterers, speckle, AWGN).

The scene (scatterer positions and amplitudes) is drawn on the host from a
counter-based generator (splitmix64 finaliser over seed*C + index), so it is
identical on every machine.  Echoes are built either on the GPU by the
library (mista_synth_echoes_device) or, for small cases, by the float64 numpy
model below; both stamp the same pulse
    p(t) = exp(-a t^2) cos(2 pi fc t),  a = (pi fc bw)^2 / (4 ln 2)
(Gaussian-modulated, fractional bandwidth `bw` at -6 dB, support where the
envelope is above -60 dB) at the round-trip time |e_tx - s| + |e_rx - s| / c.
Noise: white Gaussian, standard deviation rms(echoes) * 10^(-snr/20).
"""
import ctypes
import math

import numpy as np

from . import _lib

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(z):
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed, idx):
    """Uniform (0, 1) doubles for counter indices idx (same as the device generator)."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        u = _mix64(np.uint64(seed) * np.uint64(0xD1B54A32D192ED03) + idx)
    return ((u >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def gaussian(seed, n, offset=0):
    i = np.arange(offset, offset + n, dtype=np.uint64)
    u1 = uniform01(seed, 2 * i)
    u2 = uniform01(seed, 2 * i + 1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def point_scatterers(seed, n, x_range, z_range, amp_range=(0.5, 1.0), y_range=None):
    """[n][4] float32 (x, y, z, amplitude), uniform positions and amplitudes."""
    u = uniform01(seed, np.arange(4 * n)).reshape(n, 4)
    s = np.zeros((n, 4), np.float64)
    s[:, 0] = x_range[0] + (x_range[1] - x_range[0]) * u[:, 0]
    if y_range is not None:
        s[:, 1] = y_range[0] + (y_range[1] - y_range[0]) * u[:, 1]
    s[:, 2] = z_range[0] + (z_range[1] - z_range[0]) * u[:, 2]
    s[:, 3] = amp_range[0] + (amp_range[1] - amp_range[0]) * u[:, 3]
    return s.astype(np.float32)


def speckle_scatterers(seed, density_per_cm2, x_range, z_range, sigma=0.1):
    """Uniform positions at the given areal density, Rayleigh amplitudes
    (magnitude of a complex Gaussian of per-component std `sigma`)."""
    area_cm2 = (x_range[1] - x_range[0]) * (z_range[1] - z_range[0]) * 1e4
    n = int(round(density_per_cm2 * area_cm2))
    u = uniform01(seed, np.arange(2 * n)).reshape(n, 2)
    g = gaussian(seed ^ 0x5DEECE66D, 2 * n).reshape(n, 2)
    s = np.zeros((n, 4), np.float64)
    s[:, 0] = x_range[0] + (x_range[1] - x_range[0]) * u[:, 0]
    s[:, 2] = z_range[0] + (z_range[1] - z_range[0]) * u[:, 1]
    s[:, 3] = sigma * np.hypot(g[:, 0], g[:, 1])
    return s.astype(np.float32)


def linear_array_xyz(num_elements, pitch):
    """Element centres of a linear array (ArrayGeometry.h:37-49), y = z = 0."""
    p = np.zeros((num_elements, 3), np.float32)
    p[:, 0] = (np.arange(num_elements) * pitch - (num_elements - 1) * pitch / 2).astype(np.float32)
    return p


def pulse_a(fc, bw):
    ref = 10.0 ** (-6.0 / 20.0)
    return -(math.pi * fc * bw) ** 2 / (4.0 * math.log(ref))


def echoes_cpu(tx_xyz, rx_xyz, scat, num_samples, fs, c, fc, bw):
    """float64 numpy model of the echo synthesis (small cases only)."""
    a = pulse_a(fc, bw)
    half = math.sqrt(-math.log(1e-3) / a) * fs
    ntx, nrx = len(tx_xyz), len(rx_xyz)
    out = np.zeros((ntx, nrx, num_samples), np.float64)
    S = scat.astype(np.float64)
    dt_ = np.linalg.norm(S[None, :, :3] - tx_xyz[:, None, :].astype(np.float64), axis=-1)  # [ntx][ns]
    dr_ = np.linalg.norm(S[None, :, :3] - rx_xyz[:, None, :].astype(np.float64), axis=-1)  # [nrx][ns]
    w = 2.0 * math.pi * fc / fs
    for t in range(ntx):
        cs = (dt_[t][None, :] + dr_) / c * fs      # [nrx][nscat] centre sample
        for si in range(S.shape[0]):
            amp = S[si, 3]
            for r in range(nrx):
                c0 = cs[r, si]
                n0 = max(int(math.ceil(c0 - half)), 0)
                n1 = min(int(math.floor(c0 + half)), num_samples - 1)
                if n1 < n0:
                    continue
                u = np.arange(n0, n1 + 1) - c0
                out[t, r, n0:n1 + 1] += amp * np.exp(-a / (fs * fs) * u * u) * np.cos(w * u)
    return out.astype(np.float32)


def add_noise_cpu(echoes, snr_db, seed):
    rms = math.sqrt(float(np.mean(echoes.astype(np.float64) ** 2)))
    sigma = rms * 10.0 ** (-snr_db / 20.0)
    n = echoes.size
    noise = gaussian(seed, n).reshape(echoes.shape)
    return (echoes + sigma * noise).astype(np.float32), sigma


def echoes_gpu(tx_xyz, rx_xyz, scat, num_samples, fs, c, fc, bw, d_out, snr_db=None, seed=0, stream=None):
    """Synthesize into the device buffer d_out ([ntx][nrx][ns] float32; a
    torch tensor or a device pointer).  Returns the noise sigma used."""
    L = _lib.lib()
    p = _lib.SynthParams()
    p.num_tx, p.num_rx, p.num_samples = len(tx_xyz), len(rx_xyz), num_samples
    p.num_scatterers = len(scat)
    p.sampling_frequency, p.propagation_speed, p.center_frequency, p.bandwidth = fs, c, fc, bw
    p.noise_std = 0.0
    p.seed = seed
    tx = np.ascontiguousarray(tx_xyz, np.float32)
    rx = np.ascontiguousarray(rx_xyz, np.float32)
    sc = np.ascontiguousarray(scat, np.float32)
    ptr = d_out.data_ptr() if hasattr(d_out, "data_ptr") else int(d_out)
    sptr = stream.cuda_stream if stream is not None and hasattr(stream, "cuda_stream") else stream
    _lib.check(L.mista_synth_echoes_device(ctypes.byref(p), tx.ctypes.data, rx.ctypes.data,
                                           sc.ctypes.data if len(sc) else None, ptr, sptr))
    sigma = 0.0
    if snr_db is not None:
        n = len(tx_xyz) * len(rx_xyz) * num_samples
        ss = ctypes.c_double()
        _lib.check(L.mista_sum_squares_device(ptr, n, ctypes.byref(ss), sptr))
        sigma = math.sqrt(ss.value / n) * 10.0 ** (-snr_db / 20.0)
        _lib.check(L.mista_add_noise_device(ptr, n, sigma, seed, sptr))
    return sigma
