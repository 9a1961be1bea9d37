"""Named STA/TFM workloads with their seeded synthetic scenes.

sta2d_32_small   32 elements, pitch 0.3 mm, c 1540 m/s, fs 40 MHz, 5 MHz pulse
                 (60 % bw), 32 tx x 32 rx x 2048 samples, 20 point scatterers
                 (x in [-8, 8] mm, z in [10, 50] mm, amplitude U[0.5, 1]),
                 AWGN 30 dB SNR; U = 4, Hann rx/tx apodization, no coherence
                 factor; grid 256 x 256 over x [-10, 10] mm, z [5, 55] mm.
sta2d_128_pcf    128 elements, pitch 0.3 mm, 5 MHz pulse, fs 40 MHz,
                 128 x 128 x 4096 samples from 200 point scatterers plus
                 speckle (5000 scatterers/cm^2 over the image, Rayleigh
                 amplitudes), AWGN 30 dB SNR; U = 8, PCF (gamma = 1), Hann rx
                 apodization; grid 512 x 1024 over x [-19, 19] mm, z [5, 65] mm.
sta3d_sparse     32 x 32 matrix array (1024 rx), pitch 0.3 mm, 64 seeded tx
                 elements, fs 40 MHz, 3 MHz pulse, 64 x 1024 x 4096 samples from
                 500 scatterers in a 20 x 20 x 60 mm volume, 35 dB SNR; U = 4,
                 SCF; volume 128 x 128 x 256.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import imaging, synth

MM = 1e-3


@dataclass
class Workload2D:
    name: str
    num_elements: int
    num_samples: int
    upsampling: int
    pitch: float = 0.3 * MM
    fs: float = 40e6
    c: float = 1540.0
    fc: float = 5e6
    bw: float = 0.6
    snr_db: float = 30.0
    cf: str = "none"
    cf_param: float = 1.0
    rx_apod: str = "hanning"
    tx_apod: str = None
    grid_x: tuple = (-10 * MM, 10 * MM)
    grid_z: tuple = (5 * MM, 55 * MM)
    n_x: int = 256
    n_z: int = 256
    points: int = 20
    point_x: tuple = (-8 * MM, 8 * MM)
    point_z: tuple = (10 * MM, 50 * MM)
    speckle_density: float = 0.0
    seed_points: int = 1
    seed_speckle: int = 3
    seed_noise: int = 11
    first_tx: int = 0
    last_tx: int = -1
    peak_offset: float = 0.0
    dead_zone_m: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def num_tx(self):
        last = self.num_elements - 1 if self.last_tx < 0 else self.last_tx
        return last - self.first_tx + 1

    @property
    def pixel_pairs(self):
        return self.n_x * self.n_z * self.num_tx * self.num_elements

    def sta_config(self):
        from .sta import STAConfiguration
        return STAConfiguration(num_elements=self.num_elements, pitch=self.pitch, sampling_frequency=self.fs,
                                propagation_speed=self.c, center_frequency=self.fc,
                                first_tx_elem=self.first_tx,
                                last_tx_elem=(self.num_elements - 1 if self.last_tx < 0 else self.last_tx),
                                dead_zone_m=self.dead_zone_m)

    def coherence(self):
        from .sta import CoherenceFactor
        return CoherenceFactor(self.cf, self.cf_param)

    def grid(self):
        return imaging.rectangular_grid(self.grid_x[0], self.grid_x[1], self.n_x, self.grid_z[0], self.grid_z[1], self.n_z)

    def rx_apodization(self):
        return imaging.window(self.rx_apod, self.num_elements)

    def tx_apodization(self):
        return None if self.tx_apod is None else imaging.window(self.tx_apod, self.num_tx)

    def scatterers(self):
        pts = synth.point_scatterers(self.seed_points, self.points, self.point_x, self.point_z)
        if self.speckle_density > 0:
            spk = synth.speckle_scatterers(self.seed_speckle, self.speckle_density, self.grid_x, self.grid_z)
            pts = np.concatenate([pts, spk])
        return pts

    def element_xyz(self):
        return synth.linear_array_xyz(self.num_elements, self.pitch)

    def echoes_cpu(self):
        """[ntx][nrx][ns] float32 from the numpy model (small workloads only)."""
        xyz = self.element_xyz()
        tx = xyz[self.first_tx:self.first_tx + self.num_tx]
        e = synth.echoes_cpu(tx, xyz, self.scatterers(), self.num_samples, self.fs, self.c, self.fc, self.bw)
        if self.snr_db is not None:
            e, _ = synth.add_noise_cpu(e, self.snr_db, self.seed_noise)
        return e

    def echoes_gpu(self, d_out, stream=None):
        xyz = self.element_xyz()
        tx = xyz[self.first_tx:self.first_tx + self.num_tx]
        return synth.echoes_gpu(tx, xyz, self.scatterers(), self.num_samples, self.fs, self.c, self.fc, self.bw,
                                d_out, snr_db=self.snr_db, seed=self.seed_noise, stream=stream)

    def oracle_cfg(self):
        import oracle  # test infrastructure
        cf = {"none": oracle.CF_NONE, "phase_coherence_factor": oracle.CF_PCF,
              "sign_coherence_factor": oracle.CF_SCF}[self.cf]
        return oracle.make_cfg(self.num_elements, self.num_samples, self.upsampling, self.pitch, self.fs, self.c,
                               self.fc, first_tx=self.first_tx,
                               last_tx=(self.num_elements - 1 if self.last_tx < 0 else self.last_tx),
                               dead_zone_m=self.dead_zone_m, peak_offset=self.peak_offset, cf_kind=cf,
                               cf_param=self.cf_param)


def sta2d_32_small():
    return Workload2D(name="sta2d_32_small", num_elements=32, num_samples=2048, upsampling=4,
                      rx_apod="hanning", tx_apod="hanning", cf="none")


def sta2d_128_pcf(storage_note=""):
    return Workload2D(name="sta2d_128_pcf", num_elements=128, num_samples=4096, upsampling=8,
                      cf="phase_coherence_factor", cf_param=1.0, rx_apod="hanning",
                      grid_x=(-19 * MM, 19 * MM), grid_z=(5 * MM, 65 * MM), n_x=512, n_z=1024,
                      points=200, point_x=(-19 * MM, 19 * MM), point_z=(5 * MM, 65 * MM),
                      speckle_density=5000.0, seed_points=2, seed_speckle=3, seed_noise=12)


def tiny(num_elements=8, num_samples=256, upsampling=4, cf="none", n_x=16, n_z=16, seed=5):
    """Small case for brute-force checks: few elements, short records."""
    return Workload2D(name="tiny", num_elements=num_elements, num_samples=num_samples, upsampling=upsampling,
                      cf=cf, grid_x=(-2 * MM, 2 * MM), grid_z=(2 * MM, 8 * MM), n_x=n_x, n_z=n_z, points=3,
                      point_x=(-1.5 * MM, 1.5 * MM), point_z=(3 * MM, 7 * MM), seed_points=seed, rx_apod="hanning",
                      snr_db=30.0)


WORKLOADS = {"sta2d_32_small": sta2d_32_small, "sta2d_128_pcf": sta2d_128_pcf}


def db_error(env, env_ref):
    """20 log10(max|env - env_ref| / max|env_ref|)."""
    num = float(np.max(np.abs(env.astype(np.float64) - env_ref.astype(np.float64))))
    den = float(np.max(np.abs(env_ref)))
    if num == 0.0:
        return -math.inf
    return 20.0 * math.log10(num / den)
