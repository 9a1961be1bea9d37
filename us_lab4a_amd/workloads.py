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
sta2d_192_long   192 elements, 8192 samples, U = 8 (FFT size 67 500), PCF,
                 512 x 512 grid: the serving mix's long records.
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


@dataclass
class Workload3D:
    """Sparse 3-D STA on a 2-D matrix array (Vectorial3DSTAProcessor): all rx
    elements active, a seeded subset of tx elements; volume grid laid out as
    n1 = nx * ny columns (y fastest) by n2 = nz rows."""
    name: str
    array_nx: int
    array_ny: int
    num_tx: int
    num_samples: int
    upsampling: int
    pitch: float = 0.3 * MM
    fs: float = 40e6
    c: float = 1540.0
    fc: float = 3e6
    bw: float = 0.6
    snr_db: float = 35.0
    cf: str = "sign_coherence_factor"
    cf_param: float = 1.0
    extent_xy: tuple = (-10 * MM, 10 * MM)
    extent_z: tuple = (10 * MM, 70 * MM)
    n_x: int = 128
    n_y: int = 128
    n_z: int = 256
    points: int = 500
    seed_tx: int = 21
    seed_points: int = 22
    seed_noise: int = 23
    peak_offset: float = 0.0
    dead_zone_m: float = 0.0

    @property
    def num_rx(self):
        return self.array_nx * self.array_ny

    @property
    def pixel_pairs(self):
        return self.n_x * self.n_y * self.n_z * self.num_tx * self.num_rx

    def element_xy(self):
        return imaging.element_positions_2d_array(self.pitch, self.array_nx, 0.0, self.pitch, self.array_ny, 0.0)

    def active_tx(self):
        """Seeded choice of num_tx distinct elements, in ascending order."""
        u = synth.uniform01(self.seed_tx, np.arange(self.num_rx))
        return sorted(int(i) for i in np.argsort(u, kind="stable")[:self.num_tx])

    def sa3d_config(self):
        from .sta import SA3DConfiguration
        pos = self.element_xy()
        return SA3DConfiguration(sampling_frequency=self.fs, propagation_speed=self.c, center_frequency=self.fc,
                                 tx_elem_pos=pos, rx_elem_pos=pos, active_tx_elem=self.active_tx(),
                                 active_rx_elem=list(range(self.num_rx)), dead_zone_m=self.dead_zone_m)

    def coherence(self):
        from .sta import CoherenceFactor
        return CoherenceFactor(self.cf, self.cf_param)

    def rx_apodization(self):
        return np.ones(self.num_rx, np.float32)

    def grid_xyz(self):
        """[nx*ny][nz][3] float32 (x, y, z), spacing as fillSequenceFromStartToEndWithSize."""
        xs = imaging.fill_sequence_start_end_size(self.extent_xy[0], self.extent_xy[1], self.n_x)
        ys = imaging.fill_sequence_start_end_size(self.extent_xy[0], self.extent_xy[1], self.n_y)
        zs = imaging.fill_sequence_start_end_size(self.extent_z[0], self.extent_z[1], self.n_z)
        g = np.empty((self.n_x, self.n_y, self.n_z, 3), np.float32)
        g[..., 0] = np.asarray(xs, np.float32)[:, None, None]
        g[..., 1] = np.asarray(ys, np.float32)[None, :, None]
        g[..., 2] = np.asarray(zs, np.float32)[None, None, :]
        return g.reshape(self.n_x * self.n_y, self.n_z, 3)

    def scatterers(self):
        return synth.point_scatterers(self.seed_points, self.points, self.extent_xy, self.extent_z,
                                      y_range=self.extent_xy)

    def positions_xyz(self):
        pos = self.element_xy()
        xyz = np.zeros((self.num_rx, 3), np.float32)
        xyz[:, :2] = pos
        return xyz[self.active_tx()], xyz

    def echoes_cpu(self):
        tx, rx = self.positions_xyz()
        e = synth.echoes_cpu(tx, rx, self.scatterers(), self.num_samples, self.fs, self.c, self.fc, self.bw)
        if self.snr_db is not None:
            e, _ = synth.add_noise_cpu(e, self.snr_db, self.seed_noise)
        return e

    def echoes_gpu(self, d_out, stream=None):
        tx, rx = self.positions_xyz()
        return synth.echoes_gpu(tx, rx, self.scatterers(), self.num_samples, self.fs, self.c, self.fc, self.bw,
                                d_out, snr_db=self.snr_db, seed=self.seed_noise, stream=stream)



def sta3d_sparse():
    return Workload3D(name="sta3d_sparse", array_nx=32, array_ny=32, num_tx=64, num_samples=4096, upsampling=4)


def sta3d_small(cf="sign_coherence_factor", n=(12, 10, 24)):
    """8 x 8 array, 16 tx, 1024 samples: small enough for the oracle."""
    return Workload3D(name="sta3d_small", array_nx=8, array_ny=8, num_tx=16, num_samples=1024, upsampling=4,
                      cf=cf, extent_xy=(-3 * MM, 3 * MM), extent_z=(5 * MM, 17 * MM), n_x=n[0], n_y=n[1], n_z=n[2],
                      points=6)


def sta2d_serve_long(num_elements=192):
    """The serving mix's long records: 192 elements, 8192 samples, U = 8
    (FFT size 67 500: the four-step preparation), PCF, grid 512 x 512 over the
    aperture's width and z 5..65 mm, points plus speckle (2000 / cm^2)."""
    half = num_elements * 0.3 * MM / 2
    return Workload2D(name=f"sta2d_{num_elements}_long", num_elements=num_elements, num_samples=8192, upsampling=8,
                      cf="phase_coherence_factor", grid_x=(-half, half), grid_z=(5 * MM, 65 * MM), n_x=512, n_z=512,
                      points=50, point_x=(-half, half), point_z=(5 * MM, 65 * MM), speckle_density=2000.0,
                      seed_points=492, seed_speckle=592, seed_noise=692)


WORKLOADS = {"sta2d_32_small": sta2d_32_small, "sta2d_128_pcf": sta2d_128_pcf, "sta3d_sparse": sta3d_sparse,
             "sta2d_192_long": sta2d_serve_long}


def db_error(env, env_ref):
    """20 log10(max|env - env_ref| / max|env_ref|)."""
    num = float(np.max(np.abs(env.astype(np.float64) - env_ref.astype(np.float64))))
    den = float(np.max(np.abs(env_ref)))
    if num == 0.0:
        return -math.inf
    return 20.0 * math.log10(num / den)
