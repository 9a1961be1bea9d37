"""STA processor with the reference's interface, running on the mista HIP library.

Mirrors VectorialSTAProcessor / VectorialSTACUDAProcessor
(src/user/processor/VectorialSTAProcessor.h:58-341,
VectorialSTACUDAProcessor.h:40-127): constructor(config, acquisition,
upsampling_factor, coherence_factor, peak_offset, rx_apod), prepare(base) and
process(grid_data) writing value/factor into a Matrix<XYZValueFactor<float>>
equivalent (numpy structured array, imaging.XYZ_VALUE_FACTOR).

All arithmetic runs in libmista_hip.so; this module only moves arguments.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib
from .imaging import grid_xz

_STORAGE = {"f32": _lib.STORAGE_F32, "f16": _lib.STORAGE_F16, "bf16": _lib.STORAGE_BF16}


@dataclass
class STAConfiguration:
    """Fields of STAConfiguration<float> on the path (STAConfiguration.h:118-165)."""
    num_elements: int
    pitch: float
    sampling_frequency: float
    propagation_speed: float
    center_frequency: float
    first_tx_elem: int = 0
    last_tx_elem: int = -1          # -1: num_elements - 1
    dead_zone_m: float = 0.0
    max_frequency: float = 0.0

    def __post_init__(self):
        if self.last_tx_elem < 0:
            self.last_tx_elem = self.num_elements - 1

    @property
    def num_tx(self):
        return self.last_tx_elem - self.first_tx_elem + 1


@dataclass
class CoherenceFactor:
    """AnalyticSignalCoherenceFactorProcessor selection (CoherenceFactor.h:139-152):
    method "none" | "phase_coherence_factor" (param = gamma) |
    "sign_coherence_factor" (param = p, applied to Re of the per-rx sums)."""
    method: str = "none"
    param: float = 1.0

    @property
    def kind(self):
        return {"none": _lib.CF_NONE, "phase_coherence_factor": _lib.CF_PCF,
                "sign_coherence_factor": _lib.CF_SCF}[self.method]

    def enabled(self):
        return self.method != "none"


def _ptr(a):
    """Pointer of a numpy array or a torch tensor (host or device), or an int."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


def _stream_ptr(stream):
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return stream.cuda_stream
    return int(stream)


class HipSTAProcessor:
    """ArrayProcessor<XYZValueFactor<float>> on MI355X (ArrayProcessor.h:27-40)."""

    def __init__(self, config, acquisition, upsampling_factor, coherence_factor, peak_offset, rx_apod,
                 tx_apod=None, storage="f32", device=0):
        self.config = config
        self.acquisition = acquisition
        self.coherence_factor = coherence_factor if coherence_factor is not None else CoherenceFactor()
        c = _lib.StaConfig()
        c.num_elements = config.num_elements
        c.first_tx_elem = config.first_tx_elem
        c.last_tx_elem = config.last_tx_elem
        c.upsampling_factor = upsampling_factor
        c.pitch = config.pitch
        c.sampling_frequency = config.sampling_frequency
        c.propagation_speed = config.propagation_speed
        c.center_frequency = config.center_frequency
        c.dead_zone_m = config.dead_zone_m
        c.peak_offset = peak_offset
        c.cf_kind = self.coherence_factor.kind
        c.cf_param = self.coherence_factor.param
        c.storage = _STORAGE[storage]
        c.device = device
        self._cfg = c
        h = ctypes.c_void_p()
        check(lib().mista_create(ctypes.byref(c), ctypes.byref(h)))
        self._h = h
        self.upsampling_factor = upsampling_factor
        self.storage = storage
        self._grid_shape = None
        self.set_rx_apod(rx_apod)
        if tx_apod is not None:
            self.set_tx_apod(tx_apod)

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().mista_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- configuration ----------------------------------------------------
    def set_rx_apod(self, rx_apod):
        a = np.ascontiguousarray(rx_apod, dtype=np.float32)
        check(lib().mista_set_rx_apod(self._h, a.ctypes.data, a.size), self._h)

    def set_tx_apod(self, tx_apod):
        if tx_apod is None:
            check(lib().mista_set_tx_apod(self._h, None, 0), self._h)
            return
        a = np.ascontiguousarray(tx_apod, dtype=np.float32)
        check(lib().mista_set_tx_apod(self._h, a.ctypes.data, a.size), self._h)

    def set_grid(self, grid):
        """grid: structured XYZValueFactor array [n1][n2] or float32 [n1][n2][2] (x, z)."""
        xz = grid_xz(grid) if grid.dtype.names else np.ascontiguousarray(grid, dtype=np.float32)
        n1, n2 = xz.shape[:2]
        check(lib().mista_set_grid(self._h, xz.ctypes.data, n1, n2), self._h)
        self._grid_shape = (n1, n2)
        self._grid_xz = xz

    def set_prep(self, method):
        """0 automatic, 1 library FFT (hipFFT), 2 fused LDS kernel,
        3 four-step FFTs (long records)."""
        check(lib().mista_set_prep(self._h, method), self._h)

    def set_kernel(self, variant, p0=0, p1=0):
        check(lib().mista_set_kernel(self._h, variant, p0, p1), self._h)

    def set_graph(self, enable=True):
        """Replay steady device frames (process_device) as one hipGraph."""
        check(lib().mista_set_graph(self._h, 1 if enable else 0), self._h)

    def graph_stats(self):
        """(captures, replays) of the frame graph."""
        a, b = ctypes.c_uint64(0), ctypes.c_uint64(0)
        check(lib().mista_graph_stats(self._h, ctypes.byref(a), ctypes.byref(b)), self._h)
        return a.value, b.value

    def last_kernel(self):
        """(tile shape, chunking) of the last DAS launch; shape -1: gather
        kernel, -2: rx-parallel form (no tile fits)."""
        a, b = ctypes.c_int32(0), ctypes.c_int32(0)
        check(lib().mista_last_kernel(self._h, ctypes.byref(a), ctypes.byref(b)), self._h)
        return a.value, b.value

    def tuned_kernel(self):
        """(tile shape, chunking) the automatic variant chose; (-1, 0) when
        nothing was tuned or no tile fits."""
        a, b = ctypes.c_int32(0), ctypes.c_int32(0)
        check(lib().mista_tuned_kernel(self._h, ctypes.byref(a), ctypes.byref(b)), self._h)
        return a.value, b.value

    # -- reference interface ----------------------------------------------
    def prepare(self, base_element):
        """ArrayProcessor::prepare (VectorialSTAProcessor.h:151-156)."""
        self.acquisition.prepare(base_element)
        check(lib().mista_prepare(self._h, base_element), self._h)

    def process(self, grid_data):
        """VectorialSTAProcessor::process: pulls every tx from the acquisition,
        forms the image and writes value/factor into grid_data in place."""
        xz = grid_xz(grid_data)
        if self._grid_shape != xz.shape[:2] or not np.array_equal(xz, self._grid_xz):
            self.set_grid(grid_data)
        cfg = self.config
        first = self.acquisition.execute(cfg.first_tx_elem)
        echoes = np.empty((cfg.num_tx,) + first.shape, dtype=np.float32)
        echoes[0] = first
        for t in range(1, cfg.num_tx):
            echoes[t] = self.acquisition.execute(cfg.first_tx_elem + t)
        value, factor = self.process_echoes(echoes)
        grid_data["value"] = value
        grid_data["factor"] = factor

    def process_echoes(self, echoes):
        """Host frame [ntx][nrx][ns] float32 -> (value, factor) [n1][n2]."""
        echoes = np.ascontiguousarray(echoes, dtype=np.float32)
        if self._grid_shape is None:
            raise _lib.InvalidStateError(_lib.MISTA_E_INVALID_STATE, "The image grid has not been set.")
        n1, n2 = self._grid_shape
        value = np.empty((n1, n2), np.float32)
        factor = np.empty((n1, n2), np.float32)
        check(lib().mista_process(self._h, echoes.ctypes.data, echoes.shape[-1], value.ctypes.data,
                                  factor.ctypes.data), self._h)
        return value, factor

    # -- device-resident path ---------------------------------------------
    def process_device(self, d_echoes, num_samples, d_value, d_factor=None, stream=None):
        check(lib().mista_process_device(self._h, _ptr(d_echoes), num_samples, _ptr(d_value), _ptr(d_factor),
                                         _stream_ptr(stream)), self._h)

    def prepare_signals_device(self, d_echoes, num_samples, stream=None):
        check(lib().mista_prepare_signals_device(self._h, _ptr(d_echoes), num_samples, _stream_ptr(stream)), self._h)

    def das_device(self, d_value, d_factor=None, stream=None):
        check(lib().mista_das_device(self._h, _ptr(d_value), _ptr(d_factor), _stream_ptr(stream)), self._h)

    def das_device_noalloc(self, d_value, d_factor=None, stream=None):
        """das_device without device allocation (window tables in the
        library's static scratch; see mista_das_device_noalloc)."""
        check(lib().mista_das_device_noalloc(self._h, _ptr(d_value), _ptr(d_factor), _stream_ptr(stream)), self._h)

    def upload_analytic(self, analytic):
        """analytic: complex [ntx][nrx][L] (complex64/128) or float32 [..][L][2]."""
        a = np.asarray(analytic)
        if np.iscomplexobj(a):
            a = np.stack([a.real, a.imag], axis=-1)
        a = np.ascontiguousarray(a, dtype=np.float32)
        check(lib().mista_upload_analytic(self._h, a.ctypes.data, a.shape[2]), self._h)

    def download_analytic(self):
        L = self.query()[0]
        out = np.empty((self.config.num_tx, self.config.num_elements, L, 2), np.float32)
        check(lib().mista_download_analytic(self._h, out.ctypes.data), self._h)
        return out

    def download_delays(self):
        n1, n2 = self._grid_shape
        out = np.empty((self.config.num_elements, n1 * n2), np.float32)
        check(lib().mista_download_delays(self._h, out.ctypes.data), self._h)
        return out

    def _delay_rows(self):
        return self.config.num_elements

    def download_delays_at(self, points):
        """Delay-table columns of the given point indices: [rows][len(points)]."""
        idx = np.ascontiguousarray(points, dtype=np.uint64)
        out = np.empty((self._delay_rows(), idx.size), np.float32)
        check(lib().mista_download_delays_at(self._h, idx.ctypes.data, idx.size, out.ctypes.data), self._h)
        return out

    def query(self):
        L, nf, npts = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        check(lib().mista_query(self._h, ctypes.byref(L), ctypes.byref(nf), ctypes.byref(npts)), self._h)
        return L.value, nf.value, npts.value

    def last_timing(self):
        p, d = ctypes.c_float(), ctypes.c_float()
        check(lib().mista_last_timing(self._h, ctypes.byref(p), ctypes.byref(d)), self._h)
        return p.value, d.value


def das_batch_device(processors, d_values, d_factors=None, stream=None):
    """DAS of several prepared processors: compatible tuned frames share one
    launch of the tiled kernel (mista_das_batch_device); each writes the
    images it writes alone."""
    n = len(processors)
    if n == 0 or len(d_values) != n or (d_factors is not None and len(d_factors) != n):
        raise _lib.InvalidParameterError(_lib.MISTA_E_INVALID_PARAM, "one value (and factor) buffer per processor")
    hs = (ctypes.c_void_p * n)(*[p._h for p in processors])
    vs = (ctypes.c_void_p * n)(*[_ptr(v) for v in d_values])
    fs = None if d_factors is None else (ctypes.c_void_p * n)(*[_ptr(f) for f in d_factors])
    check(lib().mista_das_batch_device(hs, n, vs, fs, _stream_ptr(stream)), processors[0]._h)


@dataclass
class SA3DConfiguration:
    """Fields of SA3DConfiguration<float> on the path (SA3DConfiguration.h:37-58):
    element positions (x, y) per array element and the active tx/rx lists,
    relative to the base element."""
    sampling_frequency: float
    propagation_speed: float
    center_frequency: float
    tx_elem_pos: np.ndarray          # [num tx elements][2]
    rx_elem_pos: np.ndarray          # [num rx elements][2]
    active_tx_elem: list
    active_rx_elem: list
    dead_zone_m: float = 0.0

    @property
    def num_tx(self):
        return len(self.active_tx_elem)

    @property
    def num_rx(self):
        return len(self.active_rx_elem)

    # the 2-D processor's tensor shape is [num_tx][num_elements][L]
    @property
    def num_elements(self):
        return self.num_rx

    def positions(self, base_element=0):
        """Positions of the active elements: txElemPos[base + activeTxElem[i]]
        (Vectorial3DSTAProcessor.h:204-211)."""
        tx = np.asarray(self.tx_elem_pos, np.float32)[base_element + np.asarray(self.active_tx_elem)]
        rx = np.asarray(self.rx_elem_pos, np.float32)[base_element + np.asarray(self.active_rx_elem)]
        return np.ascontiguousarray(tx), np.ascontiguousarray(rx)


class HipSTA3DProcessor(HipSTAProcessor):
    """Vectorial3DSTAProcessor on MI355X (Vectorial3DSTAProcessor.h:50-253):
    same constructor, prepare(base) and process(grid_data) on points with
    x, y, z; rx_apod has one entry per active rx element."""

    def __init__(self, config, acquisition, upsampling_factor, coherence_factor, peak_offset, rx_apod,
                 tx_apod=None, storage="f32", device=0):
        self.config = config
        self.acquisition = acquisition
        self.coherence_factor = coherence_factor if coherence_factor is not None else CoherenceFactor()
        c = _lib.Sta3dConfig()
        c.num_tx = config.num_tx
        c.num_rx = config.num_rx
        c.upsampling_factor = upsampling_factor
        c.sampling_frequency = config.sampling_frequency
        c.propagation_speed = config.propagation_speed
        c.center_frequency = config.center_frequency
        c.dead_zone_m = config.dead_zone_m
        c.peak_offset = peak_offset
        c.cf_kind = self.coherence_factor.kind
        c.cf_param = self.coherence_factor.param
        c.storage = _STORAGE[storage]
        c.device = device
        tx, rx = config.positions(0)
        h = ctypes.c_void_p()
        check(lib().mista_create_3d(ctypes.byref(c), tx.ctypes.data, rx.ctypes.data, ctypes.byref(h)))
        self._h = h
        self._base = 0
        self.upsampling_factor = upsampling_factor
        self.storage = storage
        self._grid_shape = None
        self._grid_xyz = None
        self.set_rx_apod(rx_apod)
        if tx_apod is not None:
            self.set_tx_apod(tx_apod)

    def prepare(self, base_element):
        """Vectorial3DSTAProcessor::prepare (:133-138): a new base element moves
        the active elements."""
        if self.acquisition is not None:
            self.acquisition.prepare(base_element)
        if base_element != self._base:
            tx, rx = self.config.positions(base_element)
            check(lib().mista_set_elements_3d(self._h, tx.ctypes.data, rx.ctypes.data), self._h)
            self._base = base_element
        check(lib().mista_prepare(self._h, base_element), self._h)

    def set_grid(self, grid, ny=1):
        """grid: structured XYZValueFactor array [n1][n2] or float32 [n1][n2][3]
        (x, y, z); ny > 1 declares a volume with n1 = nx * ny (y fastest)."""
        if grid.dtype.names:
            xyz = np.ascontiguousarray(np.stack([grid["x"], grid["y"], grid["z"]], axis=-1), dtype=np.float32)
        else:
            xyz = np.ascontiguousarray(grid, dtype=np.float32)
        n1, n2 = xyz.shape[:2]
        check(lib().mista_set_grid_xyz(self._h, xyz.ctypes.data, n1, n2, ny), self._h)
        self._grid_shape = (n1, n2)
        self._grid_xyz = xyz

    def process(self, grid_data):
        xyz = np.ascontiguousarray(np.stack([grid_data["x"], grid_data["y"], grid_data["z"]], axis=-1), np.float32)
        if self._grid_shape != xyz.shape[:2] or not np.array_equal(xyz, self._grid_xyz):
            self.set_grid(xyz)
        cfg = self.config
        echoes = np.stack([self.acquisition.execute(t) for t in cfg.active_tx_elem]).astype(np.float32)
        value, factor = self.process_echoes(echoes)
        grid_data["value"] = value
        grid_data["factor"] = factor

    def _delay_rows(self):
        return self.config.num_rx + self.config.num_tx

    def download_delays(self):
        n1, n2 = self._grid_shape
        out = np.empty((self.config.num_rx + self.config.num_tx, n1 * n2), np.float32)
        check(lib().mista_download_delays(self._h, out.ctypes.data), self._h)
        return out


class HipSTARFProcessor(HipSTAProcessor):
    """DefaultSTAProcessor on MI355X (DefaultSTAProcessor.h:50-226): real
    samples upsampled by 4 with the fixed 81-tap filter, no apodization;
    process() writes value = signed sum / (Ntx*Nrx) and factor = SCF (or 1).
    envelope() is the row envelope STAMethod takes afterwards."""

    def __init__(self, config, acquisition, coherence_factor, peak_offset, device=0):
        self.config = config
        self.acquisition = acquisition
        self.coherence_factor = coherence_factor if coherence_factor is not None else CoherenceFactor()
        c = _lib.StaConfig()
        c.num_elements = config.num_elements
        c.first_tx_elem = config.first_tx_elem
        c.last_tx_elem = config.last_tx_elem
        c.upsampling_factor = 4
        c.pitch = config.pitch
        c.sampling_frequency = config.sampling_frequency
        c.propagation_speed = config.propagation_speed
        c.center_frequency = config.center_frequency
        c.dead_zone_m = config.dead_zone_m
        c.peak_offset = peak_offset
        c.cf_kind = self.coherence_factor.kind
        c.cf_param = self.coherence_factor.param
        c.storage = _lib.STORAGE_F32
        c.device = device
        h = ctypes.c_void_p()
        check(lib().mista_create_rf(ctypes.byref(c), ctypes.byref(h)))
        self._h = h
        self.upsampling_factor = 4
        self.storage = "rf"
        self._grid_shape = None

    def upload_analytic(self, tensor):
        """Real tensor [ntx][nrx][L] (the prepared RF signals)."""
        a = np.ascontiguousarray(tensor, dtype=np.float32)
        check(lib().mista_upload_analytic(self._h, a.ctypes.data, a.shape[2]), self._h)

    def download_analytic(self):
        L = self.query()[0]
        out = np.empty((self.config.num_tx, self.config.num_elements, L), np.float32)
        check(lib().mista_download_analytic(self._h, out.ctypes.data), self._h)
        return out

    def envelope(self, values):
        """Row envelope of a host float32 [n1][n2] image (returns a copy)."""
        v = np.array(values, dtype=np.float32, order="C", copy=True)
        check(lib().mista_envelope(self._h, v.ctypes.data, v.shape[0], v.shape[1]), self._h)
        return v

    def envelope_device(self, d_values, n1, n2, stream=None):
        check(lib().mista_envelope_device(self._h, _ptr(d_values), n1, n2, _stream_ptr(stream)), self._h)
