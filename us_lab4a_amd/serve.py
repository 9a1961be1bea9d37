"""Serving engine over the mista C ABI (mista_engine_*, include/mista.h).

    with Engine(max_inflight=4) as eng:
        t = eng.submit(job)            # Job2D: configuration, grid, apodization, echoes, outputs
        stats = eng.wait(t)            # JobStats: status, queued/latency/device ms, footprint

Jobs of uneven shapes are admitted as in-flight slots and device memory
allow (arrival order, except that a job that can start may pass one that must
wait), run on resident processor handles (delay tables, window tables, tuned
tile shape and FFT plans kept in HBM per shape; `handles_per_shape` handles
per shape so that one frame's preparation overlaps the previous DAS) and
overlap on the device.  Handles are built on the engine's builder thread.
The images of a job do not depend on the other jobs in flight.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import lib
from .imaging import grid_xz
from .sta import CoherenceFactor, _STORAGE


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


@dataclass
class Job2D:
    """One 2-D frame: STA configuration (an STAConfiguration), upsampling,
    coherence factor, grid (XYZValueFactor array or [n1][n2][2] x, z),
    apodization, echoes [ntx][nrx][ns] float32 and output buffers [n1*n2]
    (numpy on the host or torch tensors on the device)."""
    config: object
    upsampling: int
    coherence: CoherenceFactor
    grid: np.ndarray
    rx_apod: np.ndarray
    echoes: object
    value: object
    factor: object = None
    tx_apod: np.ndarray = None
    storage: str = "f32"
    peak_offset: float = 0.0

    def to_c(self):
        c = _lib.StaConfig()
        cfg = self.config
        c.num_elements = cfg.num_elements
        c.first_tx_elem = cfg.first_tx_elem
        c.last_tx_elem = cfg.last_tx_elem
        c.upsampling_factor = self.upsampling
        c.pitch = cfg.pitch
        c.sampling_frequency = cfg.sampling_frequency
        c.propagation_speed = cfg.propagation_speed
        c.center_frequency = cfg.center_frequency
        c.dead_zone_m = cfg.dead_zone_m
        c.peak_offset = self.peak_offset
        c.cf_kind = self.coherence.kind
        c.cf_param = self.coherence.param
        c.storage = _STORAGE[self.storage]
        c.device = 0
        self._keep = []   # host copies referenced by the C struct
        xz = grid_xz(self.grid) if self.grid.dtype.names else np.ascontiguousarray(self.grid, np.float32)
        rx = np.ascontiguousarray(self.rx_apod, np.float32)
        tx = None if self.tx_apod is None else np.ascontiguousarray(self.tx_apod, np.float32)
        self._keep += [xz, rx, tx]
        j = _lib.Job()
        j.config = c
        j.grid_xz = xz.ctypes.data
        j.n1, j.n2 = xz.shape[0], xz.shape[1]
        j.rx_apod = rx.ctypes.data
        j.tx_apod = tx.ctypes.data if tx is not None else None
        on_dev = hasattr(self.echoes, "data_ptr") and getattr(self.echoes, "is_cuda", False)
        if not on_dev:
            self.echoes = np.ascontiguousarray(self.echoes, np.float32)
        j.echoes = _ptr(self.echoes)
        j.num_samples = self.echoes.shape[-1]
        j.echoes_on_device = 1 if on_dev else 0
        out_dev = hasattr(self.value, "data_ptr") and getattr(self.value, "is_cuda", False)
        j.value = _ptr(self.value)
        j.factor = _ptr(self.factor)
        j.outputs_on_device = 1 if out_dev else 0
        return j


class Engine:
    def __init__(self, max_inflight=4, max_queued=256, hbm_budget=0, device=0, handles_per_shape=0):
        c = _lib.EngineConfig()
        c.device = device
        c.max_inflight = max_inflight
        c.max_queued = max_queued
        c.hbm_budget = hbm_budget
        c.handles_per_shape = handles_per_shape
        h = ctypes.c_void_p()
        rc = lib().mista_engine_create(ctypes.byref(c), ctypes.byref(h))
        if rc != _lib.MISTA_OK:
            raise _lib._ERR.get(rc, _lib.MistaError)(rc, "mista_engine_create failed")
        self._e = h
        self._jobs = {}

    def _err(self, rc):
        msg = lib().mista_engine_last_error(self._e)
        raise _lib._ERR.get(rc, _lib.MistaError)(rc, msg.decode() if msg else "")

    def submit(self, job):
        """Queues a Job2D; raises BusyError when the admission queue is full."""
        cj = job.to_c()
        t = ctypes.c_uint64()
        rc = lib().mista_engine_submit(self._e, ctypes.byref(cj), ctypes.byref(t))
        if rc != _lib.MISTA_OK:
            self._err(rc)
        self._jobs[t.value] = job   # keeps the buffers alive until wait()
        return t.value

    def wait(self, ticket):
        st = _lib.JobStats()
        rc = lib().mista_engine_wait(self._e, ticket, ctypes.byref(st))
        self._jobs.pop(ticket, None)
        return {"status": st.status, "das_tile": st.das_tile, "queued_ms": st.queued_ms, "latency_ms": st.latency_ms,
                "device_ms": st.device_ms, "footprint": st.footprint, "ok": rc == _lib.MISTA_OK}

    def done(self, ticket):
        d = ctypes.c_int()
        rc = lib().mista_engine_query(self._e, ticket, ctypes.byref(d))
        if rc != _lib.MISTA_OK:
            self._err(rc)
        return bool(d.value)

    def stats(self):
        r, i, q = ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32()
        lib().mista_engine_stats(self._e, ctypes.byref(r), ctypes.byref(i), ctypes.byref(q))
        return {"resident_bytes": r.value, "inflight": i.value, "queued": q.value}

    def counters(self):
        """Totals since creation: handles built, autotuner runs, evictions;
        and handles being built now."""
        h, t, e, b = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint32()
        lib().mista_engine_counters(self._e, ctypes.byref(h), ctypes.byref(t), ctypes.byref(e), ctypes.byref(b))
        return {"handles_created": h.value, "tunes": t.value, "evictions": e.value, "building": b.value}

    def idle(self, timeout_s=600.0):
        """Waits until no handle is being built (after a warm-up)."""
        import time
        t0 = time.perf_counter()
        while self.counters()["building"] and time.perf_counter() - t0 < timeout_s:
            time.sleep(0.005)
        return self.counters()["building"] == 0

    def close(self):
        if getattr(self, "_e", None):
            lib().mista_engine_destroy(self._e)
            self._e = None
            self._jobs.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
