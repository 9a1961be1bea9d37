"""ctypes binding of libmista_hip.so (the C ABI declared in include/mista.h).

There is no CPU fallback: if the HIP library is missing, every entry point
raises.  Build it with ``python -m us_lab4a_amd.build``.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmista_hip.so")

MISTA_OK = 0
MISTA_E_INVALID_PARAM = -1
MISTA_E_INVALID_VALUE = -2
MISTA_E_INVALID_STATE = -3
MISTA_E_HIP = -4
MISTA_E_NO_MEMORY = -5
MISTA_E_BUSY = -6

CF_NONE, CF_PCF, CF_SCF = 0, 1, 2
STORAGE_F32, STORAGE_F16, STORAGE_BF16 = 0, 1, 2


class StaConfig(ctypes.Structure):
    """mista_sta_config (include/mista.h)."""
    _fields_ = [
        ("num_elements", ctypes.c_uint32),
        ("first_tx_elem", ctypes.c_uint32),
        ("last_tx_elem", ctypes.c_uint32),
        ("upsampling_factor", ctypes.c_uint32),
        ("pitch", ctypes.c_float),
        ("sampling_frequency", ctypes.c_float),
        ("propagation_speed", ctypes.c_float),
        ("center_frequency", ctypes.c_float),
        ("dead_zone_m", ctypes.c_float),
        ("peak_offset", ctypes.c_float),
        ("cf_kind", ctypes.c_int32),
        ("cf_param", ctypes.c_float),
        ("storage", ctypes.c_int32),
        ("device", ctypes.c_int32),
    ]


class Sta3dConfig(ctypes.Structure):
    """mista_sta3d_config (include/mista.h)."""
    _fields_ = [
        ("num_tx", ctypes.c_uint32),
        ("num_rx", ctypes.c_uint32),
        ("upsampling_factor", ctypes.c_uint32),
        ("sampling_frequency", ctypes.c_float),
        ("propagation_speed", ctypes.c_float),
        ("center_frequency", ctypes.c_float),
        ("dead_zone_m", ctypes.c_float),
        ("peak_offset", ctypes.c_float),
        ("cf_kind", ctypes.c_int32),
        ("cf_param", ctypes.c_float),
        ("storage", ctypes.c_int32),
        ("device", ctypes.c_int32),
    ]


class EngineConfig(ctypes.Structure):
    """mista_engine_config (include/mista.h)."""
    _fields_ = [
        ("device", ctypes.c_int32),
        ("max_inflight", ctypes.c_uint32),
        ("max_queued", ctypes.c_uint32),
        ("handles_per_shape", ctypes.c_uint32),
        ("hbm_budget", ctypes.c_uint64),
    ]


class Job(ctypes.Structure):
    """mista_job (include/mista.h)."""
    _fields_ = [
        ("config", StaConfig),
        ("grid_xz", ctypes.c_void_p),
        ("n1", ctypes.c_uint32),
        ("n2", ctypes.c_uint32),
        ("rx_apod", ctypes.c_void_p),
        ("tx_apod", ctypes.c_void_p),
        ("echoes", ctypes.c_void_p),
        ("num_samples", ctypes.c_uint32),
        ("echoes_on_device", ctypes.c_int32),
        ("value", ctypes.c_void_p),
        ("factor", ctypes.c_void_p),
        ("outputs_on_device", ctypes.c_int32),
        ("pad_", ctypes.c_uint32),
    ]


class JobStats(ctypes.Structure):
    """mista_job_stats (include/mista.h)."""
    _fields_ = [
        ("status", ctypes.c_int32),
        ("das_tile", ctypes.c_int32),
        ("queued_ms", ctypes.c_double),
        ("latency_ms", ctypes.c_double),
        ("device_ms", ctypes.c_double),
        ("footprint", ctypes.c_uint64),
    ]


class SynthParams(ctypes.Structure):
    """mista_synth_params (include/mista.h)."""
    _fields_ = [
        ("num_tx", ctypes.c_uint32),
        ("num_rx", ctypes.c_uint32),
        ("num_samples", ctypes.c_uint32),
        ("num_scatterers", ctypes.c_uint32),
        ("sampling_frequency", ctypes.c_float),
        ("propagation_speed", ctypes.c_float),
        ("center_frequency", ctypes.c_float),
        ("bandwidth", ctypes.c_float),
        ("noise_std", ctypes.c_float),
        ("pad_", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
    ]


class MistaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mista error {code}: {msg}")
        self.code = code


class InvalidParameterError(MistaError, ValueError):
    """InvalidParameterException analogue."""


class InvalidValueError(MistaError, ValueError):
    """InvalidValueException analogue."""


class InvalidStateError(MistaError):
    """InvalidStateException analogue."""


class BusyError(MistaError):
    """Serving admission refused."""


_ERR = {
    MISTA_E_INVALID_PARAM: InvalidParameterError,
    MISTA_E_INVALID_VALUE: InvalidValueError,
    MISTA_E_INVALID_STATE: InvalidStateError,
    MISTA_E_BUSY: BusyError,
}

# Every symbol declared in include/mista.h with its ctypes signature.
_vp = ctypes.c_void_p
_fp = ctypes.POINTER(ctypes.c_float)
_u32, _u64, _i32, _f32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float
SIGNATURES = {
    "mista_abi_version": (_i32, []),
    "mista_create": (_i32, [ctypes.POINTER(StaConfig), ctypes.POINTER(_vp)]),
    "mista_destroy": (_i32, [_vp]),
    "mista_last_error": (ctypes.c_char_p, [_vp]),
    "mista_prepare": (_i32, [_vp, _u32]),
    "mista_set_rx_apod": (_i32, [_vp, _vp, _u32]),
    "mista_set_tx_apod": (_i32, [_vp, _vp, _u32]),
    "mista_set_grid": (_i32, [_vp, _vp, _u32, _u32]),
    "mista_set_grid_points": (_i32, [_vp, _vp, _u32, _u32]),
    "mista_set_grid_xyz": (_i32, [_vp, _vp, _u32, _u32, _u32]),
    "mista_set_prep": (_i32, [_vp, _i32]),
    "mista_create_rf": (_i32, [ctypes.POINTER(StaConfig), ctypes.POINTER(_vp)]),
    "mista_envelope_device": (_i32, [_vp, _vp, _u32, _u32, _vp]),
    "mista_envelope": (_i32, [_vp, _vp, _u32, _u32]),
    "mista_engine_create": (_i32, [ctypes.POINTER(EngineConfig), ctypes.POINTER(_vp)]),
    "mista_engine_destroy": (_i32, [_vp]),
    "mista_engine_submit": (_i32, [_vp, ctypes.POINTER(Job), ctypes.POINTER(_u64)]),
    "mista_engine_wait": (_i32, [_vp, _u64, ctypes.POINTER(JobStats)]),
    "mista_engine_query": (_i32, [_vp, _u64, ctypes.POINTER(ctypes.c_int)]),
    "mista_engine_stats": (_i32, [_vp, ctypes.POINTER(_u64), ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "mista_engine_counters": (_i32, [_vp, ctypes.POINTER(_u64), ctypes.POINTER(_u64), ctypes.POINTER(_u64),
                                     ctypes.POINTER(_u32)]),
    "mista_engine_last_error": (ctypes.c_char_p, [_vp]),
    "mista_create_3d": (_i32, [ctypes.POINTER(Sta3dConfig), _vp, _vp, ctypes.POINTER(_vp)]),
    "mista_set_elements_3d": (_i32, [_vp, _vp, _vp]),
    "mista_process": (_i32, [_vp, _vp, _u32, _vp, _vp]),
    "mista_process_points": (_i32, [_vp, _vp, _u32, _vp]),
    "mista_process_device": (_i32, [_vp, _vp, _u32, _vp, _vp, _vp]),
    "mista_prepare_signals_device": (_i32, [_vp, _vp, _u32, _vp]),
    "mista_das_device": (_i32, [_vp, _vp, _vp, _vp]),
    "mista_das_device_noalloc": (_i32, [_vp, _vp, _vp, _vp]),
    "mista_das_batch_device": (_i32, [_vp, _u32, _vp, _vp, _vp]),
    "mista_upload_analytic": (_i32, [_vp, _vp, _u32]),
    "mista_download_analytic": (_i32, [_vp, _vp]),
    "mista_download_delays": (_i32, [_vp, _vp]),
    "mista_download_delays_at": (_i32, [_vp, _vp, _u64, _vp]),
    "mista_query": (_i32, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32), ctypes.POINTER(_u64)]),
    "mista_set_kernel": (_i32, [_vp, _i32, _i32, _i32]),
    "mista_tuned_kernel": (_i32, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "mista_last_kernel": (_i32, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "mista_set_graph": (_i32, [_vp, _i32]),
    "mista_graph_stats": (_i32, [_vp, ctypes.POINTER(_u64), ctypes.POINTER(_u64)]),
    "mista_last_timing": (_i32, [_vp, _fp, _fp]),
    "mista_synth_echoes_device": (_i32, [ctypes.POINTER(SynthParams), _vp, _vp, _vp, _vp, _vp]),
    "mista_sum_squares_device": (_i32, [_vp, _u64, ctypes.POINTER(ctypes.c_double), _vp]),
    "mista_add_noise_device": (_i32, [_vp, _u64, _f32, _u64, _vp]),
    "mista_read_bandwidth_device": (_i32, [_vp, _u64, _i32, ctypes.POINTER(ctypes.c_double), _vp]),
}

_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"mista HIP library not found at {LIB_PATH}; build it with "
                "`python -m us_lab4a_amd.build` (there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, handle=None):
    if rc != MISTA_OK:
        msg = lib().mista_last_error(handle)
        msg = msg.decode() if msg else ""
        raise _ERR.get(rc, MistaError)(rc, msg)
    return rc
