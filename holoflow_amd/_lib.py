"""ctypes binding of libhfx.so (the C ABI declared in include/hfx.h).

There is no fallback: if the shared library is missing or fails to load, every entry
point raises.  Device buffers are passed as raw integer addresses (for example
``torch.Tensor.data_ptr()``) and streams as raw hipStream_t handles.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
# HFX_LIBRARY selects another build of the same library (same ABI), e.g. for A/B timing.
LIB_PATH = os.environ.get("HFX_LIBRARY") or os.path.join(_HERE, "libhfx.so")

HFX_OK = 0
HFX_E_INVALID = 1
HFX_E_UNSUPPORTED = 2
HFX_E_HIP = 3
HFX_E_NOMEM = 4
HFX_NOT_READY = 5
HFX_EOF = 6

PROP_FRESNEL = 0
PROP_ANGULAR_SPECTRUM = 1

PREC_FP32 = 0
PREC_FP16 = 1
PREC_BF16 = 2
PREC_FP16_SPLIT = 3
PRECISIONS = {"fp32": PREC_FP32, "fp16": PREC_FP16, "bf16": PREC_BF16, "fp16_split": PREC_FP16_SPLIT}

FLAG_FFTSHIFT = 1
FLAG_POWER2 = 2
FLAG_BOTH_SIDEBANDS = 4

DTYPE_U16 = 1
DTYPE_F32 = 2
DTYPE_CF32 = 3
DTYPE_U8 = 4

REDUCE_POWER2 = 1
REDUCE_ABS = 2


class HfxError(RuntimeError):
    """Raised for HIP / unsupported failures reported by libhfx."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"hfx error {code}: {msg}")
        self.code = code


class HfxInvalidArgument(HfxError, ValueError):
    """HFX_E_INVALID: the reference's std::invalid_argument."""


class PlanDesc(C.Structure):
    _fields_ = [
        ("height", C.c_int32), ("width", C.c_int32),
        ("window", C.c_int32), ("stride", C.c_int32),
        ("fs_hz", C.c_double),
        ("lambda_m", C.c_double), ("dx_m", C.c_double), ("dy_m", C.c_double), ("z_m", C.c_double),
        ("bin_begin", C.c_int32), ("bin_end", C.c_int32),
        ("prop", C.c_int32), ("precision", C.c_int32), ("flags", C.c_uint32),
        ("pad_height", C.c_int32), ("pad_width", C.c_int32),
        ("filter_enabled", C.c_int32),
        ("r_inner", C.c_double), ("r_outer", C.c_double), ("s_inner", C.c_double), ("s_outer", C.c_double),
        ("split_bin", C.c_int32),
        ("weight_fs_hz", C.c_double),
        ("flatfield_enabled", C.c_int32),
        ("flatfield_sigma_y", C.c_float), ("flatfield_sigma_x", C.c_float),
        ("batch_windows", C.c_int32),
        ("batch_lanes", C.c_int32),
    ]


class PlanInfo(C.Structure):
    _fields_ = [
        ("out_height", C.c_int32), ("out_width", C.c_int32),
        ("n_outputs", C.c_int32), ("n_planes", C.c_int32),
        ("workspace_bytes", C.c_int64),
        ("batch_lanes", C.c_int32), ("groupable", C.c_int32),
    ]


class AsmDesc(C.Structure):
    _fields_ = [
        ("batch", C.c_int64),
        ("height", C.c_int32), ("width", C.c_int32),
        ("lambda_m", C.c_double), ("dx_m", C.c_double), ("dy_m", C.c_double), ("z_m", C.c_double),
        ("in_real", C.c_int32),
        ("pad_height", C.c_int32), ("pad_width", C.c_int32),
        ("filter_enabled", C.c_int32),
        ("r_inner", C.c_double), ("r_outer", C.c_double), ("s_inner", C.c_double), ("s_outer", C.c_double),
    ]


class SynthDesc(C.Structure):
    _fields_ = [
        ("height", C.c_int32), ("width", C.c_int32),
        ("disk_radius_px", C.c_double),
        ("fs_hz", C.c_double),
        ("ref_amp", C.c_double), ("obj_sigma", C.c_double), ("gain", C.c_double),
        ("carrier_cyc_per_px", C.c_double), ("phase_noise_rad", C.c_double), ("fd_max_hz", C.c_double),
        ("seed", C.c_uint32), ("stream_id", C.c_uint32),
    ]


_VP = C.c_void_p
_P = C.POINTER
# name -> (restype, argtypes); every symbol declared in include/hfx.h and include/hfx_nodes.h
SIGNATURES = {
    "hfx_last_error": (C.c_char_p, []),
    "hfx_abi_version": (C.c_int32, []),
    "hfx_plan_desc_init": (None, [_P(PlanDesc)]),
    "hfx_plan_create": (C.c_int, [_P(PlanDesc), _VP, _P(_VP)]),
    "hfx_plan_destroy": (C.c_int, [_VP]),
    "hfx_plan_get_info": (C.c_int, [_VP, _P(PlanInfo)]),
    "hfx_plan_update_z": (C.c_int, [_VP, C.c_double, _VP]),
    "hfx_moments": (C.c_int, [_VP, _VP, C.c_int64, C.c_int64, _VP, _VP]),
    "hfx_moments_batch": (C.c_int, [_VP, _VP, C.c_int64, C.c_int64, C.c_int32, C.c_int64, _VP, _VP]),
    "hfx_moments_group": (C.c_int, [_P(_VP), _P(_VP), _P(C.c_int64), _P(C.c_int64), _P(_VP), C.c_int32, _VP]),
    "hfx_plan_time_dft": (C.c_int, [_VP, _VP, C.c_int64, C.c_int64, _VP, _VP]),
    "hfx_plan_run_stage": (C.c_int, [_VP, C.c_int32, _VP, C.c_int64, C.c_int64, _VP, _VP]),
    "hfx_dft_axis": (C.c_int, [_VP, C.c_int32, C.c_int64, C.c_int32, C.c_int64, C.c_int32, C.c_int32,
                               C.c_float, _VP, _VP]),
    "hfx_fresnel_create": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_int32, C.c_int32, C.c_int32, _VP, _P(_VP)]),
    "hfx_fresnel_update_z": (C.c_int, [_VP, C.c_double, _VP]),
    "hfx_fresnel_execute": (C.c_int, [_VP, _VP, _VP, _VP]),
    "hfx_fresnel_destroy": (C.c_int, [_VP]),
    "hfx_asm_create": (C.c_int, [_P(AsmDesc), _VP, _P(_VP)]),
    "hfx_asm_execute": (C.c_int, [_VP, _VP, _VP, _VP]),
    "hfx_asm_destroy": (C.c_int, [_VP]),
    "hfx_filter2d_create": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                      C.c_int32, _VP, _P(_VP)]),
    "hfx_filter2d_execute": (C.c_int, [_VP, _VP, _VP, _VP]),
    "hfx_filter2d_destroy": (C.c_int, [_VP]),
    "hfx_flatfield_create": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_float, _VP, _P(_VP),
                                       _P(C.c_int32)]),
    "hfx_flatfield_execute": (C.c_int, [_VP, _VP, _VP, _VP]),
    "hfx_flatfield_destroy": (C.c_int, [_VP]),
    "hfx_weighted_mean": (C.c_int, [_VP, C.c_int32, C.c_int64, C.c_int64, _VP, C.c_uint32, _VP, _VP]),
    "hfx_strided_copy": (C.c_int, [_VP, _P(C.c_int64), _P(C.c_int64), C.c_int64, C.c_int32, C.c_int32, _VP,
                                   C.c_int64, C.c_int64, _VP]),
    "hfx_multiply": (C.c_int, [_VP, C.c_int32, _P(C.c_int64), _VP, C.c_int32, _P(C.c_int64), C.c_int32,
                               _P(C.c_int64), _VP, _VP]),
    "hfx_fftshift": (C.c_int, [_VP, _VP, C.c_int32, C.c_int32, _P(C.c_int64), _P(C.c_int32), _VP]),
    "hfx_convert_u16": (C.c_int, [_VP, _VP, C.c_int64, C.c_int32, _VP]),
    "hfx_convert_scratch_bytes": (C.c_int64, [C.c_int64]),
    "hfx_convert": (C.c_int, [_VP, C.c_int32, _VP, C.c_int32, C.c_int32, C.c_int64, _VP, _VP]),
    "hfx_convert_noalloc_scratch_bytes": (C.c_int64, []),
    "hfx_convert_noalloc": (C.c_int, [_VP, C.c_int32, _VP, C.c_int32, C.c_int32, C.c_int64, _VP]),
    "hfx_hbm_probe": (C.c_int, [C.c_int32, _VP, _VP, C.c_int64, C.c_int32, C.c_int32, C.c_int32, _VP]),
    "hfx_slide_avg_step": (C.c_int, [_VP, _VP, _VP, C.c_int64, C.c_int32, _VP]),
    "hfx_synth_desc_init": (None, [_P(SynthDesc), C.c_int32, C.c_int32]),
    "hfx_synth_frames": (C.c_int, [_P(SynthDesc), _VP, C.c_int64, C.c_int64, _VP, _VP]),
    # include/hfx_nodes.h
    "hfx_node_kinds": (C.c_int, [C.c_char_p, C.c_int64]),
    "hfx_node_infer": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int64]),
    "hfx_node_create": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, _VP, _P(_VP)]),
    "hfx_node_update": (C.c_int, [_VP, C.c_char_p, C.c_char_p, _VP, _P(C.c_int32)]),
    "hfx_node_execute": (C.c_int, [_VP, _P(_VP), C.c_int32, _P(_VP), C.c_int32]),
    "hfx_node_acquire_input": (C.c_int, [_VP, _P(_VP)]),
    "hfx_node_push": (C.c_int, [_VP, _P(_VP), C.c_int32]),
    "hfx_node_pop": (C.c_int, [_VP, _P(_VP), C.c_int32]),
    "hfx_node_release_output": (C.c_int, [_VP, C.c_int32]),
    "hfx_node_destroy": (C.c_int, [_VP]),
}



class ServerDesc(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("hbm_reserve_bytes", C.c_int64),
                ("max_streams", C.c_int32), ("n_lanes", C.c_int32), ("use_graphs", C.c_int32),
                ("policy", C.c_int32), ("lane_depth", C.c_int32)]


SCHED_FIFO, SCHED_EDF, SCHED_GROUP = 0, 1, 2


class StreamDesc(C.Structure):
    _fields_ = [("plan", PlanDesc), ("queue_windows", C.c_int32), ("camera", SynthDesc)]


class WindowDone(C.Structure):
    _fields_ = [("stream_id", C.c_int32), ("n_outputs", C.c_int32), ("window", C.c_int64),
                ("latency_us", C.c_float), ("gpu_us", C.c_float), ("wait_us", C.c_float)]


class ServerInfo(C.Structure):
    _fields_ = [("active_streams", C.c_int32), ("max_streams", C.c_int32), ("bytes_held", C.c_int64),
                ("bytes_free", C.c_int64), ("windows_launched", C.c_int64), ("windows_retired", C.c_int64),
                ("clock_error_us", C.c_double), ("group_launches", C.c_int64), ("grouped_windows", C.c_int64)]


class WorkloadDesc(C.Structure):
    _fields_ = [("server", ServerDesc), ("n_streams", C.c_int32), ("streams", C.POINTER(StreamDesc)),
                ("arrival_s", C.POINTER(C.c_double)), ("window_rate_hz", C.POINTER(C.c_double)),
                ("windows_per_stream", C.c_int32), ("seed", C.c_uint64), ("capture_stream", C.c_int32),
                ("capture_out", C.POINTER(C.c_float)), ("time_limit_s", C.c_double), ("arrivals", C.c_int32)]


ARRIVALS_PERIODIC, ARRIVALS_POISSON = 0, 1


class WorkloadResult(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("windows_retired", C.c_int64), ("max_concurrent", C.c_int32),
                ("admissions_deferred", C.c_int32), ("ingest_deferred", C.c_int64),
                ("latency_us", C.POINTER(C.c_float)), ("gpu_us", C.POINTER(C.c_float)),
                ("admit_s", C.POINTER(C.c_double)), ("finish_s", C.POINTER(C.c_double)),
                ("wait_us", C.POINTER(C.c_float)), ("windows_ingested_late", C.c_int64),
                ("clock_error_us", C.c_double), ("group_launches", C.c_int64), ("grouped_windows", C.c_int64)]


SIGNATURES.update({
    # include/hfx_serve.h
    "hfx_server_desc_init": (None, [C.POINTER(ServerDesc)]),
    "hfx_stream_desc_init": (None, [C.POINTER(StreamDesc), C.c_int32, C.c_int32]),
    "hfx_server_stream_bytes": (C.c_int, [C.POINTER(StreamDesc), C.POINTER(C.c_int64)]),
    "hfx_server_create": (C.c_int, [C.POINTER(ServerDesc), C.POINTER(C.c_void_p)]),
    "hfx_server_destroy": (C.c_int, [C.c_void_p]),
    "hfx_server_admit": (C.c_int, [C.c_void_p, C.POINTER(StreamDesc), C.POINTER(C.c_int32)]),
    "hfx_server_clock": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "hfx_server_stream_state": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    "hfx_server_ingest": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64]),
    "hfx_server_ingest_due": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_double]),
    "hfx_server_schedule": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "hfx_server_poll": (C.c_int, [C.c_void_p, C.POINTER(WindowDone), C.c_int32, C.POINTER(C.c_int32)]),
    "hfx_server_window_output": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_void_p)]),
    "hfx_server_release": (C.c_int, [C.c_void_p, C.c_int32]),
    "hfx_server_info_get": (C.c_int, [C.c_void_p, C.POINTER(ServerInfo)]),
    "hfx_serve_workload": (C.c_int, [C.POINTER(WorkloadDesc), C.POINTER(WorkloadResult)]),
})


class HolofileInfo(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("bits_per_pixel", C.c_uint16),
                ("version", C.c_uint16), ("frame_count", C.c_uint32), ("data_size", C.c_uint64),
                ("frame_bytes", C.c_uint64), ("has_footer", C.c_int32)]


SIGNATURES.update({
    # include/hfx_graph.h
    "hfx_graph_compile": (C.c_int, [C.c_char_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "hfx_graph_destroy": (C.c_int, [C.c_void_p]),
    "hfx_graph_describe": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "hfx_graph_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "hfx_graph_start": (C.c_int, [C.c_void_p]),
    "hfx_graph_stop": (C.c_int, [C.c_void_p]),
    "hfx_graph_metrics": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "hfx_graph_output": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.POINTER(C.c_void_p), C.c_char_p,
                                   C.c_int64]),
    "hfx_graph_dot": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int64]),
})


SIGNATURES.update({
    # include/hfx_holofile.h
    "hfx_holofile_open": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "hfx_holofile_close": (C.c_int, [C.c_void_p]),
    "hfx_holofile_get_info": (C.c_int, [C.c_void_p, C.POINTER(HolofileInfo)]),
    "hfx_holofile_footer": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "hfx_holofile_seek": (C.c_int, [C.c_void_p, C.c_int64]),
    "hfx_holofile_tell": (C.c_int64, [C.c_void_p]),
    "hfx_holofile_read": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "hfx_holofile_read_to_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "hfx_holofile_write": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint16, C.c_void_p, C.c_int64,
                                     C.c_char_p]),
})

_lib = None
_lock = threading.Lock()


def lib() -> C.CDLL:
    """Loads libhfx.so once; raises if it is missing (there is no fallback path)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with `python holoflow_amd/build.py` "
                    "(hipcc, gfx950). The HIP path has no CPU fallback.")
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(code: int) -> None:
    if code == HFX_OK:
        return
    msg = lib().hfx_last_error().decode(errors="replace")
    if code == HFX_E_INVALID:
        raise HfxInvalidArgument(code, msg)
    raise HfxError(code, msg)
