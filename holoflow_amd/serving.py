"""Python side of the serving loop (include/hfx_serve.h): stream descriptions for the
mixed-stream load, per-window cost calibration and the native Poisson workload driver.

The loop itself (admission by free HBM, device-resident rings, hipGraph replay per
window position, lanes that overlap the windows of different streams, latency events)
runs in C++ inside libhfx; this module only prepares arguments and summarises results.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, lib

# Frame sizes (H, W) of the mixed load: 512^2, 768^2, 1024^2 and 2048 wide x 512 high.
SIZES = [(512, 512), (768, 768), (1024, 1024), (512, 2048)]
WINDOWS = [32, 64, 128]


@dataclass
class StreamSpec:
    height: int
    width: int
    window: int
    stride: int
    bin_begin: int
    bin_end: int
    z_m: float
    fs_hz: float
    seed: int
    stream_id: int
    prop: str = "asm"
    queue_windows: int = 4

    @property
    def window_period_s(self) -> float:
        """Time the camera takes to deliver one stride of frames."""
        return self.stride / self.fs_hz

    def to_desc(self) -> _lib.StreamDesc:
        d = _lib.StreamDesc()
        lib().hfx_stream_desc_init(C.byref(d), self.height, self.width)
        p = d.plan
        p.window, p.stride = self.window, self.stride
        p.bin_begin, p.bin_end = self.bin_begin, self.bin_end
        p.z_m = self.z_m
        p.fs_hz = self.fs_hz
        p.weight_fs_hz = self.fs_hz
        p.prop = _lib.PROP_ANGULAR_SPECTRUM if self.prop == "asm" else _lib.PROP_FRESNEL
        d.queue_windows = self.queue_windows
        cam = d.camera
        cam.disk_radius_px = 200.0 / 512.0 * min(self.height, self.width)
        cam.fs_hz = self.fs_hz
        cam.seed = self.seed
        cam.stream_id = self.stream_id
        return d


def band_bins(f_lo, f_hi, fs, T):
    """[ceil(f_lo/df), floor(f_hi/df) + 1) clipped to the one-sided DFT (df = fs/T)."""
    df = fs / T
    kb = int(math.ceil(f_lo / df - 1e-12))
    ke = int(math.floor(f_hi / df + 1e-12)) + 1
    return kb, min(ke, T // 2 + 1)


def mixed_streams(n: int, seed: int = 0, sizes=None, windows=None) -> List[StreamSpec]:
    """n streams of the mixed load: frame size from SIZES, T from {32, 64, 128}, S = T/4,
    z ~ U(0.2, 0.6) m, f_s ~ U(20, 75) kHz, band 6-15 kHz clipped to f_s/2."""
    rng = np.random.default_rng(seed)
    sizes = sizes or SIZES
    windows = windows or WINDOWS
    out = []
    for i in range(n):
        H, W = sizes[int(rng.integers(len(sizes)))]
        T = int(windows[int(rng.integers(len(windows)))])
        fs = float(rng.uniform(20e3, 75e3))
        kb, ke = band_bins(6e3, 15e3, fs, T)
        out.append(StreamSpec(H, W, T, T // 4, kb, ke, float(rng.uniform(0.2, 0.6)), fs,
                              seed=1000 + seed * 97 + i, stream_id=i))
    return out


def stream_bytes(spec: StreamSpec) -> int:
    b = C.c_int64()
    d = spec.to_desc()
    check(lib().hfx_server_stream_bytes(C.byref(d), C.byref(b)))
    return b.value


def window_cost_us(spec: StreamSpec, reps: int = 20) -> float:
    """GPU time of one window of this stream alone (HIP events around hfx_moments)."""
    import torch

    from . import hfx
    cfg = hfx.MomentConfig(spec.height, spec.width, spec.window, spec.stride, spec.bin_begin, spec.bin_end,
                           spec.prop, z_m=spec.z_m, fs_hz=spec.fs_hz, weight_fs_hz=spec.fs_hz)
    plan = hfx.MomentPlan(cfg)
    ring = torch.zeros((spec.window + spec.stride * 3, spec.height, spec.width), dtype=torch.int16,
                       device="cuda")
    out = plan.new_output()
    for w in range(3):
        plan.moments(ring, w * spec.stride, out)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for i in range(reps):
        plan.moments(ring, (i % 4) * spec.stride, out)
    e1.record()
    torch.cuda.synchronize()
    cost = e0.elapsed_time(e1) / reps * 1e3
    plan.close()
    return cost


def server_desc(max_streams=16, n_lanes=4, budget_bytes=0, reserve_bytes=8 << 30, use_graphs=True,
                policy="fifo", lane_depth=1):
    d = _lib.ServerDesc()
    lib().hfx_server_desc_init(C.byref(d))
    d.max_streams, d.n_lanes = max_streams, n_lanes
    d.hbm_budget_bytes, d.hbm_reserve_bytes = budget_bytes, reserve_bytes
    d.use_graphs = int(use_graphs)
    d.policy = {"fifo": _lib.SCHED_FIFO, "edf": _lib.SCHED_EDF, "group": _lib.SCHED_GROUP}[policy]
    d.lane_depth = lane_depth
    return d


@dataclass
class WorkloadResult:
    wall_s: float
    windows_retired: int
    max_concurrent: int
    admissions_deferred: int
    ingest_deferred: int
    latency_us: np.ndarray      # [n_streams, windows_per_stream]: due time -> moments complete
    gpu_us: np.ndarray          # launch -> complete on its lane
    wait_us: np.ndarray         # due time -> start of its work
    admit_s: np.ndarray
    finish_s: np.ndarray
    capture: Optional[np.ndarray]
    windows_ingested_late: int = 0
    clock_error_us: float = 0.0
    group_launches: int = 0      # launches that carried windows of several streams (policy "group")
    grouped_windows: int = 0


def run_workload(specs: Sequence[StreamSpec], arrival_s: Sequence[float], window_rate_hz: Sequence[float],
                 windows_per_stream: int, server=None, seed: int = 0, capture_stream: int = -1,
                 time_limit_s: float = 120.0, arrivals: str = "poisson") -> WorkloadResult:
    """Runs the native workload driver: stream i arrives at arrival_s[i] and its camera delivers
    windows at window_rate_hz[i], periodically ("periodic": one stride every 1/rate) or as a
    Poisson process ("poisson")."""
    n = len(specs)
    descs = (_lib.StreamDesc * n)(*[s.to_desc() for s in specs])
    arr = (C.c_double * n)(*[float(a) for a in arrival_s])
    rates = (C.c_double * n)(*[float(r) for r in window_rate_hz])
    lat = np.zeros((n, windows_per_stream), np.float32)
    gpu = np.zeros((n, windows_per_stream), np.float32)
    wait = np.zeros((n, windows_per_stream), np.float32)
    adm = np.zeros(n, np.float64)
    fin = np.zeros(n, np.float64)
    cap = None
    d = _lib.WorkloadDesc()
    d.server = server if server is not None else server_desc()
    d.n_streams = n
    d.streams = descs
    d.arrival_s = arr
    d.window_rate_hz = rates
    d.windows_per_stream = windows_per_stream
    d.seed = seed
    d.capture_stream = capture_stream
    d.time_limit_s = time_limit_s
    d.arrivals = {"periodic": _lib.ARRIVALS_PERIODIC, "poisson": _lib.ARRIVALS_POISSON}[arrivals]
    if capture_stream >= 0:
        s = specs[capture_stream]
        n_out = 3
        cap = np.zeros((windows_per_stream, n_out, s.height, s.width), np.float32)
        d.capture_out = cap.ctypes.data_as(C.POINTER(C.c_float))
    r = _lib.WorkloadResult()
    r.latency_us = lat.ctypes.data_as(C.POINTER(C.c_float))
    r.gpu_us = gpu.ctypes.data_as(C.POINTER(C.c_float))
    r.admit_s = adm.ctypes.data_as(C.POINTER(C.c_double))
    r.finish_s = fin.ctypes.data_as(C.POINTER(C.c_double))
    r.wait_us = wait.ctypes.data_as(C.POINTER(C.c_float))
    check(lib().hfx_serve_workload(C.byref(d), C.byref(r)))
    return WorkloadResult(r.wall_s, r.windows_retired, r.max_concurrent, r.admissions_deferred,
                          r.ingest_deferred, lat, gpu, wait, adm, fin, cap, r.windows_ingested_late,
                          r.clock_error_us, r.group_launches, r.grouped_windows)


def summarize(specs: Sequence[StreamSpec], res: WorkloadResult, periods_s=None):
    """Throughput and latency percentiles of a workload run; `periods_s` are the streams' window
    periods at the rates they ran at (default: their cameras' own stride / f_s)."""
    lat = res.latency_us
    if periods_s is None:
        periods_s = [s.window_period_s for s in specs]
    periods_us = np.asarray(periods_s, np.float64)[:, None] * 1e6
    ratio = (lat / periods_us).ravel()
    pix = np.array([s.height * s.width for s in specs], np.float64)
    out = {
        "streams": len(specs),
        "windows": int(res.windows_retired),
        "wall_s": round(res.wall_s, 4),
        "windows_per_s": round(res.windows_retired / res.wall_s, 1) if res.wall_s > 0 else None,
        "moment_megapixels_per_s": round(float((pix * lat.shape[1]).sum()) / res.wall_s / 1e6, 1)
        if res.wall_s > 0 else None,
        "latency_us_p50": round(float(np.percentile(lat, 50)), 1),
        "latency_us_p99": round(float(np.percentile(lat, 99)), 1),
        "gpu_us_p50": round(float(np.percentile(res.gpu_us, 50)), 1),
        "wait_us_p50": round(float(np.percentile(res.wait_us, 50)), 1),
        "wait_us_p99": round(float(np.percentile(res.wait_us, 99)), 1),
        "latency_over_window_period_p50": round(float(np.percentile(ratio, 50)), 4),
        "latency_over_window_period_p99": round(float(np.percentile(ratio, 99)), 4),
        "frac_windows_within_period": round(float((ratio < 1.0).mean()), 4),
        "max_concurrent_streams": int(res.max_concurrent),
        "admissions_deferred": int(res.admissions_deferred),
        "ingest_deferred": int(res.ingest_deferred),
        "windows_ingested_late": int(res.windows_ingested_late),
        "clock_error_us": round(float(res.clock_error_us), 1),
        "group_launches": int(res.group_launches),
        "grouped_windows": int(res.grouped_windows),
    }
    return out


class Server:
    """Direct use of the server API (admit / ingest / schedule / poll / release)."""

    def __init__(self, desc=None):
        self._h = C.c_void_p()
        d = desc if desc is not None else server_desc()
        check(lib().hfx_server_create(C.byref(d), C.byref(self._h)))

    def admit(self, spec: StreamSpec, wait: bool = True, timeout_s: float = 60.0) -> Optional[int]:
        """Stream id, or None when no slot or memory is free now.  The server sets the stream
        up on its worker thread; with wait, returns once it is ready to take frames."""
        import time
        sid = C.c_int32(-1)
        d = spec.to_desc()
        rc = lib().hfx_server_admit(self._h, C.byref(d), C.byref(sid))
        if rc == _lib.HFX_NOT_READY:
            return None
        check(rc)
        t0 = time.monotonic()
        while wait and self.state(sid.value) == 0:
            if time.monotonic() - t0 > timeout_s:
                raise TimeoutError("stream set-up did not finish")
            time.sleep(1e-3)
        return sid.value

    def state(self, sid: int) -> int:
        """0 being set up, 1 ready (a failed set-up raises)."""
        st = C.c_int32(0)
        check(lib().hfx_server_stream_state(self._h, sid, C.byref(st)))
        return st.value

    def clock(self) -> float:
        """Seconds on the server's clock."""
        t = C.c_double()
        check(lib().hfx_server_clock(self._h, C.byref(t)))
        return t.value

    def ingest(self, sid: int, n_frames: int, due_s: Optional[float] = None) -> bool:
        """False when the ring is full (windows not retired yet).  due_s: server-clock time the
        camera delivered the last of these frames (default: now)."""
        if due_s is None:
            rc = lib().hfx_server_ingest(self._h, sid, n_frames)
        else:
            rc = lib().hfx_server_ingest_due(self._h, sid, n_frames, float(due_s))
        if rc == _lib.HFX_NOT_READY:
            return False
        check(rc)
        return True

    def schedule(self) -> int:
        n = C.c_int32(0)
        check(lib().hfx_server_schedule(self._h, C.byref(n)))
        return n.value

    def poll(self, cap: int = 256):
        buf = (_lib.WindowDone * cap)()
        n = C.c_int32(0)
        check(lib().hfx_server_poll(self._h, buf, cap, C.byref(n)))
        return [(buf[i].stream_id, buf[i].window, buf[i].latency_us, buf[i].gpu_us, buf[i].wait_us)
                for i in range(n.value)]

    def window_output(self, sid: int, window: int, shape) -> np.ndarray:
        """Copies a retired window's moments ([n_out][OH][OW] float32) to the host."""
        import torch
        ptr = C.c_void_p()
        check(lib().hfx_server_window_output(self._h, sid, window, C.byref(ptr)))
        out = torch.empty(shape, dtype=torch.float32, device="cuda")
        n = int(np.prod(shape))
        from . import hfx
        src_shape = (C.c_int64 * 1)(n)
        strides = (C.c_int64 * 1)(4)
        check(lib().hfx_strided_copy(ptr, src_shape, strides, 0, 1, 4, C.c_void_p(out.data_ptr()), 0, 0,
                                     hfx._stream(None)))
        return out.cpu().numpy()

    def release(self, sid: int):
        check(lib().hfx_server_release(self._h, sid))

    def info(self) -> dict:
        i = _lib.ServerInfo()
        check(lib().hfx_server_info_get(self._h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in _lib.ServerInfo._fields_}

    def close(self):
        if self._h and lib is not None:      # module globals are gone at interpreter exit
            lib().hfx_server_destroy(self._h)
            self._h = C.c_void_p()

    __del__ = close
