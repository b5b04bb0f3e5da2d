"""Whole pipelines from graph JSON, run by libhfx's holoflow runtime (include/hfx_graph.h).

``Graph`` compiles a GraphSpec JSON (src/holoflow/src/core/graph_spec.cc:66-170) with the
library's node kinds and runs it: ``step()`` drives every ready section once on the calling
thread (deterministic), ``start()`` / ``stop()`` run one thread per section as the
reference's Scheduler does (graph_exec.cc:160-433).  ``recompile`` consumes a graph and
reuses its buffers, streams and tasks (compiler.cc:663-767, 904-1040).

``doppler_graph`` writes the graph the reference's GraphBuilder wires for the xy moment
view (graph_builder.cc:93-305, 522-647) for a Holofile source, node by node, with the
same kinds and settings; ``fused_doppler_graph`` the same acquisition ahead of the fused
``DopplerMoments`` kind.
"""
from __future__ import annotations

import ctypes as C
import json
from typing import Dict, List, Optional

import numpy as np

from ._lib import check, lib
from .nodes import DTYPES, _ESZ

_BUF = 1 << 20


def _call_text(fn, *args) -> str:
    buf = C.create_string_buffer(_BUF)
    check(fn(*args, buf, len(buf)))
    return buf.value.decode()


def to_dot(spec: Dict) -> str:
    return _call_text(lib().hfx_graph_dot, json.dumps(spec).encode())


class Graph:
    """A compiled graph; ``prev`` (consumed) lends its buffers, streams and tasks."""

    def __init__(self, spec: Dict, prev: Optional["Graph"] = None):
        self.spec = spec
        h = C.c_void_p()
        old = None
        if prev is not None:
            old, prev._h = prev._h, C.c_void_p()
        check(lib().hfx_graph_compile(json.dumps(spec).encode(), old, C.byref(h)))
        self._h = h

    def recompile(self, spec: Dict) -> "Graph":
        return Graph(spec, prev=self)

    def describe(self) -> Dict:
        return json.loads(_call_text(lib().hfx_graph_describe, self._h))

    def step(self) -> int:
        n = C.c_int32(0)
        check(lib().hfx_graph_step(self._h, C.byref(n)))
        return n.value

    def start(self) -> None:
        check(lib().hfx_graph_start(self._h))

    def stop(self) -> None:
        check(lib().hfx_graph_stop(self._h))

    def metrics(self) -> Dict:
        return json.loads(_call_text(lib().hfx_graph_metrics, self._h))

    def output(self, node: str, idx: int = 0):
        """A copy of output `idx` of `node` (torch tensor on its memory side), or None when
        the node owns that output and holds no view right now."""
        import torch
        from .nodes import _copy_raw
        ptr = C.c_void_p()
        buf = C.create_string_buffer(1 << 14)
        check(lib().hfx_graph_output(self._h, node.encode(), idx, C.byref(ptr), buf, len(buf)))
        d = json.loads(buf.value.decode())
        if not ptr.value:
            return None
        if d["mem_loc"] == "Host":
            n = int(np.prod(d["shape"])) * _ESZ[d["dtype"]]
            host = np.empty(n, np.uint8)
            C.memmove(host.ctypes.data, ptr.value, n)
            return torch.from_numpy(host).view(DTYPES[d["dtype"]]).reshape(d["shape"])
        out = torch.empty(d["shape"], dtype=DTYPES[d["dtype"]], device="cuda")
        _copy_raw(ptr.value, d, out.data_ptr(), None)
        torch.cuda.synchronize()
        return out

    def close(self) -> None:
        if self._h and lib is not None:      # module globals are gone at interpreter exit
            lib().hfx_graph_destroy(self._h)
            self._h = C.c_void_p()

    __del__ = close


class _Builder:
    """Adds nodes named like GraphBuilderTasks' helpers (graph_builder_tasks.cc:26-80)."""

    def __init__(self):
        self.nodes: Dict[str, Dict] = {}
        self.edges: List[Dict] = []
        self._count: Dict[str, int] = {}

    def add(self, base: str, kind: str, params: Dict, inputs=()) -> str:
        i = self._count.get(base, 0)
        self._count[base] = i + 1
        name = f"{base}_{i}"
        self.nodes[name] = {"type": kind, "params": params}
        for k, src in enumerate(inputs):
            node, out = src if isinstance(src, tuple) else (src, 0)
            self.edges.append({"from": node, "to": name, "out": out, "in": k})
        return name

    def spec(self) -> Dict:
        return {"nodes": self.nodes, "edges": self.edges}


def doppler_graph(path: str, height: int, width: int, window: int, bin_begin: int, bin_end: int,
                  end_frame: int, prop: str = "fresnel", lambda_m: float = 852e-9, pixel_m: float = 12e-6,
                  z_m: float = 0.3, time_method: str = "RFFT", moment: str = "M0", fftshift: bool = False,
                  load_kind: str = "CPUCached", start_frame: int = 0, load_batch: Optional[int] = None,
                  gpu_in_size: Optional[int] = None, n_pre: int = 8, pp_accumulation: int = 1,
                  gpu_out_size: int = 4, filter2d: Optional[Dict] = None,
                  flatfield_sigma: Optional[tuple] = None, asm_padding: Optional[tuple] = None) -> Dict:
    """The reference GraphBuilder's node chain for one moment image, as graph JSON.

    Mirrors build_acquisition / build_preprocessing / build_time_frequency_analysis /
    build_spatial_filter / build_spatial_propagation / build_xy_view (graph_builder.cc:171-191,
    232-305, 522-647), up to the U8 display conversion (the display queue and widget sink
    are UI).  Weights use the reference's 37 kHz (graph_builder.cc:578-600).  The sliding
    average ring needs target_capacity >= 2 at window 1 (slide_avg.cu:148-156 keeps one
    slot free beyond the window), hence gpu_out_size 4 by default."""
    T = window
    b = _Builder()
    H = b.add("source", "Holofile", {
        "path": path, "load_kind": load_kind, "start_frame": start_frame,
        "end_frame": end_frame, "batch_size": load_batch or T, "keep_cursor": False})
    # build_preprocessing (graph_builder.cc:232-243)
    if load_kind != "GPUCached":
        H = b.add("memcpy", "Memcpy", {"target": "Device"}, [H])
        H = b.add("batched_queue", "BatchQueue", {"target_capacity": gpu_in_size or 2 * T,
                                                  "output_size": T, "output_stride": T}, [H])
    H = b.add("conversion", "Conversion", {"target": "F32", "strategy": "Real"}, [H])
    # build_time_frequency_analysis (graph_builder.cc:245-305)
    H = b.add("reshape", "Reshape", {"shape": [1, T, height, width], "copy": False}, [H])
    H = b.add("batched_queue", "BatchQueue", {"target_capacity": 2 * n_pre, "output_size": n_pre,
                                              "output_stride": n_pre}, [H])
    slice_ = lambda a, z: {"slices": [{}, {"start": a, "stop": z}, {}, {}]}  # noqa: E731
    if time_method == "RFFT":
        F = b.add("rfft", "RFFT", {"axis": 1}, [H])
        F = b.add("slice", "Slice", slice_(bin_begin, bin_end), [F])
        FH = b.add("copy", "Copy", {}, [F])
    else:
        F = b.add("fft", "FFT", {"axis": 1}, [H])
        pos = b.add("slice", "Slice", slice_(bin_begin, bin_end), [F])
        neg = b.add("slice", "Slice", slice_(T - bin_end, T - bin_begin), [F])
        FH = b.add("concatenate", "Concatenate", {"axis": 1}, [pos, neg])
    if filter2d is not None and prop == "fresnel":
        FH = b.add("filter_2d", "Filter2D", dict(filter2d), [FH])
    q = b.add("dual_reader_batch_queue", "DualReaderBatchQueue",
              {"target_capacity": max(2, 2 * n_pre), "window_size": pp_accumulation}, [FH])
    cur = (q, 0)
    # build_spatial_propagation (graph_builder.cc:522-566)
    if prop == "fresnel":
        Z = b.add("fresnel_diffraction", "FresnelDiffraction", {
            "lambda": lambda_m, "dx": pixel_m, "dy": pixel_m, "z": z_m, "axes": [-2, -1],
            "output_magnitude": True}, [cur])
    else:
        st = {"lambda": lambda_m, "dx": pixel_m, "dy": pixel_m, "z": z_m}
        st["filter"] = dict(filter2d) if filter2d is not None else None
        st["padding"] = {"height": asm_padding[0], "width": asm_padding[1]} if asm_padding else None
        Z = b.add("angular_spectrum", "AngularSpectrum", st, [cur])
    magnitude = prop == "fresnel"
    nf = (bin_end - bin_begin) * (2 if time_method == "FFT" else 1)
    # build_xy_view (graph_builder.cc:604-647)
    if moment == "M0":
        R = b.add("mean", "Mean", {"axis": [-3], "keepdims": True}, [Z]) if magnitude else \
            b.add("mean_abs", "MeanAbs", {"axis": [-3], "keepdims": True}, [Z])
    else:
        fs = 37e3
        df = fs / T
        f0, f1 = bin_begin * df, bin_end * df
        w = b.add("arange", "Arange", {"start": f0, "stop": f1, "step": df, "dtype": "F32"})
        if time_method == "FFT":
            wn = b.add("arange", "Arange", {"start": f0 - fs, "stop": f1 - fs, "step": df, "dtype": "F32"})
            w = b.add("concatenate", "Concatenate", {"axis": 0}, [w, wn])
        w = b.add("reshape", "Reshape", {"shape": [1, nf, 1, 1]}, [w])
        if moment == "M2":
            w = b.add("multiply", "Multiply", {}, [w, w])
        A = Z if magnitude else b.add("abs", "Abs", {}, [Z])
        R = b.add("multiply", "Multiply", {}, [w, A])
        R = b.add("mean", "Mean", {"axis": [-3], "keepdims": True}, [R])
    if fftshift:
        R = b.add("fftshift", "FFTShiftNp", {"axes": [-2, -1]}, [R])
    if flatfield_sigma is not None:
        R = b.add("flatfield", "Flatfield", {"sigma_y": flatfield_sigma[0], "sigma_x": flatfield_sigma[1]}, [R])
    R = b.add("mean", "Mean", {"axis": [0], "keepdims": False}, [R])
    R = b.add("slide_avg", "SlidingAverage", {"target_capacity": max(1, gpu_out_size),
                                              "window_size": pp_accumulation, "discard_first": 0}, [R])
    b.add("conversion", "Conversion", {"target": "U8", "strategy": "Scaled"}, [R])
    return b.spec()


def fused_doppler_graph(path: str, window: int, bin_begin: int, bin_end: int, end_frame: int,
                        load_kind: str = "CPUCached", load_batch: Optional[int] = None, **moments) -> Dict:
    """Holofile -> H->D copy -> BatchQueue{T, T} (graph_builder.cc:236-239) -> DopplerMoments
    (the three moment images of each window from one fused node); a GPUCached source feeds
    its T-frame batches to the moments directly."""
    T = window
    b = _Builder()
    H = b.add("source", "Holofile", {"path": path, "load_kind": load_kind, "start_frame": 0,
                                     "end_frame": end_frame, "batch_size": load_batch or T,
                                     "keep_cursor": False})
    if load_kind != "GPUCached":  # a device-cached source hands out windows directly
        H = b.add("memcpy", "Memcpy", {"target": "Device"}, [H])
        H = b.add("batched_queue", "BatchQueue", {"target_capacity": 2 * T, "output_size": T,
                                                  "output_stride": T}, [H])
    settings = {"time_z_begin": bin_begin, "time_z_end": bin_end, "stride": T}
    settings.update(moments)
    b.add("doppler_moments", "DopplerMoments", settings, [H])
    return b.spec()
