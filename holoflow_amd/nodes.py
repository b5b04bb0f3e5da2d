"""Python runner for holoflow node kinds over libhfx's node ABI (include/hfx_nodes.h).

Mirrors the reference's single-node harness (test/holonp/support/sync_task_runner.hh:
26-53): ``infer`` a kind's output descriptors from input descriptors and settings, create
the task on a stream, execute it on device tensors, and ``update`` it with new settings.
Descriptors are dicts {"shape", "dtype", "mem_loc", "strides" (bytes), "offset" (bytes)}.
``infer`` never touches the GPU, so the validation rules can be exercised on any host.
"""
from __future__ import annotations

import ctypes as C
import json
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

DTYPES = {"U8": torch.uint8, "U16": torch.int16, "F32": torch.float32, "CF32": torch.complex64}
_NAMES = {torch.uint8: "U8", torch.int16: "U16", torch.uint16: "U16", torch.float32: "F32",
          torch.complex64: "CF32"}
_ESZ = {"U8": 1, "U16": 2, "F32": 4, "CF32": 8}


def desc(shape: Sequence[int], dtype: str, strides: Optional[Sequence[int]] = None, offset: int = 0,
         mem_loc: str = "Device") -> Dict:
    """A tensor descriptor; strides default to C-contiguous (bytes)."""
    shape = [int(s) for s in shape]
    if strides is None:
        strides, acc = [0] * len(shape), _ESZ[dtype]
        for i in range(len(shape) - 1, -1, -1):
            strides[i] = acc
            acc *= shape[i]
    return {"shape": shape, "dtype": dtype, "mem_loc": mem_loc, "strides": [int(s) for s in strides],
            "offset": int(offset)}


def desc_of(t: torch.Tensor) -> Dict:
    """Descriptor of a device tensor (a view keeps its strides and storage offset)."""
    name = _NAMES[t.dtype]
    e = t.element_size()
    return desc(t.shape, name, [s * e for s in t.stride()], t.storage_offset() * e,
                "Device" if t.is_cuda else "Host")


def _enc(obj) -> bytes:
    return json.dumps(obj).encode()


def kinds() -> List[str]:
    buf = C.create_string_buffer(4096)
    check(lib().hfx_node_kinds(buf, len(buf)))
    return json.loads(buf.value.decode())


def infer(kind: str, settings: Dict, inputs: Sequence[Dict]) -> Dict:
    """factory.infer: raises HfxInvalidArgument where the reference throws invalid_argument."""
    buf = C.create_string_buffer(1 << 16)
    check(lib().hfx_node_infer(kind.encode(), _enc(settings), _enc(list(inputs)), buf, len(buf)))
    return json.loads(buf.value.decode())


def _storage_view(t: torch.Tensor, d: Dict) -> torch.Tensor:
    """The tensor a descriptor describes inside t's storage."""
    e = _ESZ[d["dtype"]]
    return t.as_strided(d["shape"], [s // e for s in d["strides"]], d["offset"] // e)


class Node:
    """A created node: factory.create on `stream`, then execute / update."""

    def __init__(self, kind: str, settings: Dict, inputs: Sequence[Dict], stream=None):
        from .hfx import _stream
        self.kind = kind
        self.settings = dict(settings)
        self.inputs = list(inputs)
        self._h = C.c_void_p()
        self._stream = _stream(stream)
        self.info = infer(kind, settings, inputs)
        check(lib().hfx_node_create(kind.encode(), _enc(settings), _enc(self.inputs), self._stream,
                                    C.byref(self._h)))

    @property
    def output_descs(self) -> List[Dict]:
        return self.info["output_descs"]

    def update(self, settings: Dict, inputs: Optional[Sequence[Dict]] = None, stream=None) -> bool:
        """factory.update; returns True when the existing task was kept."""
        from .hfx import _stream
        inputs = self.inputs if inputs is None else list(inputs)
        reused = C.c_int32(0)
        self._stream = _stream(stream)
        check(lib().hfx_node_update(self._h, _enc(settings), _enc(inputs), self._stream, C.byref(reused)))
        self.settings, self.inputs = dict(settings), inputs
        self.info = infer(self.kind, settings, inputs)
        return bool(reused.value)

    def execute(self, *tensors: torch.Tensor) -> List[torch.Tensor]:
        """Runs the task on tensors whose descriptors match the node's inputs.

        Outputs aliased to an input (in_place) are returned as views of that input's
        storage; outputs the node owns (a cached Holofile source's views into its cache) are
        copied out and released; the others are freshly allocated contiguous tensors, on the
        CPU for host-memory outputs."""
        if len(tensors) != len(self.inputs):
            raise ValueError(f"{self.kind}: expected {len(self.inputs)} inputs, got {len(tensors)}")
        for t, d in zip(tensors, self.inputs):
            if desc_of(t)["shape"] != d["shape"] or _NAMES[t.dtype] != d["dtype"]:
                raise ValueError(f"{self.kind}: tensor {tuple(t.shape)} {t.dtype} does not match {d}")
        aliased = {o: i for i, o in self.info["in_place"]}
        owned = self.info["owned_outputs"]
        outs, optrs = [], []
        for k, od in enumerate(self.output_descs):
            if k in aliased:
                src = tensors[aliased[k]]
                outs.append(_storage_view(src, od))
                optrs.append(src.untyped_storage().data_ptr())
            elif k < len(owned) and owned[k]:
                outs.append(None)
                optrs.append(None)
            else:
                dev = "cpu" if od["mem_loc"] == "Host" else "cuda"
                t = torch.empty(od["shape"], dtype=DTYPES[od["dtype"]], device=dev)
                outs.append(t)
                optrs.append(t.data_ptr())
        iptrs = [t.untyped_storage().data_ptr() for t in tensors]
        ia = (C.c_void_p * max(1, len(iptrs)))(*iptrs)
        oa = (C.c_void_p * max(1, len(optrs)))(*optrs)
        check(lib().hfx_node_execute(self._h, ia, len(iptrs), oa, len(optrs)))
        if any(o is None for o in outs):
            for k, od in enumerate(self.output_descs):
                if outs[k] is None:
                    outs[k] = _copy_owned(oa[k], od, self._stream)
            torch.cuda.synchronize()
            for k, o in enumerate(owned):
                if o:
                    check(lib().hfx_node_release_output(self._h, k))
        return outs

    def close(self):
        if self._h and lib is not None:      # module globals are gone at interpreter exit
            lib().hfx_node_destroy(self._h)
            self._h = C.c_void_p()

    __del__ = close


def _copy_raw(src_ptr: int, d: Dict, dst_ptr: int, stream) -> None:
    nd = max(1, len(d["shape"]))
    shape = (C.c_int64 * nd)(*(d["shape"] or [1]))
    strides = (C.c_int64 * nd)(*(d["strides"] or [_ESZ[d["dtype"]]]))
    check(lib().hfx_strided_copy(C.c_void_p(src_ptr), shape, strides, d["offset"], len(d["shape"]) or 1,
                                 _ESZ[d["dtype"]], C.c_void_p(dst_ptr), 0, 0, stream))


def _copy_owned(ptr: int, od: Dict, stream) -> torch.Tensor:
    """A new tensor holding the contiguous output a node owns at `ptr`."""
    if od["mem_loc"] == "Host":
        n = int(np.prod(od["shape"])) * _ESZ[od["dtype"]]
        host = np.empty(n, np.uint8)
        C.memmove(host.ctypes.data, ptr, n)
        return torch.from_numpy(host).view(DTYPES[od["dtype"]]).reshape(od["shape"])
    out = torch.empty(od["shape"], dtype=DTYPES[od["dtype"]], device="cuda")
    _copy_raw(ptr, desc(od["shape"], od["dtype"]), out.data_ptr(), stream)
    return out


class AsyncNode(Node):
    """An asynchronous node (BatchQueue, SlidingAverage, DualReaderBatchQueue): push input
    batches into owned slots, pop outputs (copied out, then released)."""

    def push(self, t: torch.Tensor, *extra: int) -> bool:
        """Copies t into the next owned input slot and pushes it; False when full.
        `extra` are the further inputs (host U8 scalars), as 0/1 ints."""
        slot = C.c_void_p()
        rc = lib().hfx_node_acquire_input(self._h, C.byref(slot))
        if rc == _lib.HFX_NOT_READY:
            return False
        check(rc)
        t = t.contiguous()
        _copy_raw(t.data_ptr(), desc(t.shape, _NAMES[t.dtype]), slot.value, self._stream)
        flags = [(C.c_uint8 * 1)(int(v)) for v in extra]
        arr = (C.c_void_p * max(1, len(flags)))(*[C.cast(f, C.c_void_p).value for f in flags])
        check(lib().hfx_node_push(self._h, arr, len(flags)))
        return True

    def pop(self):
        """None when not ready; else the outputs: owned ones copied to new device tensors,
        host U8 flags as ints."""
        ods = self.output_descs
        owned = self.info["owned_outputs"]
        flags = {i: (C.c_uint8 * 1)(0) for i, o in enumerate(owned) if not o}
        ptrs = (C.c_void_p * len(ods))(*[C.cast(flags[i], C.c_void_p).value if i in flags else None
                                          for i in range(len(ods))])
        rc = lib().hfx_node_pop(self._h, ptrs, len(ods))
        if rc == _lib.HFX_NOT_READY:
            return None
        check(rc)
        outs = []
        for i, od in enumerate(ods):
            if i in flags:
                outs.append(int(flags[i][0]))
                continue
            outs.append(_copy_owned(ptrs[i], od, self._stream))
        torch.cuda.synchronize()
        for i, o in enumerate(owned):
            if o:
                check(lib().hfx_node_release_output(self._h, i))
        return outs[0] if len(outs) == 1 else outs
