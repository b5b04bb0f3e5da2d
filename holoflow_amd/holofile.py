"""The .holo recording format over libhfx (include/hfx_holofile.h).

Mirrors the reference's ``holofile::Reader`` / ``Writer`` (src/holofile/include/holofile/
holofile.hh:87-153, src/holofile/src/holofile.cc:41-268): a 64-byte packed header (magic
0x4F4C4F48, version 66, 8 or 16 bits per pixel, width, height, frame count, data size),
the raw frames, and an optional JSON footer holding the pipeline settings.  Malformed files
raise ``HfxInvalidArgument`` with the reference's messages ("Holofile: Invalid magic
number", ...); reading past the last frame raises ``EOFError``.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from ._lib import check, lib

MAGIC = 0x4F4C4F48
VERSION = 66
HEADER_BYTES = 64


@dataclass
class Header:
    width: int
    height: int
    bits_per_pixel: int
    version: int
    frame_count: int
    data_size: int
    frame_bytes: int

    @property
    def dtype(self):
        return np.uint8 if self.bits_per_pixel == 8 else np.uint16


def _check_read(code: int) -> None:
    if code == _lib.HFX_EOF:
        raise EOFError(lib().hfx_last_error().decode(errors="replace"))
    check(code)


class Reader:
    """Reader(path): header, footer, seek / tell and read_frames."""

    def __init__(self, path: str):
        self._h = C.c_void_p()
        check(lib().hfx_holofile_open(os.fsencode(path), C.byref(self._h)))
        info = _lib.HolofileInfo()
        check(lib().hfx_holofile_get_info(self._h, C.byref(info)))
        self.header = Header(info.width, info.height, info.bits_per_pixel, info.version, info.frame_count,
                             info.data_size, info.frame_bytes)
        self._has_footer = bool(info.has_footer)

    @property
    def footer(self) -> Optional[dict]:
        """The footer's pipeline settings, or None when the file has no valid footer."""
        if not self._has_footer:
            return None
        buf = C.create_string_buffer(1 << 20 | 1)
        check(lib().hfx_holofile_footer(self._h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def seek(self, frame: int) -> None:
        check(lib().hfx_holofile_seek(self._h, int(frame)))

    def tell(self) -> int:
        return int(lib().hfx_holofile_tell(self._h))

    def read_frames(self, n: int) -> np.ndarray:
        """The next n frames as a [n, height, width] array."""
        h = self.header
        out = np.empty((n, h.height, h.width), h.dtype)
        _check_read(lib().hfx_holofile_read(self._h, out.ctypes.data, int(n)))
        return out

    def read_frames_to_device(self, n: int, stream=None):
        """The next n frames as a [n, height, width] device tensor (uint8, or int16 holding
        the uint16 bit pattern), uploaded through the reader's pinned staging buffers."""
        import torch
        from .hfx import _stream
        h = self.header
        out = torch.empty((n, h.height, h.width), dtype=torch.uint8 if h.bits_per_pixel == 8 else torch.int16,
                          device="cuda")
        _check_read(lib().hfx_holofile_read_to_device(self._h, C.c_void_p(out.data_ptr()), int(n), _stream(stream)))
        return out

    def close(self) -> None:
        if self._h and lib is not None:      # module globals are gone at interpreter exit
            lib().hfx_holofile_close(self._h)
            self._h = C.c_void_p()

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def write(path: str, frames: np.ndarray, footer: Optional[dict] = None) -> None:
    """Writes frames [n, height, width] (uint8 or uint16) with an optional footer JSON."""
    frames = np.ascontiguousarray(frames)
    if frames.ndim != 3 or frames.dtype not in (np.uint8, np.uint16):
        raise ValueError("frames must be a [n, height, width] uint8 or uint16 array")
    n, height, width = frames.shape
    text = json.dumps(footer, separators=(",", ":")).encode() if footer is not None else None
    check(lib().hfx_holofile_write(os.fsencode(path), width, height, 8 * frames.itemsize,
                                   frames.ctypes.data if n else None, n, text))
