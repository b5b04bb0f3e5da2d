"""Oracle: the .holo recording format and the Holofile source's batch schedule, in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of libhfx's reader/writer and
Holofile source node; the product path never imports it.

Restates:
  * the 64-byte packed header (src/holofile/include/holofile/holofile.hh:88-103):
    u32 magic_number, u16 version, u16 bits_per_pixel, u32 frame_width, u32 frame_height,
    u32 frame_count, u64 data_size_in_bytes, u8 endianness, 35 bytes padding; the magic
    is 0x4F4C4F48 on a little-endian host (holofile.cc:67) and CURRENT_VERSION is 66
    (holofile.hh:99-102);
  * the footer rule (holofile.cc:97-163): the bytes after header + data_size, kept when
    there are 1..1 MiB of them, the first is '{' and they parse as JSON;
  * the reader's checks (holofile.cc:41-91): a short header, a wrong magic, or a frame of
    a bit count not divisible by 8 is an error;
  * the source's batch schedule (sources/holofile.cc:201-235): a batch that would pass
    end_frame restarts at start_frame; each batch is frames [idx, idx + batch_size).

Pinning: the header layout is fixed by the struct above (64 bytes, no implicit padding,
checked by struct.calcsize) and the magic's byte order by the ASCII spelling "HOLO" of
0x4F4C4F48 stored little-endian; the schedule is checked by enumeration in
tests/test_holofile.py.
"""
from __future__ import annotations

import json
import struct
from typing import List, Optional, Tuple

import numpy as np

HEADER_FMT = "<IHHIIIQB35x"
MAGIC = 0x4F4C4F48
VERSION = 66
MAX_FOOTER = 1 << 20


class FormatError(ValueError):
    pass


def pack(frames: np.ndarray, footer: Optional[dict] = None, *, magic: int = MAGIC, version: int = VERSION,
         bits_per_pixel: Optional[int] = None, frame_count: Optional[int] = None,
         data_size: Optional[int] = None) -> bytes:
    """The bytes of a .holo file holding frames [n, h, w]; keyword overrides forge headers."""
    n, h, w = frames.shape
    bpp = bits_per_pixel if bits_per_pixel is not None else 8 * frames.itemsize
    data = np.ascontiguousarray(frames).astype(frames.dtype.newbyteorder("<"), copy=False).tobytes()
    head = struct.pack(HEADER_FMT, magic, version, bpp, w, h, n if frame_count is None else frame_count,
                       len(data) if data_size is None else data_size, 0)
    tail = json.dumps(footer, separators=(",", ":")).encode() if footer is not None else b""
    return head + data + tail


def unpack(raw: bytes) -> Tuple[dict, np.ndarray, Optional[dict]]:
    """(header fields, frames [frame_count, h, w], footer or None) of a .holo file's bytes."""
    if len(raw) < struct.calcsize(HEADER_FMT):
        raise FormatError("Holofile: Incomplete header")
    magic, version, bpp, w, h, count, size, endian = struct.unpack_from(HEADER_FMT, raw)
    if magic != MAGIC:
        raise FormatError("Holofile: Invalid magic number")
    if (w * h * bpp) % 8:
        raise FormatError("Holofile: Invalid frame size")
    hdr = dict(magic=magic, version=version, bits_per_pixel=bpp, width=w, height=h, frame_count=count,
               data_size=size, endianness=endian)
    frame_bytes = w * h * bpp // 8
    dt = np.dtype("<u1") if bpp == 8 else np.dtype("<u2")
    body = raw[64:64 + count * frame_bytes]
    frames = np.frombuffer(body, dt).reshape(count, h, w) if bpp in (8, 16) else None
    footer = None
    tail = raw[64 + size:]
    if 0 < len(tail) <= MAX_FOOTER and tail[:1] == b"{":
        try:
            footer = json.loads(tail.decode())
        except (ValueError, UnicodeDecodeError):
            footer = None
    return hdr, frames, footer


def batch_starts(start: int, end: int, batch: int, n_batches: int, cursor: Optional[int] = None) -> List[int]:
    """First frame of each of n_batches consecutive source batches."""
    idx = start if cursor is None else cursor
    out = []
    for _ in range(n_batches):
        if idx + batch > end:
            idx = start
        out.append(idx)
        idx += batch
    return out
