"""The .holo format (libhfx reader/writer) and the Holofile source node, on the CPU.

The checker is oracle/holofile.py: the reference's packed header (holofile.hh:88-103),
footer rule (holofile.cc:97-163), reader checks (holofile.cc:41-91) and the source's batch
schedule (sources/holofile.cc:201-235).  Files are built by the oracle (independently of
libhfx's writer) and read by libhfx, and the other way round.  Cached load kinds allocate
pinned or device memory and are covered by tests/test_gpu_holofile.py.
"""
import json
import os
import struct
import time

import numpy as np
import pytest

from holoflow_amd import holofile, nodes
from holoflow_amd._lib import HfxInvalidArgument
from oracle import holofile as OH


def _frames(n, h, w, dtype, seed=0):
    rng = np.random.default_rng(seed)
    return rng.integers(0, np.iinfo(dtype).max + 1, size=(n, h, w), dtype=np.uint64).astype(dtype)


def _put(tmp_path, raw, name="f.holo"):
    p = tmp_path / name
    p.write_bytes(raw)
    return str(p)


# ------------------------------------------------------------------ oracle pins
def test_oracle_header_layout_is_the_packed_struct():
    assert struct.calcsize(OH.HEADER_FMT) == 64
    raw = OH.pack(np.zeros((1, 2, 3), np.uint8))
    assert raw[:4] == b"HOLO"                      # 0x4F4C4F48 stored little-endian
    assert raw[4:6] == (66).to_bytes(2, "little")
    assert len(raw) == 64 + 6


def test_oracle_batch_schedule_by_enumeration():
    # start 1, end 7, batch 4: [1..4], then 5+4 > 7 wraps to 1
    assert OH.batch_starts(1, 7, 4, 4) == [1, 1, 1, 1]
    assert OH.batch_starts(0, 6, 2, 5) == [0, 2, 4, 0, 2]
    assert OH.batch_starts(0, 7, 3, 4) == [0, 3, 0, 3]
    assert OH.batch_starts(0, 7, 2, 2, cursor=5) == [5, 0]


# ------------------------------------------------------------------ writer -> oracle
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_writer_bytes_match_oracle(tmp_path, dtype):
    fr = _frames(5, 4, 6, dtype, seed=1)
    footer = {"version": 1, "compute": {"batch": [1, 2]}, "name": "é"}
    p = str(tmp_path / "w.holo")
    holofile.write(p, fr, footer)
    raw = open(p, "rb").read()
    hdr, frames, foot = OH.unpack(raw)
    assert hdr == dict(magic=OH.MAGIC, version=66, bits_per_pixel=8 * fr.itemsize, width=6, height=4,
                       frame_count=5, data_size=fr.nbytes, endianness=0)
    np.testing.assert_array_equal(frames, fr)
    assert foot == footer
    assert raw[64 + fr.nbytes:] == json.dumps(footer, separators=(",", ":")).encode()
    assert raw[45:64] == bytes(19)                 # padding is zero


def test_writer_without_footer_and_empty(tmp_path):
    p = str(tmp_path / "e.holo")
    holofile.write(p, np.zeros((0, 3, 4), np.uint16))
    raw = open(p, "rb").read()
    assert len(raw) == 64
    hdr, frames, foot = OH.unpack(raw)
    assert hdr["frame_count"] == 0 and hdr["data_size"] == 0 and foot is None
    with pytest.raises(ValueError):
        holofile.write(p, np.zeros((2, 3), np.uint16))


# ------------------------------------------------------------------ oracle -> reader
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_reader_reads_oracle_file(tmp_path, dtype):
    fr = _frames(9, 8, 5, dtype, seed=2)
    p = _put(tmp_path, OH.pack(fr, {"a": 1}))
    with holofile.Reader(p) as r:
        h = r.header
        assert (h.width, h.height, h.bits_per_pixel, h.version, h.frame_count) == (5, 8, 8 * fr.itemsize, 66, 9)
        assert h.data_size == fr.nbytes and h.frame_bytes == fr[0].nbytes
        assert r.footer == {"a": 1}
        assert r.tell() == 0
        np.testing.assert_array_equal(r.read_frames(3), fr[:3])
        assert r.tell() == 3
        r.seek(7)
        np.testing.assert_array_equal(r.read_frames(2), fr[7:9])
        assert r.tell() == 9
        r.seek(1)
        np.testing.assert_array_equal(r.read_frames(8), fr[1:9])


def test_reader_eof_advances_by_frames_read(tmp_path):
    fr = _frames(4, 3, 3, np.uint8, seed=3)
    p = _put(tmp_path, OH.pack(fr))
    r = holofile.Reader(p)
    r.seek(2)
    with pytest.raises(EOFError, match="End of file"):
        r.read_frames(3)
    assert r.tell() == 4                            # Reader::read_frames counts what it read
    r.close()


@pytest.mark.parametrize("tail,expect", [
    (b"", None),                                    # no footer
    (b"[1,2]", None),                               # not an object
    (b"{\"a\": ", None),                            # does not parse
    (b"{\"k\": [1, 2.5, \"x\"], \"n\": null}", {"k": [1, 2.5, "x"], "n": None}),
])
def test_footer_rule(tmp_path, tail, expect):
    fr = _frames(2, 2, 2, np.uint16, seed=4)
    raw = OH.pack(fr) + tail
    p = _put(tmp_path, raw)
    assert OH.unpack(raw)[2] == expect
    assert holofile.Reader(p).footer == expect


def test_footer_over_one_mib_is_dropped(tmp_path):
    fr = _frames(1, 2, 2, np.uint8)
    big = b"{\"x\":\"" + b"a" * (1 << 20) + b"\"}"
    raw = OH.pack(fr) + big
    assert OH.unpack(raw)[2] is None
    assert holofile.Reader(_put(tmp_path, raw)).footer is None


@pytest.mark.parametrize("raw_fn,msg", [
    (lambda: OH.pack(np.zeros((1, 2, 2), np.uint8))[:40], "Incomplete header"),
    (lambda: OH.pack(np.zeros((1, 2, 2), np.uint8), magic=0x484F4C4F), "Invalid magic number"),
    (lambda: OH.pack(np.zeros((1, 1, 1), np.uint8), bits_per_pixel=12), "Invalid frame size"),
])
def test_reader_rejects_malformed_files(tmp_path, raw_fn, msg):
    raw = raw_fn()
    with pytest.raises(OH.FormatError, match=msg):
        OH.unpack(raw)
    with pytest.raises(HfxInvalidArgument, match=msg):
        holofile.Reader(_put(tmp_path, raw))


def test_reader_missing_file(tmp_path):
    with pytest.raises(HfxInvalidArgument, match="Failed to open"):
        holofile.Reader(str(tmp_path / "nope.holo"))


# ------------------------------------------------------------------ Holofile source node
def _st(path, kind="Live", start=0, end=None, batch=2, **kw):
    s = {"path": path, "load_kind": kind, "start_frame": start, "end_frame": end, "batch_size": batch}
    s.update(kw)
    return s


@pytest.fixture
def holo16(tmp_path):
    fr = _frames(7, 4, 6, np.uint16, seed=5)
    return _put(tmp_path, OH.pack(fr, {"v": 1}), "s16.holo"), fr


def test_infer_output_descs(holo16, tmp_path):
    p, fr = holo16
    for kind, loc, owned in (("Live", "Host", False), ("CPUCached", "Host", True), ("GPUCached", "Device", True)):
        r = nodes.infer("Holofile", _st(p, kind, 0, 7, 3), [])
        assert r["kind"] == "Sync"
        d = r["output_descs"][0]
        assert (d["shape"], d["dtype"], d["mem_loc"]) == ([3, 4, 6], "U16", loc)
        assert r["owned_outputs"] == [owned]
    p8 = _put(tmp_path, OH.pack(_frames(2, 3, 5, np.uint8)), "s8.holo")
    assert nodes.infer("Holofile", _st(p8, end=2, batch=1), [])["output_descs"][0]["dtype"] == "U8"


@pytest.mark.parametrize("over,msg", [
    (dict(start=3, end=3), "Invalid frame range"),
    (dict(start=-1, end=3), "Invalid frame range"),
    (dict(batch=0), "Batch size must be positive"),
    (dict(max_fps=0), "max_fps must be positive"),
    (dict(end=8), "end_frame exceeds total frames"),
    (dict(start=5, end=7, batch=3), "Batch size exceeds available frames"),
    (dict(kind="Disk"), "Invalid LoadKind string"),
])
def test_infer_validation(holo16, over, msg):
    p, _ = holo16
    s = _st(p, over.pop("kind", "Live"), over.pop("start", 0), over.pop("end", 7), over.pop("batch", 2), **over)
    with pytest.raises(HfxInvalidArgument, match=msg):
        nodes.infer("Holofile", s, [])


def test_infer_rejects_inputs_and_unsupported_depth(holo16, tmp_path):
    p, _ = holo16
    with pytest.raises(HfxInvalidArgument, match="must have no inputs"):
        nodes.infer("Holofile", _st(p, end=7), [nodes.desc([1], "U8")])
    p32 = _put(tmp_path, OH.pack(np.zeros((2, 2, 2), np.uint32)), "s32.holo")
    with pytest.raises(HfxInvalidArgument, match="Unsupported bits_per_pixel: 32"):
        nodes.infer("Holofile", _st(p32, end=2, batch=1), [])


def test_live_source_schedule_and_update(holo16):
    p, fr = holo16
    n = nodes.Node("Holofile", _st(p, "Live", 1, 7, 4), [], stream=0)
    for start in OH.batch_starts(1, 7, 4, 3):
        out = n.execute()[0]
        assert out.device.type == "cpu"
        np.testing.assert_array_equal(out.numpy().view(np.uint16), fr[start:start + 4])
    # keep_cursor (default true): the cursor (5) stays inside the new range [0, 7)
    n.update(_st(p, "Live", 0, 7, 2), stream=0)
    for start in OH.batch_starts(0, 7, 2, 3, cursor=5):
        np.testing.assert_array_equal(n.execute()[0].numpy().view(np.uint16), fr[start:start + 2])
    # cursor 3 falls outside [4, 7): back to start_frame
    n.update(_st(p, "Live", 4, 7, 1), stream=0)
    np.testing.assert_array_equal(n.execute()[0].numpy().view(np.uint16), fr[4:5])
    # keep_cursor false restarts at start_frame
    n.update(_st(p, "Live", 0, 7, 2, keep_cursor=False), stream=0)
    np.testing.assert_array_equal(n.execute()[0].numpy().view(np.uint16), fr[0:2])
    with pytest.raises(HfxInvalidArgument, match="Cannot release output for Live"):
        from holoflow_amd._lib import check, lib
        check(lib().hfx_node_release_output(n._h, 0))
    n.close()


def test_live_source_max_fps_paces_batches(holo16):
    p, _ = holo16
    n = nodes.Node("Holofile", _st(p, "Live", 0, 7, 2, max_fps=200), [], stream=0)  # 10 ms a batch
    t0 = time.perf_counter()
    for _ in range(6):
        n.execute()
    dt = time.perf_counter() - t0
    assert dt >= 0.045, dt                          # 5 periods after the first batch
    n.close()
