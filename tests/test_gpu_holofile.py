"""Holofile ingest on the GPU: staged uploads, cached source kinds, and a recording fed
through the source -> BatchQueue -> DopplerMoments chain against the oracle.

The chain is the reference's acquisition and preprocessing wiring (graph_builder.cc:171-191,
232-243) ahead of the fused moments node; frames come from the seeded synthetic camera,
written to a .holo file by the oracle's packer.
"""
import numpy as np
import pytest
import torch

from holoflow_amd import holofile
from holoflow_amd.nodes import AsyncNode, Node, desc
from oracle import holofile as OH
from oracle import oracle as O
from oracle.synth import SynthParams, synth_frames
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _st(path, kind, start, end, batch):
    return {"path": path, "load_kind": kind, "start_frame": start, "end_frame": end, "batch_size": batch}


def test_read_to_device_spans_staging_chunks(gpu, tmp_path):
    # 40 frames of 1 MiB: two 32 MiB staging chunks
    rng = np.random.default_rng(11)
    fr = rng.integers(0, 65536, size=(40, 512, 1024), dtype=np.uint32).astype(np.uint16)
    p = tmp_path / "big.holo"
    p.write_bytes(OH.pack(fr))
    with holofile.Reader(str(p)) as r:
        r.seek(3)
        d = r.read_frames_to_device(37)
        torch.cuda.synchronize()
        assert r.tell() == 40
        np.testing.assert_array_equal(d.cpu().numpy().view(np.uint16), fr[3:])
        r.seek(38)
        with pytest.raises(EOFError):
            r.read_frames_to_device(3)


@pytest.mark.parametrize("kind", ["CPUCached", "GPUCached"])
def test_cached_source_schedule(gpu, tmp_path, kind):
    rng = np.random.default_rng(12)
    fr = rng.integers(0, 65536, size=(11, 16, 24), dtype=np.uint32).astype(np.uint16)
    p = tmp_path / "c.holo"
    p.write_bytes(OH.pack(fr, {"v": 2}))
    # start_frame > 0: each view is the frame the cursor names
    n = Node("Holofile", _st(str(p), kind, 2, 11, 4), [])
    starts = OH.batch_starts(2, 11, 4, 5)
    for start in starts:
        out = n.execute()[0]
        assert out.device.type == ("cuda" if kind == "GPUCached" else "cpu")
        np.testing.assert_array_equal(out.cpu().numpy().view(np.uint16), fr[start:start + 4])
    # same range: the cache is kept and the cursor too
    n.update(_st(str(p), kind, 2, 11, 3))
    for start in OH.batch_starts(2, 11, 3, 3, cursor=starts[-1] + 4):
        np.testing.assert_array_equal(n.execute()[0].cpu().numpy().view(np.uint16), fr[start:start + 3])
    # a new range reloads the cache from the file
    n.update(_st(str(p), kind, 0, 5, 5))
    np.testing.assert_array_equal(n.execute()[0].cpu().numpy().view(np.uint16), fr[0:5])
    n.close()


@pytest.mark.parametrize("kind", ["Live", "GPUCached"])
def test_recording_through_source_and_moments(gpu, tmp_path, kind):
    """Live: source batches of 8 frames -> H->D copy -> BatchQueue{T, T} (graph_builder.cc:
    236-239); GPUCached (the reference's LOAD_IN_GPU): source batches of T frames feed the
    moments directly."""
    T, H, W, n_win = 32, 64, 64, 3
    kb, ke = O.band_bins(6e3, 15e3, 37e3, T)
    p = O.PathParams(H, W, T, T, kb, ke, "fresnel", z_m=0.3, fftshift=True)
    frames, _ = synth_frames(SynthParams(H, W, 20.0, seed=9, stream_id=2), 0, n_win * T)
    path = tmp_path / "rec.holo"
    path.write_bytes(OH.pack(frames, {"version": 1}))

    moments = Node("DopplerMoments", {"lambda": p.lambda_m, "dx": p.dx_m, "z": p.z_m, "time_z_begin": kb,
                                      "time_z_end": ke, "stride": T, "prop": "fresnel", "fftshift": True},
                   [desc([T, H, W], "U16")])
    windows = []
    if kind == "Live":
        b = 8
        src = Node("Holofile", _st(str(path), kind, 0, n_win * T, b), [])
        q = AsyncNode("BatchQueue", {"target_capacity": 2 * T, "output_size": T, "output_stride": T},
                      [desc([b, H, W], "U16")])
        for _ in range(n_win * T // b):
            assert q.push(src.execute()[0].cuda())
            w = q.pop()
            if w is not None:
                windows.append(moments.execute(w)[0].cpu().numpy())
    else:
        src = Node("Holofile", _st(str(path), kind, 0, n_win * T, T), [])
        for _ in range(n_win):
            windows.append(moments.execute(src.execute()[0])[0].cpu().numpy())
    torch.cuda.synchronize()
    assert len(windows) == n_win
    for i, out in enumerate(windows):
        ref = O.doppler_moments(frames[i * T:(i + 1) * T], p, "f64")
        assert rel_l2(out, ref) < 1e-5, f"window {i}"
