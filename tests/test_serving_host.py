"""Serving descriptors and admission sizing (host arithmetic only, no GPU)."""
import pytest

from holoflow_amd import serving
from holoflow_amd._lib import HfxInvalidArgument


def test_mixed_streams_follow_the_load_spec():
    specs = serving.mixed_streams(64, seed=3)
    assert {(s.height, s.width) for s in specs} <= set(serving.SIZES)
    for s in specs:
        assert s.window in (32, 64, 128) and s.stride == s.window // 4
        assert 0.2 <= s.z_m <= 0.6 and 20e3 <= s.fs_hz <= 75e3
        df = s.fs_hz / s.window
        assert s.bin_begin == pytest.approx(6e3 / df, abs=1) and s.bin_end <= s.window // 2 + 1
        assert (s.bin_begin - 1) * df < 6e3 <= s.bin_begin * df
    # e.g. T = 64 at 20 kHz -> bins [20, 33): the 15 kHz edge is clipped to T/2 + 1
    assert serving.band_bins(6e3, 15e3, 20e3, 64) == (20, 33)


def test_stream_bytes_counts_ring_pool_outputs_and_plan():
    s = serving.StreamSpec(1024, 1024, 64, 16, 11, 26, 0.3, 37e3, seed=1, stream_id=0)
    N = 1024 * 1024
    R = 64 + 4 * 16
    n_pos = R // 16
    expect = R * N * 2 + 64 * N * 2 + n_pos * 3 * N * 4 + (2 * 15 * N * 8 + N * 8 + (1 << 20)) + N * 8
    assert serving.stream_bytes(s) == expect


def test_stream_bytes_rejects_bad_geometry():
    s = serving.StreamSpec(512, 512, 32, 12, 6, 13, 0.3, 37e3, seed=1, stream_id=0)   # T % S != 0
    with pytest.raises(HfxInvalidArgument):
        serving.stream_bytes(s)


def test_summarize_uses_the_given_periods():
    import numpy as np
    specs = [serving.StreamSpec(512, 512, 32, 8, 5, 9, 0.3, 40e3, seed=1, stream_id=0),
             serving.StreamSpec(512, 512, 32, 8, 5, 9, 0.3, 20e3, seed=2, stream_id=1)]
    lat = np.array([[100.0, 300.0], [100.0, 500.0]], np.float32)
    res = serving.WorkloadResult(1.0, 4, 2, 0, 0, lat, lat / 2, lat / 4, np.zeros(2), np.ones(2), None)
    # camera periods: 200 us and 400 us -> ratios 0.5, 1.5, 0.25, 1.25
    out = serving.summarize(specs, res)
    assert out["frac_windows_within_period"] == 0.5
    # periods at a slower rate (1 ms each): every window within
    out = serving.summarize(specs, res, periods_s=[1e-3, 1e-3])
    assert out["frac_windows_within_period"] == 1.0
    assert out["latency_over_window_period_p99"] < 0.5

