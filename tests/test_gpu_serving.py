"""GPU tests of the serving loop (include/hfx_serve.h): a stream's results are the same
alone and inside a mix of concurrent streams (and the same as eager launches of its plan
on the same frames), admission waits for memory, the ring applies backpressure.

Tolerance: bit-exact.  Every window runs the same kernels on the same frames whether it is
replayed from a graph, launched eagerly, alone or overlapped with other streams."""
import numpy as np
import pytest
import torch

from holoflow_amd import hfx, serving



SMALL = dict(sizes=[(512, 512), (768, 768), (512, 2048)], windows=[32, 64])


def _eager(spec: serving.StreamSpec, n_windows: int) -> np.ndarray:
    """The stream's windows through hfx_moments on a ring rebuilt from its camera frames."""
    cam = spec.to_desc().camera
    P = spec.window                      # the server's synthetic camera plays T frames cyclically
    pool, _ = hfx.synth_frames(cam, 0, P)
    plan = hfx.MomentPlan(hfx.MomentConfig(spec.height, spec.width, spec.window, spec.stride,
                                           spec.bin_begin, spec.bin_end, spec.prop, z_m=spec.z_m,
                                           fs_hz=spec.fs_hz, weight_fs_hz=spec.fs_hz))
    outs = []
    for w in range(n_windows):
        idx = torch.tensor([(w * spec.stride + n) % P for n in range(spec.window)], device="cuda")
        win = pool.index_select(0, idx).contiguous()
        outs.append(plan.moments(win, 0).cpu().numpy())
    return np.stack(outs)


# seed 0 of the full mixed load: every frame size (512^2, 768^2, 1024^2, 2048 wide x 512) and
# every window (32, 64, 128) appears; stream 3 is 1024^2 with T = 128, stream 0 is 2048x512
FULL_MIX_SEED = 0


def test_full_mix_covers_every_size_and_window():  # host only: the load's composition
    specs = serving.mixed_streams(6, seed=FULL_MIX_SEED)
    assert {(s.height, s.width) for s in specs} == set(serving.SIZES)
    assert {s.window for s in specs} == set(serving.WINDOWS)
    assert (specs[3].height, specs[3].width, specs[3].window) == (1024, 1024, 128)
    assert (specs[0].height, specs[0].width) == (512, 2048)


@pytest.mark.gpu
@pytest.mark.parametrize("policy", ["fifo", "edf"])
def test_alone_and_in_mix_results_identical(gpu, policy):
    specs = serving.mixed_streams(6, seed=11, **SMALL)
    k, wps = 2, 12
    arrivals = np.concatenate([[0.0], np.cumsum(np.random.default_rng(1).exponential(0.002, 5))])
    rates = [3000.0] * len(specs)
    mix = serving.run_workload(specs, arrivals, rates, wps, capture_stream=k,
                               server=serving.server_desc(max_streams=4, n_lanes=3, policy=policy),
                               time_limit_s=60)
    assert mix.windows_retired == len(specs) * wps
    assert mix.max_concurrent >= 2
    alone = serving.run_workload([specs[k]], [0.0], [rates[k]], wps, capture_stream=0,
                                 server=serving.server_desc(max_streams=1, n_lanes=1), time_limit_s=60)
    eager = _eager(specs[k], wps)
    assert np.array_equal(alone.capture, mix.capture)
    assert np.array_equal(alone.capture, eager)
    assert (mix.latency_us > 0).all() and (mix.gpu_us > 0).all()


@pytest.mark.gpu
def test_eager_lanes_match_graph_replay(gpu):
    specs = serving.mixed_streams(3, seed=5, **SMALL)
    wps = 10
    g = serving.run_workload(specs, [0.0, 0.0, 0.0], [4000.0] * 3, wps, capture_stream=1,
                             server=serving.server_desc(max_streams=3, n_lanes=2, use_graphs=True), time_limit_s=60)
    e = serving.run_workload(specs, [0.0, 0.0, 0.0], [4000.0] * 3, wps, capture_stream=1,
                             server=serving.server_desc(max_streams=3, n_lanes=2, use_graphs=False), time_limit_s=60)
    assert np.array_equal(g.capture, e.capture)


@pytest.mark.gpu
def test_admission_waits_for_memory(gpu):
    specs = [serving.StreamSpec(512, 512, 32, 8, 5, 9, 0.3, 37e3, seed=i, stream_id=i) for i in range(4)]
    need = serving.stream_bytes(specs[0])
    srv = serving.server_desc(max_streams=8, n_lanes=2, budget_bytes=int(2.5 * need), reserve_bytes=0)
    res = serving.run_workload(specs, [0.0, 0.0, 0.0, 0.0], [2000.0] * 4, 8, server=srv, time_limit_s=60)
    assert res.windows_retired == 32
    assert res.max_concurrent == 2
    assert res.admissions_deferred > 0
    # the third and fourth streams were admitted only after an earlier one was released
    assert min(res.admit_s[2], res.admit_s[3]) >= min(res.finish_s[0], res.finish_s[1])


@pytest.mark.gpu
def test_server_api_backpressure_and_outputs(gpu):
    spec = serving.StreamSpec(512, 512, 32, 8, 5, 9, 0.3, 37e3, seed=3, stream_id=0)
    srv = serving.Server(serving.server_desc(max_streams=2, n_lanes=1))
    sid = srv.admit(spec)
    assert sid is not None
    assert srv.ingest(sid, 32)                      # window 0 ready
    for _ in range(4):
        assert srv.ingest(sid, 8)                   # windows 1..4 ready; ring (64 frames) full
    assert not srv.ingest(sid, 8)                   # would overwrite window 0's frames
    assert srv.schedule() == 5
    done = []
    for _ in range(10000):
        done += srv.poll()
        if len(done) == 5:
            break
    assert [w for _, w, _, _, _ in done] == [0, 1, 2, 3, 4]
    assert srv.ingest(sid, 8)                       # retired windows free the ring
    out = srv.window_output(sid, 4, (3, 512, 512))
    ref = _eager(spec, 5)[4]
    assert np.array_equal(out, ref)
    info = srv.info()
    assert info["active_streams"] == 1 and info["windows_retired"] == 5
    srv.release(sid)
    assert srv.info()["active_streams"] == 0
    srv.close()


@pytest.mark.gpu
@pytest.mark.parametrize("k", [3, 0])
def test_full_mix_alone_and_in_mix_identical(gpu, k):
    """The full mixed load (all four frame sizes, windows 32/64/128): the 1024^2 T = 128 stream
    and a 2048x512 stream give the same moments alone, inside the mix and through eager
    hfx_moments calls on the same frames."""
    specs = serving.mixed_streams(6, seed=FULL_MIX_SEED)
    wps = 8
    arrivals = np.concatenate([[0.0], np.cumsum(np.random.default_rng(2).exponential(0.002, 5))])
    rates = [2000.0] * len(specs)
    mix = serving.run_workload(specs, arrivals, rates, wps, capture_stream=k,
                               server=serving.server_desc(max_streams=6, n_lanes=4), time_limit_s=90)
    assert mix.windows_retired == len(specs) * wps
    assert mix.max_concurrent >= 3
    alone = serving.run_workload([specs[k]], [0.0], [rates[k]], wps, capture_stream=0,
                                 server=serving.server_desc(max_streams=1, n_lanes=1), time_limit_s=90)
    assert np.array_equal(alone.capture, mix.capture)
    assert np.array_equal(alone.capture, _eager(specs[k], wps))


@pytest.mark.gpu
def test_latency_counts_from_the_camera_due_time(gpu):
    """A window whose frames the camera delivered 10 ms before they were ingested (as when the
    ring was full) reports a latency from that due time: it includes the 10 ms, and latency =
    wait before its work starts + its GPU time, all read on the server's clock."""
    spec = serving.StreamSpec(512, 512, 32, 8, 5, 9, 0.3, 37e3, seed=3, stream_id=0)
    srv = serving.Server(serving.server_desc(max_streams=1, n_lanes=1))
    sid = srv.admit(spec)
    t = srv.clock()
    assert srv.ingest(sid, 32, due_s=t - 0.010)
    assert srv.schedule() == 1
    done = []
    for _ in range(100000):
        done += srv.poll()
        if done:
            break
    (_, w, lat, gpu_us, wait), = done
    assert w == 0
    assert wait >= 10_000.0
    assert lat >= 10_000.0 + gpu_us
    assert abs((lat - wait) - gpu_us) < 5.0
    info = srv.info()
    assert 0.0 < info["clock_error_us"] < 10_000.0
    srv.release(sid)
    srv.close()



def _rig(n, H=512, W=512, T=32, prop="asm", fs=37e3):
    """n acquisitions of one plan shape (frame size, window, band) with different z and
    cameras, as a multi-camera rig delivers them."""
    kb, ke = serving.band_bins(6e3, 15e3, fs, T)
    return [serving.StreamSpec(H, W, T, T // 4, kb, ke, 0.2 + 0.07 * i, fs, seed=500 + i, stream_id=i, prop=prop)
            for i in range(n)]


@pytest.mark.gpu
@pytest.mark.parametrize("prop,k", [("asm", 0), ("asm", 3), ("fresnel", 2)])
def test_grouped_launches_match_each_stream_alone(gpu, prop, k):
    """Policy "group": ready windows of streams of one plan shape go out as one
    hfx_moments_group launch; each stream's moments equal those of the stream served alone
    and of eager hfx_moments calls on its frames."""
    specs = _rig(4, prop=prop)
    wps = 10
    mix = serving.run_workload(specs, [0.0] * 4, [3000.0] * 4, wps, capture_stream=k, arrivals="periodic",
                               server=serving.server_desc(max_streams=4, n_lanes=2, policy="group"),
                               time_limit_s=60)
    assert mix.windows_retired == 4 * wps
    assert mix.group_launches > 0 and mix.grouped_windows >= 2 * mix.group_launches
    alone = serving.run_workload([specs[k]], [0.0], [3000.0], wps, capture_stream=0, arrivals="periodic",
                                 server=serving.server_desc(max_streams=1, n_lanes=1), time_limit_s=60)
    assert np.array_equal(alone.capture, mix.capture)
    assert np.array_equal(alone.capture, _eager(specs[k], wps))
