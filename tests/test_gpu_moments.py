"""GPU parity of the fused Doppler-moment path (hfx_moments) against the oracle.

FP32 path: relative L2 <= 1e-5 against the FP32 restatement (torch-CPU/MKL) and against
the FP64 oracle, per moment image.  Mixed-precision time DFT: errors against FP64 are
reported and bounded (FP16 split <= 1e-5, FP16 <= 1e-3, BF16 <= 2e-2 on M0)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL_FP32 = 1e-5


def _frames(H, W, n, fs=37e3, seed=7, radius=None):
    from holoflow_amd import hfx

    d = hfx.synth_desc(H, W, disk_radius_px=radius or 0.39 * min(H, W), fs_hz=fs, seed=seed)
    ring, _ = hfx.synth_frames(d, 0, n)
    return ring


def _cfg(p: O.PathParams, precision="fp32"):
    from holoflow_amd import hfx

    return hfx.MomentConfig(p.height, p.width, p.window, p.stride, p.bin_begin, p.bin_end, p.prop,
                            p.lambda_m, p.dx_m, p.dy_m, p.z_m, weight_fs_hz=p.weight_fs_hz,
                            precision=precision, fftshift=p.fftshift, power2=p.power2,
                            both_sidebands=p.both_sidebands, pad=p.pad, filt=p.filt,
                            split_bin=p.split_bin)


def _check(p: O.PathParams, ring, first, precision="fp32", against=("f32", "f64"), tol=TOL_FP32):
    from holoflow_amd import hfx

    plan = hfx.MomentPlan(_cfg(p, precision))
    out = plan.moments(ring, first).cpu().numpy()
    host = ring.cpu().numpy().view(np.uint16)
    idx = (first + np.arange(p.window)) % host.shape[0]
    win = host[idx]
    errs = {}
    for prec in against:
        ref = O.doppler_moments(win, p, prec)
        assert out.shape == ref.shape
        errs[prec] = [rel_l2(out[i], ref[i]) for i in range(ref.shape[0])]
    print(p, precision, errs)
    for prec, e in errs.items():
        assert max(e) < tol, (prec, e)
    return out, errs


def test_fresnel_small_wrapping_ring(gpu):
    p = O.PathParams(64, 64, 32, 8, 6, 13, "fresnel", 852e-9, 12e-6, 12e-6, 0.45, fftshift=True)
    ring = _frames(64, 64, 48)
    _check(p, ring, 40)


def test_asm_small_padding_filter_split(gpu):
    p = O.PathParams(48, 64, 32, 8, 5, 14, "asm", 852e-9, 12e-6, 12e-6, 0.3, pad=(64, 96),
                     filt=(4, 40, 2, 6), split_bin=9)
    _check(p, _frames(48, 64, 32), 0)


def test_fresnel_both_sidebands_and_power(gpu):
    p = O.PathParams(64, 96, 64, 16, 11, 26, "fresnel", 852e-9, 12e-6, 12e-6, 0.3, both_sidebands=True,
                     power2=True)
    _check(p, _frames(64, 96, 64), 0)


def test_dc_bin_in_band(gpu):
    p = O.PathParams(32, 32, 32, 32, 0, 6, "asm", 852e-9, 12e-6, 12e-6, 0.2)
    _check(p, _frames(32, 32, 32), 0)


def test_cfg0_512_asm(gpu):
    p = O.config_params("cfg0_512_asm")
    _check(p, _frames(512, 512, 32), 0)


def test_cfg1_1024_fresnel_one_window(gpu):
    p = O.config_params("cfg1_1024_fresnel")
    ring = _frames(1024, 1024, 80, radius=400)
    _check(p, ring, 16)


def test_cfg4_2048_asm_band_split_one_window(gpu):
    # 2048^2, T=128, f_s 10 kHz, bins [13, 65) with the band split at 52 (5 outputs):
    # FP32 and FP16-split time DFT against the FP64 oracle
    p = O.config_params("cfg4_2048_asm")
    assert (p.bin_begin, p.bin_end, p.split_bin) == (13, 65, 52)
    ring = _frames(2048, 2048, 128, fs=10e3, radius=900)
    from holoflow_amd import hfx

    win = ring.cpu().numpy().view(np.uint16)
    ref = O.doppler_moments(win, p, "f64")
    for precision in ("fp32", "fp16_split"):
        plan = hfx.MomentPlan(_cfg(p, precision))
        out = plan.moments(ring, 0).cpu().numpy()
        plan.close()
        errs = [rel_l2(out[i], ref[i]) for i in range(5)]
        print("cfg4", precision, errs)
        assert max(errs) < TOL_FP32, (precision, errs)


@pytest.mark.parametrize("precision,tol", [("fp16_split", 1e-5), ("fp16", 1e-3), ("bf16", 2e-2)])
def test_mixed_precision_time_dft_vs_fp64(gpu, precision, tol):
    p = O.config_params("cfg1_1024_fresnel")
    p = O.PathParams(256, 256, 64, 16, p.bin_begin, p.bin_end, "fresnel", p.lambda_m, p.dx_m, p.dy_m, p.z_m,
                     fftshift=True)
    ring = _frames(256, 256, 64, radius=100)
    _, errs = _check(p, ring, 0, precision, against=("f64",), tol=tol)
    if precision == "fp16_split":
        assert errs["f64"][0] <= 1e-4   # the mixed path's M0 bar


@pytest.mark.parametrize("precision,tol", [("fp16_split", 1e-5), ("fp16", 1e-3), ("bf16", 2e-2)])
@pytest.mark.parametrize("H,W,T,S,kb,ke,prop", [(512, 512, 64, 16, 11, 26, "fresnel"),
                                                (64, 1024, 32, 8, 6, 13, "asm")])
def test_fused_16bit_front_vs_fp64(gpu, precision, tol, H, W, T, S, kb, ke, prop):
    """The fused time-DFT + row-FFT kernel in the 16-bit modes (32-frame MFMA steps, several rows
    per workgroup at 512 rows) against the FP64 oracle, on a window that wraps around the ring's
    end."""
    p = O.PathParams(H, W, T, S, kb, ke, prop, 852e-9, 12e-6, 12e-6, 0.45, fftshift=True)
    F = T + 2 * S
    ring = _frames(H, W, F, radius=0.39 * min(H, W))
    _, errs = _check(p, ring, F - S, precision, against=("f64",), tol=tol)
    if precision == "fp16_split":
        assert errs["f64"][0] <= 1e-4


@pytest.mark.parametrize("precision,tol", [("fp16_split", 1e-5), ("fp16", 1e-3), ("bf16", 2e-2)])
@pytest.mark.parametrize("W,kb,ke,both", [(512, 3, 9, False), (512, 2, 16, False), (768, 3, 9, False),
                                          (768, 2, 16, False), (1024, 3, 9, False), (1024, 2, 16, False),
                                          (768, 2, 16, True)])  # 30 planes: more than one LDS group
def test_every_16bit_front_variant_vs_fp64(gpu, precision, tol, W, kb, ke, both):
    """Each compiled 16-bit fused kernel (row widths 512/768/1024, one or two bin tiles, and the
    several-group form of a two-sided band) against the FP64 oracle, with the bars of the modes."""
    H, T = 256, 64
    prop = "asm" if W == 768 else "fresnel"
    p = O.PathParams(H, W, T, T // 4, kb, ke, prop, 852e-9, 12e-6, 12e-6, 0.4, fftshift=True,
                     both_sidebands=both)
    F = T + T // 4
    _check(p, _frames(H, W, F, seed=W + kb), T // 4, precision, against=("f64",), tol=tol)


@pytest.mark.parametrize("W,T,kb,ke,prop,both", [(512, 32, 0, 12, "fresnel", False),
                                                 (768, 64, 5, 17, "asm", False),
                                                 (512, 128, 20, 35, "fresnel", True),
                                                 (1024, 64, 11, 26, "fresnel", True)])
def test_folded_front_vs_oracles(gpu, W, T, kb, ke, prop, both):
    """The FP32 front kernel folds the real-input symmetry of the time DFT (x_p + x_{T-p} on the
    cosine rows, x_p - x_{T-p} on the sine rows, x_0 and x_{T/2} apart) for 9..16 bins: window
    lengths 32/64/128, the DC bin in the band (centring offset), both sidebands (16 bins)."""
    p = O.PathParams(64, W, T, T // 4, kb, ke, prop, 852e-9, 12e-6, 12e-6, 0.4, fftshift=True,
                     both_sidebands=both)
    F = T + T // 2
    _check(p, _frames(64, W, F), F - T // 4)


@pytest.mark.parametrize("W", [512, 768, 1024])
@pytest.mark.parametrize("T", [32, 64, 128])
@pytest.mark.parametrize("kb,ke", [(3, 9), (2, 16)])
def test_every_front_variant_vs_fp64(gpu, W, T, kb, ke):
    """Each compiled FP32 front kernel (row widths 512/768/1024, windows 32/64/128, up to 8 bins
    interleaved or 9..16 folded) against the FP64 oracle.  768 rows on at most 256 workgroups:
    each workgroup runs three rows, so the row loop runs as well as its first and last row."""
    H = 768
    prop = "fresnel" if (W // 256 + T // 32) % 2 else "asm"
    p = O.PathParams(H, W, T, T // 4, kb, ke, prop, 852e-9, 12e-6, 12e-6, 0.4, fftshift=True)
    F = T + T // 4
    _check(p, _frames(H, W, F, seed=W + T + kb), T // 4, against=("f64",))


@pytest.mark.parametrize("H,W,T,kb,ke,prop,both", [(48, 64, 64, 1, 30, "asm", False),    # 29 bins: two tiles
                                                   (48, 64, 32, 0, 16, "fresnel", True),  # 17 bins with DC
                                                   (48, 64, 128, 10, 50, "fresnel", False),  # 40: three
                                                   (4, 24, 128, 1, 60, "asm", False),  # 59: four, ragged
                                                   (4, 24, 32, 3, 9, "fresnel", False)])  # 6: interleaved
def test_folded_standalone_dft_vs_oracles(gpu, H, W, T, kb, ke, prop, both):
    """The stand-alone FP32 time DFT (widths without the fused kernel) folds the real-input
    symmetry over several 16-bin tiles (up to 8 bins: the interleaved cos/-sin contraction).
    The persistent kernels walk 64-pixel tiles; at 4 x 24 = 96 pixels the second tile is
    ragged (its lanes past the end load clamped pixels and store nothing)."""
    p = O.PathParams(H, W, T, T // 4, kb, ke, prop, 852e-9, 12e-6, 12e-6, 0.35, fftshift=True,
                     both_sidebands=both)
    F = T + T // 2
    _check(p, _frames(H, W, F), F - T // 4)


@pytest.mark.parametrize("H,W,kb,ke,prop", [(16, 512, 11, 26, "fresnel"),  # fused front, one tile
                                              (24, 64, 1, 30, "asm")])       # stand-alone, two tiles
def test_full_16bit_samples_fp32(gpu, H, W, kb, ke, prop):
    """The FP32 path takes any uint16 sample, not only 12-bit ones: the folded time DFT forms
    x_p + x_{T-p} and x_p - x_{T-p} exactly for the whole 16-bit range (uniform samples over
    [0, 65535], against both oracles)."""
    T = 64
    p = O.PathParams(H, W, T, T // 4, kb, ke, prop, 852e-9, 12e-6, 12e-6, 0.4, fftshift=True)
    rng = np.random.default_rng(16)
    host = rng.integers(0, 65536, size=(T + 8, H, W), dtype=np.uint16)
    assert host.max() > 60000
    ring = torch.from_numpy(host.view(np.int16)).cuda()
    _check(p, ring, 3)


@pytest.mark.parametrize("precision,tol", [("fp32", TOL_FP32), ("fp16_split", 1e-5), ("fp16", 1e-3),
                                           ("bf16", 2e-2)])
@pytest.mark.parametrize("kb,ke", [(1, 7), (1, 15), (1, 31), (1, 47), (1, 65)])
def test_every_standalone_dft_variant_vs_fp64(gpu, precision, tol, kb, ke):
    """The stand-alone time DFT (a width without the fused kernel) for 6, 14, 30, 46 and 64
    bins at T = 128: in FP32 the interleaved kernel and the folded one over one to four 16-bin
    tiles, in the 16-bit modes the interleaved kernel over 1, 2, 4, 6 and 8 row tiles."""
    H, W, T = 48, 64, 128
    p = O.PathParams(H, W, T, T // 4, kb, ke, "fresnel", 852e-9, 12e-6, 12e-6, 0.4, fftshift=True)
    F = T + T // 4
    _check(p, _frames(H, W, F, seed=ke), T // 4, precision, against=("f64",), tol=tol)


@pytest.mark.parametrize("prop", ["fresnel", "asm"])
@pytest.mark.parametrize("H,W", [(256, 256), (512, 768), (768, 512), (1024, 256), (256, 1024), (2048, 256),
                                 (256, 2048)])
def test_every_pass_length_vs_fp64(gpu, prop, H, W):
    """The wave-level row and column passes at every length the path compiles (256, 512, 768,
    1024, 2048 points, rows and columns), through the fused front or the stand-alone time DFT
    with the forward row pass, for both propagations, against the FP64 oracle."""
    T = 32
    p = O.PathParams(H, W, T, T // 4, 2, 9, prop, 852e-9, 12e-6, 12e-6, 0.4, fftshift=True)
    F = T + T // 4
    _check(p, _frames(H, W, F, seed=H + W), T // 4, against=("f64",))


def test_fresnel_band_split_1024_rows(gpu):
    """Band-split moments (five sums) out of the 1024-row column pass, against both oracles."""
    p = O.PathParams(1024, 256, 32, 8, 4, 12, "fresnel", 852e-9, 12e-6, 12e-6, 0.45, fftshift=True,
                     split_bin=8)
    out, _ = _check(p, _frames(1024, 256, 40, radius=100), 8)
    assert out.shape == (5, 1024, 256)


@pytest.mark.parametrize("H,W,T,kb,ke,split", [(256, 512, 64, 5, 30, 15),    # 25 bins: stand-alone DFT
                                                 (128, 1024, 32, 3, 12, 7),     # 9 bins, 1024-point rows
                                                 (2, 2048, 32, 3, 12, 7),       # 2048-point rows, 2 rows
                                                 (64, 2048, 64, 5, 30, 0)])     # 2048-point rows, no split
def test_asm_band_split_from_the_row_pass(gpu, H, W, T, kb, ke, split):
    """Angular spectrum with the band split: the moments come out of the inverse row pass,
    which runs the planes below the split and those above it as two loops (M0 low stored at
    the split).  The 25-bin band has no fused front: its time DFT runs stand-alone (folded, two
    16-bin tiles) and the forward row pass reads its natural-layout planes."""
    p = O.PathParams(H, W, T, T // 4, kb, ke, "asm", 852e-9, 12e-6, 12e-6, 0.35, fftshift=True, split_bin=split)
    out, _ = _check(p, _frames(H, W, T + T // 2, radius=0.4 * H), T // 4)
    assert out.shape == (5 if split else 3, H, W)


def test_batch_matches_single_windows(gpu):
    from holoflow_amd import hfx

    p = O.PathParams(64, 64, 32, 8, 6, 13, "fresnel", 852e-9, 12e-6, 12e-6, 0.45, fftshift=True)
    ring = _frames(64, 64, 40)
    plan = hfx.MomentPlan(_cfg(p))
    batch = plan.moments_batch(ring, 4, 5)
    for w in range(5):
        single = plan.moments(ring, (4 + 8 * w) % 40)
        assert torch.equal(batch[w], single)


@pytest.mark.parametrize("prop,W,nwin", [("fresnel", 1024, 4), ("fresnel", 512, 3), ("asm", 1024, 2)])
def test_multi_window_front_matches_single_windows(gpu, prop, W, nwin):
    """batch_windows > 1: one pipelined front launch runs across several windows' rows; every
    window must equal its own single-window result bit for bit (7 windows: a partial last
    group; first frame near the ring's end: windows wrap)."""
    from holoflow_amd import hfx

    T, S, F = 64, 16, 200
    p = O.PathParams(64, W, T, S, 11, 26, prop, 852e-9, 12e-6, 12e-6, 0.45, fftshift=True)
    ring = _frames(64, W, F)
    cfg = _cfg(p)
    cfg.batch_windows = nwin
    multi = hfx.MomentPlan(cfg)
    single = hfx.MomentPlan(_cfg(p))
    first = F - 40
    batch = multi.moments_batch(ring, first, 7)
    for w in range(7):
        ref = single.moments(ring, (first + S * w) % F)
        assert torch.equal(batch[w], ref), f"window {w}"


@pytest.mark.parametrize("prop,W,nwin,prec", [("fresnel", 1024, 1, "fp32"), ("asm", 512, 2, "fp32"),
                                               ("fresnel", 512, 1, "fp16_split"),
                                               ("asm", 2048, 1, "fp32"), ("fresnel", 2048, 1, "fp32")])
def test_two_lanes_match_one_lane(gpu, prop, W, nwin, prec):
    """batch_lanes = 2: window groups alternate between the caller's stream and the plan's own
    stream (forked from and joined into the caller's).  Every window must equal the one-lane
    result bit for bit once the caller's stream is synchronised, on a non-default caller stream
    too; at 2048 wide the time DFT runs stand-alone, each lane with its own natural-layout
    planes.  A Flatfield plan (shared scratch images) and a Filter2D plan (shared tables) keep
    one lane."""
    from holoflow_amd import hfx

    T, S, F = 64, 16, 200
    p = O.PathParams(64, W, T, S, 11, 26, prop, 852e-9, 12e-6, 12e-6, 0.45, fftshift=True)
    ring = _frames(64, W, F)
    plans = []
    for lanes in (1, 2):
        cfg = _cfg(p, prec)
        cfg.batch_windows = nwin
        cfg.batch_lanes = lanes
        plans.append(hfx.MomentPlan(cfg))
    one, two = plans
    assert (one.batch_lanes, two.batch_lanes) == (1, 2)
    assert two.workspace_bytes > one.workspace_bytes
    first = F - 40
    ref = one.moments_batch(ring, first, 7)
    assert torch.equal(two.moments_batch(ring, first, 7), ref)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        out = two.moments_batch(ring, first, 7, stream=side)
    torch.cuda.current_stream().wait_stream(side)
    assert torch.equal(out, ref)
    # one window: no fork; a partial second group: both lanes used once
    assert torch.equal(two.moments_batch(ring, first, 1)[0], ref[0])
    assert torch.equal(two.moments_batch(ring, first, nwin + 1), ref[: nwin + 1])

    # library default: two lanes
    assert hfx.MomentPlan(_cfg(p, prec)).batch_lanes == 2
    cfg = _cfg(p, prec)
    cfg.flatfield = (3.0, 3.0)
    ff_default = hfx.MomentPlan(cfg)
    assert ff_default.batch_lanes == 1
    cfg.batch_lanes = 1
    assert ff_default.workspace_bytes == hfx.MomentPlan(cfg).workspace_bytes
    if prop == "fresnel":
        cfg = _cfg(p, prec)
        cfg.filt = (4, 20, 2, 4)
        assert hfx.MomentPlan(cfg).batch_lanes == 1


def test_two_lane_batch_in_hip_graph(gpu):
    """hfx_moments_batch with two lanes under HIP graph capture: the fork to the plan's stream
    and the join back become graph edges; every replay (with eager work and synchronisation
    between replays) must equal the eager batch bit for bit."""
    from holoflow_amd import hfx

    T, S, F = 64, 16, 160
    p = O.PathParams(64, 512, T, S, 11, 26, "fresnel", 852e-9, 12e-6, 12e-6, 0.45, fftshift=True)
    ring = _frames(64, 512, F)
    plan = hfx.MomentPlan(_cfg(p))
    assert plan.batch_lanes == 2
    first = F - 24
    ref = plan.moments_batch(ring, first, 5)
    out = torch.empty_like(ref)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.moments_batch(ring, first, 5, out, stream=side)  # warm-up off the default stream
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        plan.moments_batch(ring, first, 5, out)
    for _ in range(3):
        out.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref)


@pytest.mark.parametrize("prop,W", [("fresnel", 512), ("asm", 64)])
def test_stages_compose_to_moments(gpu, prop, W):
    # stage 0 then stage 1 is hfx_moments (fused front kernel at 512, unfused at 64)
    from holoflow_amd import hfx

    p = O.PathParams(64, W, 32, 8, 6, 13, prop, 852e-9, 12e-6, 12e-6, 0.3)
    ring = _frames(64, W, 48)
    plan = hfx.MomentPlan(_cfg(p))
    ref = plan.moments(ring, 8)
    out = plan.new_output()
    plan.run_stage(0, ring, 8)
    plan.run_stage(1, None, 0, out)
    assert torch.equal(out, ref)
    from holoflow_amd._lib import HfxInvalidArgument
    with pytest.raises(HfxInvalidArgument):
        plan.run_stage(2, ring, 0, out)


def test_update_z_matches_fresh_plan(gpu):
    from holoflow_amd import hfx

    p = O.PathParams(64, 64, 32, 32, 6, 13, "asm", 852e-9, 12e-6, 12e-6, 0.2)
    ring = _frames(64, 64, 32)
    plan = hfx.MomentPlan(_cfg(p))
    plan.update_z(0.35)
    p.z_m = 0.35
    fresh = hfx.MomentPlan(_cfg(p))
    assert torch.equal(plan.moments(ring, 0), fresh.moments(ring, 0))


def test_invalid_arguments_raise(gpu):
    from holoflow_amd import hfx
    from holoflow_amd._lib import HfxError, HfxInvalidArgument

    with pytest.raises(HfxInvalidArgument):
        hfx.MomentPlan(hfx.MomentConfig(64, 64, 32, 8, 6, 40))        # bin_end > T/2+1
    with pytest.raises(HfxInvalidArgument):
        hfx.MomentPlan(hfx.MomentConfig(64, 64, 32, 8, 6, 13, dy_m=13e-6))
    with pytest.raises(HfxError):
        hfx.MomentPlan(hfx.MomentConfig(64, 64, 24, 8, 2, 9, precision="fp16"))  # T % 32
    plan = hfx.MomentPlan(hfx.MomentConfig(64, 64, 32, 8, 6, 13))
    ring = _frames(64, 64, 16)
    with pytest.raises(ValueError):
        plan.moments(ring, 0)                                          # ring < window (wrapper)
    import ctypes as C

    from holoflow_amd._lib import check, lib
    out = plan.new_output()
    with pytest.raises(HfxInvalidArgument):                            # the C ABI's own check
        check(lib().hfx_moments(plan._h, C.c_void_p(ring.data_ptr()), ring.shape[0], 0,
                                C.c_void_p(out.data_ptr()), hfx._stream(None)))
    ring = _frames(64, 64, 40)
    with pytest.raises(ValueError):
        plan.moments(ring[:, :, :32].contiguous(), 0)                  # wrong frame width
    with pytest.raises(ValueError):
        plan.moments(ring.float(), 0)                                  # not 16-bit frames
    with pytest.raises(ValueError):
        plan.moments(ring, 0, torch.empty(3, 64, 32, device="cuda"))   # output too small
    with pytest.raises(ValueError):
        plan.moments_batch(ring, 0, 2, torch.empty(1, 3, 64, 64, device="cuda"))  # one window short


def _focus_frames(H, W, T, k0, lam, dx, z, amp=1800.0):
    """uint16 frames whose time-DFT bin k0 is (amp T / 2) exp(-i pi (x^2 + y^2) / (lambda z)):
    x_n = round(2048 + amp cos(2 pi k0 n / T - phi)), phi the converging wave's phase with the
    optical axis through the sensor centre; the other bins hold only quantisation noise."""
    x = (np.arange(W) - W / 2.0) * dx
    y = (np.arange(H) - H / 2.0) * dx
    phi = np.pi * (y[:, None] ** 2 + x[None, :] ** 2) / (lam * z)
    n = np.arange(T, dtype=np.float64)[:, None, None]
    v = np.rint(2048.0 + amp * np.cos(2 * np.pi * k0 * n / T - phi[None]))
    return v.astype(np.uint16)


@pytest.mark.parametrize("H,W,T,S,kb,ke,k0,z", [(64, 64, 32, 8, 6, 13, 9, 0.45),
                                                (1024, 1024, 64, 16, 11, 26, 18, 0.45)])
def test_fused_fresnel_focuses_a_converging_wave(gpu, H, W, T, S, kb, ke, k0, z):
    """The fused plan's Fresnel path (lens applied on the time-DFT output inside the front
    kernel, fftshift on the moment store) on frames whose band bin k0 is a spherical wave
    converging to the axis at distance z: the lens cancels it (fresnel_diffraction.cu:326-341,
    CUFFT_FORWARD at :427; Goodman eq. 4-17), so M0 is one bright pixel at the fftshifted
    zero frequency (H/2, W/2) worth (amp T/2) H W / nb.  A sign slip or a centre off by a pixel
    in the lens would spread or move it."""
    from holoflow_amd import hfx

    lam, dx = 852e-9, 12e-6
    amp = 1800.0
    win = _focus_frames(H, W, T, k0, lam, dx, z, amp)
    p = O.PathParams(H, W, T, S, kb, ke, "fresnel", lam, dx, dx, z, fftshift=True)
    plan = hfx.MomentPlan(_cfg(p))
    ring = torch.from_numpy(win.view(np.int16)).cuda()
    out = plan.moments(ring, 0).cpu().numpy().astype(np.float64)
    m0 = out[0]
    nb = ke - kb
    assert np.unravel_index(np.argmax(m0), m0.shape) == (H // 2, W // 2)
    np.testing.assert_allclose(m0[H // 2, W // 2], amp * T / 2 * H * W / nb, rtol=1e-3)
    assert m0.max() / m0.sum() > 0.5
    ref = O.doppler_moments(win, p, "f64")
    errs = [rel_l2(out[i], ref[i]) for i in range(3)]
    assert max(errs) < TOL_FP32, errs


@pytest.mark.parametrize("W,T,kb,ke,precision,tol", [(512, 32, 20, 32, "fp32", TOL_FP32),       # fused, folded
                                                      (64, 32, 20, 32, "fp32", TOL_FP32),        # stand-alone, folded
                                                      (64, 32, 27, 32, "fp32", TOL_FP32),        # stand-alone, 6 bins
                                                      (1024, 64, 56, 64, "fp32", TOL_FP32),      # fused, T = 64
                                                      (512, 32, 20, 32, "fp16_split", 1e-5)])   # fused 16-bit
def test_both_sidebands_band_ending_at_window(gpu, W, T, kb, ke, precision, tol):
    """Two-sided path with bin_end == T (validation.cc:47-52 allows it for TimeMethod::FFT):
    the first negative-band plane is X[T - bin_end] = X[0], the DC bin, computed as the extra
    bin T whose centring offset must be restored like bin 0's."""
    p = O.PathParams(64, W, T, T // 4, kb, ke, "fresnel", 852e-9, 12e-6, 12e-6, 0.4, fftshift=True,
                     both_sidebands=True)
    F = T + T // 2
    _check(p, _frames(64, W, F), F - T // 4, precision, tol=tol,
           against=("f32", "f64") if precision == "fp32" else ("f64",))


def test_cfg2_1024_mixed_precision_full_size(gpu):
    """The precision workload at its own size: 1024^2 frames, T = 64, stride 16, bins [11, 26),
    Fresnel z = 0.45 m with fftshift, the time DFT in FP16, BF16 and FP16 hi/lo split (FP32
    accumulate), every moment image against the FP64 oracle on the same frames.  Bars: split
    <= 1e-5, FP16 <= 1e-3, BF16 <= 2e-2; the mixed path's M0 <= 1e-4 (split and FP16)."""
    from holoflow_amd import hfx

    p = O.config_params("cfg1_1024_fresnel")
    assert (p.height, p.window, p.stride, p.bin_begin, p.bin_end) == (1024, 64, 16, 11, 26)
    ring = _frames(1024, 1024, 96, radius=400)
    first = 32
    win = ring.cpu().numpy().view(np.uint16)[first:first + p.window]
    ref = O.doppler_moments(win, p, "f64")
    for precision, tol in (("fp16_split", 1e-5), ("fp16", 1e-3), ("bf16", 2e-2)):
        plan = hfx.MomentPlan(_cfg(p, precision))
        out = plan.moments(ring, first).cpu().numpy()
        plan.close()
        errs = [rel_l2(out[i], ref[i]) for i in range(3)]
        print("cfg2", precision, errs)
        assert max(errs) < tol, (precision, errs)
        if precision != "bf16":
            assert errs[0] <= 1e-4, (precision, errs)
