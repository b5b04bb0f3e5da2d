"""Filter2D and Flatfield on the GPU through the node ABI, against the reference's test
vectors (filter2d_test.cu, flatfield_test.cu) and the float64 oracle (oracle/stages.py).

Bars: the reference's KATs at their own tolerances; FP32 rel-L2 <= 1e-5 against the
oracle elsewhere."""
import ctypes as C

import numpy as np
import pytest
import torch

from holoflow_amd import hfx
from holoflow_amd._lib import check, lib
from holoflow_amd.nodes import Node, desc, desc_of
from oracle import oracle as O
from oracle import stages as OS
from oracle.synth import SynthParams, synth_frames
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

ALL_PASS = {"r_inner": 0, "r_outer": 100, "s_inner": 0, "s_outer": 0}


def _run(kind, settings, x):
    n = Node(kind, settings, [desc_of(x)])
    out = n.execute(x)[0]
    torch.cuda.synchronize()
    return out.cpu().numpy()


# ------------------------------------------------------------------ reference KATs
def test_filter2d_kat_real_all_pass(gpu):
    # filter2d_test.cu RealInputUsesRealOutput: [1, 2, 3, 4] -> [4, 8, 12, 16]
    x = torch.tensor([[1.0, 2.0], [3.0, 4.0]], device="cuda")
    np.testing.assert_allclose(_run("Filter2D", ALL_PASS, x), [[4, 8], [12, 16]], atol=1e-3)


def test_filter2d_complex_all_pass_is_scaled_identity(gpu):
    rng = np.random.default_rng(3)
    z = (rng.standard_normal((2, 16, 24)) + 1j * rng.standard_normal((2, 16, 24))).astype(np.complex64)
    out = _run("Filter2D", ALL_PASS, torch.from_numpy(z).cuda())
    assert rel_l2(out, z * (1 + 1j) * 16 * 24) < 1e-6


def test_flatfield_kats(gpu):
    # flatfield_test.cu: constant input subtracts to zero (spatial and Fourier forms), and the
    # anisotropic 3 x 4 case against the CPU restatement
    out = _run("Flatfield", {"sigma_y": 1.0, "sigma_x": 1.0}, torch.full((2, 2), 5.0, device="cuda"))
    np.testing.assert_allclose(out, 0.0, atol=1e-4)
    out = _run("Flatfield", {"sigma_y": 12.0, "sigma_x": 12.0}, torch.full((8, 8), 7.0, device="cuda"))
    np.testing.assert_allclose(out, 0.0, atol=1e-3)
    x = np.arange(12, dtype=np.float32).reshape(3, 4)
    out = _run("Flatfield", {"sigma_y": 0.75, "sigma_x": 1.25}, torch.from_numpy(x).cuda())
    np.testing.assert_allclose(out, OS.flatfield(x, 0.75, 1.25), atol=1e-4)


# ------------------------------------------------------------------ parity with the oracle
@pytest.mark.parametrize("shape", [(2, 64, 64), (3, 96, 128), (2, 48, 24), (2, 1024, 1024), (1, 2048, 256)])
@pytest.mark.parametrize("cplx", [True, False])
def test_filter2d_vs_oracle(gpu, shape, cplx):
    rng = np.random.default_rng(sum(shape))
    st = {"r_inner": 3, "r_outer": min(shape[1:]) // 4, "s_inner": 2, "s_outer": 5}
    if cplx:
        x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    else:
        x = rng.standard_normal(shape).astype(np.float32)
    out = _run("Filter2D", st, torch.from_numpy(x).cuda())
    ref = OS.filter2d(x, st["r_inner"], st["r_outer"], st["s_inner"], st["s_outer"])
    assert out.dtype == x.dtype
    assert rel_l2(out, ref) < 1e-5


@pytest.mark.parametrize("shape,sy,sx,fft", [
    ((3, 100, 70), 2.0, 3.5, False),            # spatial, fused LDS tile, ragged edges
    ((2, 1024, 1024), 5.0, 5.0, False),
    ((2, 1, 300), 20.0, 20.0, False),           # one-pixel-tall planes: radius 60, two passes
    ((2, 256, 256), 15.0, 11.0, True),          # Fourier form
    ((1, 1024, 1024), 40.0, 40.0, True),
])
def test_flatfield_vs_oracle(gpu, shape, sy, sx, fft):
    assert OS.flatfield_uses_fft(shape[1], shape[2], sy, sx) == fft
    h = C.c_void_p()
    flag = C.c_int32(-1)
    check(lib().hfx_flatfield_create(shape[0], shape[1], shape[2], sy, sx, None, C.byref(h), C.byref(flag)))
    lib().hfx_flatfield_destroy(h)
    assert flag.value == int(fft)
    rng = np.random.default_rng(int(sy * 10))
    x = (rng.standard_normal(shape) + 3.0).astype(np.float32)
    out = _run("Flatfield", {"sigma_y": sy, "sigma_x": sx}, torch.from_numpy(x).cuda())
    assert rel_l2(out, OS.flatfield(x, sy, sx)) < 1e-5


def test_stage_update_reuse(gpu):
    d = desc_of(torch.empty(2, 64, 64, device="cuda"))
    f = Node("Filter2D", ALL_PASS, [d])
    assert f.update(ALL_PASS)
    assert not f.update(dict(ALL_PASS, r_outer=10))
    g = Node("Flatfield", {"sigma_y": 2.0, "sigma_x": 2.0}, [d])
    assert g.update({"sigma_y": 2.0, "sigma_x": 2.0})
    assert not g.update({"sigma_y": 2.5, "sigma_x": 2.0})
    x = torch.randn(2, 64, 64, device="cuda")
    np.testing.assert_allclose(g.execute(x)[0].cpu().numpy(), OS.flatfield(x.cpu().numpy(), 2.5, 2.0),
                               atol=1e-5)


# ------------------------------------------------------------------ inside the fused plan


@pytest.mark.parametrize("prop,filt,ff", [
    ("fresnel", (3, 20, 2, 5), None),           # Filter2D before the Fresnel lens
    ("fresnel", None, (2.0, 3.0)),              # spatial flatfield on the moment images
    ("fresnel", (0, 25, 0, 4), (15.0, 12.0)),   # both; Fourier flatfield
    ("asm", (10, 25, 3, 3), (2.5, 2.5)),        # angular-spectrum filter + flatfield
])
def test_fused_plan_with_optional_stages(gpu, prop, filt, ff):
    T, H, W, S = 32, 128, 128, 8
    kb, ke = O.band_bins(6e3, 15e3, 37e3, T)
    p = O.PathParams(H, W, T, S, kb, ke, prop, z_m=0.3, fftshift=True, filt=filt, flatfield=ff)
    frames, _ = synth_frames(SynthParams(H, W, 30.0, seed=4, stream_id=3), 0, T)
    ref = O.doppler_moments(frames, p, "f64")
    plan = hfx.MomentPlan(hfx.MomentConfig(H, W, T, S, kb, ke, prop, z_m=0.3, fftshift=True, filt=filt,
                                           flatfield=ff))
    ring = torch.from_numpy(frames.view(np.int16)).cuda()
    out = plan.moments(ring, first_frame=0)
    torch.cuda.synchronize()
    out = out.cpu().numpy()
    for m in range(3):
        assert rel_l2(out[m], ref[m]) < 1e-5, f"moment {m}"


def test_moments_node_flatfield_by_cutoff_period(gpu):
    T, H, W = 32, 128, 128
    kb, ke = O.band_bins(6e3, 15e3, 37e3, T)
    lam, dx, z, period = 852e-9, 12e-6, 0.3, 2.67e-3
    f = np.float32
    pitch = f(lam) * abs(f(z)) / (f(H) * f(dx))          # Fresnel image pitch, FP32
    sigma = float(f(0.187) * f(period) / pitch)
    p = O.PathParams(H, W, T, T, kb, ke, "fresnel", lambda_m=float(f(lam)), dx_m=float(f(dx)), dy_m=float(f(dx)),
                     z_m=float(f(z)), fftshift=True, flatfield=(sigma, sigma))
    frames, _ = synth_frames(SynthParams(H, W, 30.0, seed=6, stream_id=1), 0, T)
    n = Node("DopplerMoments", {"lambda": lam, "dx": dx, "z": z, "time_z_begin": kb, "time_z_end": ke,
                                "prop": "fresnel", "fftshift": True, "flatfield": {"cutoff_period_m": period}},
             [desc([T, H, W], "U16")])
    out = n.execute(torch.from_numpy(frames.view(np.int16)).cuda())[0]
    torch.cuda.synchronize()
    ref = O.doppler_moments(frames, p, "f64")
    assert rel_l2(out.cpu().numpy(), ref) < 1e-5
    # z moves the Fresnel image pitch and with it the flatfield sigma: the task is rebuilt
    assert not n.update({"lambda": lam, "dx": dx, "z": 0.35, "time_z_begin": kb, "time_z_end": ke,
                         "prop": "fresnel", "fftshift": True, "flatfield": {"cutoff_period_m": period}})
