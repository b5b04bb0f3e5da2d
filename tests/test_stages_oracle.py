"""Pins of the Filter2D / Flatfield oracle (oracle/stages.py) by closed forms and library
routines, on the CPU."""
import math

import numpy as np
import pytest
from scipy import ndimage

from oracle import stages as OS


def _plane_wave(H, W, ky, kx, lead=()):
    y, x = np.mgrid[0:H, 0:W]
    w = np.exp(2j * np.pi * (kx * x / W + ky * y / H))
    return np.broadcast_to(w, lead + (H, W)).copy()


# ------------------------------------------------------------------ Filter2D
@pytest.mark.parametrize("H,W", [(16, 16), (24, 32), (27, 9)])
def test_filter2d_all_pass_and_all_stop(H, W):
    rng = np.random.default_rng(0)
    z = rng.standard_normal((2, H, W)) + 1j * rng.standard_normal((2, H, W))
    xr = rng.standard_normal((2, H, W))
    big = H + W                                   # beyond every distance on the grid
    np.testing.assert_allclose(OS.filter2d(z, 0, big, 0, 0), z * (1 + 1j) * H * W, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(OS.filter2d(xr, 0, big, 0, 0), xr * H * W, rtol=1e-12, atol=1e-9)
    assert np.abs(OS.filter2d(z, 0, 0, 0, 0)).max() == 0
    assert np.abs(OS.filter2d(xr, 0, 0, 0, 0)).max() < 1e-9
    # inner radius covering everything stops everything too
    assert np.abs(OS.filter2d(xr, big, big, 0, 0)).max() < 1e-9


@pytest.mark.parametrize("ky,kx", [(0, 0), (3, 1), (-2, 5), (7, -7), (8, 0), (0, -8)])
def test_filter2d_plane_waves_pick_table_entries(ky, kx):
    H, W = 16, 16
    r_in, r_out, s_in, s_out = 2, 5, 1, 3
    d = math.hypot(kx, ky)

    def m_of(dist):
        a = 1.0 if dist < r_out else (math.cos((dist - r_out) / s_out * math.pi / 2) if dist < r_out + s_out else 0.0)
        b = 1.0 if dist < r_in else (math.cos((dist - r_in) / s_in * math.pi / 2) if dist < r_in + s_in else 0.0)
        return a * (1 - b)

    # complex: the frequency (ky, kx) sits at table index ((ky mod H), (kx mod W)); after the
    # corner exchange its centred distance is |(k mod N) - N/2 shifted|, i.e. the signed
    # frequency with N/2 mapped to -N/2
    sy = ((ky % H) + H // 2) % H - H // 2
    sx = ((kx % W) + W // 2) % W - W // 2
    dc = math.hypot(sx, sy)
    w = _plane_wave(H, W, ky, kx)
    np.testing.assert_allclose(OS.filter2d(w, r_in, r_out, s_in, s_out), w * m_of(dc) * (1 + 1j) * H * W,
                               atol=1e-9)
    # real: a cosine keeps its frequency and is scaled by the gain at (|kx|, |ky|)
    c = w.real
    np.testing.assert_allclose(OS.filter2d(c, r_in, r_out, s_in, s_out), c * m_of(d) * H * W, atol=1e-9)


def test_filter2d_taper_value_by_hand():
    # distance 6 in the outer taper (r_out 5, s_out 4): cos(1/4 * pi/2); inner disk 0
    t = OS.filter2d_table(32, 32, 0, 5, 0, 4, False)
    assert t[6, 0] == pytest.approx(math.cos(0.25 * math.pi / 2), abs=1e-15)
    assert t[0, 6] == pytest.approx(math.cos(0.25 * math.pi / 2), abs=1e-15)
    assert t[32 - 6, 0] == t[6, 0]                # signed ky
    assert t[0, 0] == 1.0 and t[9, 0] == 0.0


# ------------------------------------------------------------------ Flatfield
def test_gaussian_taps_rule():
    for sigma, radius in [(1.0, 3), (2.1, 7), (2.5, 8), (10.33, 31), (10.34, 32)]:
        k = OS.gaussian_taps(sigma)
        assert len(k) == 2 * radius + 1
        assert k.dtype == np.float32 and abs(float(k.sum()) - 1) < 1e-6
        np.testing.assert_array_equal(k, k[::-1])
    assert not OS.flatfield_uses_fft(64, 64, 10.33, 2.0)
    assert OS.flatfield_uses_fft(64, 64, 2.0, 10.34)
    assert not OS.flatfield_uses_fft(1, 64, 50.0, 50.0)


@pytest.mark.parametrize("shape,sy,sx", [((2, 20, 30), 1.5, 2.5), ((9, 40), 3.0, 1.0), ((1, 50), 20.0, 20.0)])
def test_flatfield_spatial_matches_scipy_nearest(shape, sy, sx):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape)
    ky, kx = OS.gaussian_taps(sy).astype(np.float64), OS.gaussian_taps(sx).astype(np.float64)
    blur = ndimage.correlate1d(ndimage.correlate1d(x, kx, axis=-1, mode="nearest"), ky, axis=-2, mode="nearest")
    np.testing.assert_allclose(OS.flatfield(x, sy, sx), x - blur, atol=1e-12)


def test_flatfield_spatial_constants_and_ramps():
    # the FP32 taps sum to 1 only to FP32 precision: a constant c leaves c (1 - sum_y sum_x)
    k = OS.gaussian_taps(2.0).astype(np.float64)
    x = np.full((12, 17), 3.25)
    np.testing.assert_allclose(OS.flatfield(x, 2.0, 2.0), 3.25 * (1 - k.sum() ** 2), atol=1e-12)
    y, xx = np.mgrid[0:40, 0:40]
    ramp = 0.5 * xx + 0.25 * y
    out = OS.flatfield(ramp, 2.0, 2.0) - ramp * (1 - k.sum() ** 2)
    r = len(k) // 2
    assert np.abs(out[r:-r, r:-r]).max() < 1e-9    # symmetric taps preserve an interior ramp


@pytest.mark.parametrize("H,W,ky,kx", [(64, 64, 1, 2), (64, 128, 3, 0), (96, 64, -2, 5)])
def test_flatfield_fourier_closed_form(H, W, ky, kx):
    sy, sx = 12.0, 15.5
    assert OS.flatfield_uses_fft(H, W, sy, sx)
    assert np.abs(OS.flatfield(np.full((H, W), 7.0), sy, sx)).max() < 1e-9
    c = _plane_wave(H, W, ky, kx).real
    scale = float("%f" % float(np.float32(1) / np.float32(H * W)))
    fy, fx = ky / H, kx / W
    g = (1 - math.exp(-2 * math.pi ** 2 * (sx * sx * fx * fx + sy * sy * fy * fy))) * scale * H * W
    np.testing.assert_allclose(OS.flatfield(c, sy, sx), c * g, atol=1e-9)
