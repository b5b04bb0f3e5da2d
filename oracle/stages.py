"""Oracle: the optional in-path stages Filter2D and Flatfield, in numpy float64.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of libhfx's Filter2D and
Flatfield; the product path never imports it.

Filter2D (src/holotask/src/syncs/filter2d.cu:165-497):
  * table: the annulus a(1 - b), a = 1 inside r_outer, cos((d - r_outer)/s_outer * pi/2) over
    the next s_outer pixels, 0 beyond; b likewise for r_inner/s_inner (filter2d.cu:165-240);
  * CF32 input: d from the centre (W/2.0, H/2.0), table stored after swap_corners
    (filter2d.cu:242-305), value a(1-b) in both the real and imaginary part;
    out = IFFT2_unnormalised(FFT2(x) * F);
  * F32 input: half table [H][W//2+1] at (kx, signed ky) with ky = y if y <= H//2 else y - H;
    out = C2R_unnormalised(R2C(x) * F)  (filter2d.cu:421-468).
Flatfield (src/holotask/src/syncs/flatfield.cu:73-451):
  * taps: radius ceil(3 sigma), exp(-i^2 / (2 sigma^2)) normalised, in FP32 (:73-94);
  * spatial (< 65 taps, or a plane one pixel wide or tall): horizontal then vertical blur
    with clamp-at-edge indexing, out = x - blur (:189-240);
  * Fourier (>= 65 taps): out = C2R_unnormalised(R2C(x) * gain), gain = (1 - exp(-2 pi^2
    (sx^2 fx^2 + sy^2 fy^2))) * scale, fx = kx / W, fy = signed ky / H, where sigma and
    scale = 1/(W H) enter as "%f"-formatted text in the generated callback (:143-187).

Pinning (tests/test_stages_oracle.py): all-pass / all-stop annuli reduce Filter2D to a scaled
identity and to zero; plane waves pick out single table entries (a transposed, shifted or
mirrored table fails); the taper value at a chosen distance is computed by hand; the
spatial Flatfield matches scipy.ndimage.correlate1d with mode "nearest" and annihilates
constants and interior ramps; the Fourier Flatfield annihilates constants and scales a
cosine by its closed-form gain.
"""
from __future__ import annotations

import math

import numpy as np


def _taper(d: np.ndarray, r: int, s: int) -> np.ndarray:
    out = np.zeros(d.shape, np.float64)
    out[d < r] = 1.0
    if s > 0:
        m = (d >= r) & (d < r + s)
        out[m] = np.cos((d[m] - r) / s * math.pi / 2)
    return out


def filter2d_table(H: int, W: int, r_inner: int, r_outer: int, s_inner: int, s_outer: int,
                   complex_input: bool) -> np.ndarray:
    """The reference's filter table: complex [H, W] (CF32 input) or real [H, W//2+1] (F32)."""
    if complex_input:
        y, x = np.mgrid[0:H, 0:W].astype(np.float64)
        d = np.hypot(x - W / 2.0, y - H / 2.0)
        m = _taper(d, r_outer, s_outer) * (1.0 - _taper(d, r_inner, s_inner))
        # swap_corners: exchange the quadrants of the 2*(W//2) x 2*(H//2) block
        hw, hh = W // 2, H // 2
        sw = m.copy()
        blk = m[:2 * hh, :2 * hw]
        sw[:2 * hh, :2 * hw] = np.roll(np.roll(blk, -hh, axis=0), -hw, axis=1) if hh and hw else blk
        return sw * (1 + 1j)
    FW = W // 2 + 1
    y, x = np.mgrid[0:H, 0:FW].astype(np.float64)
    ky = np.where(y <= H // 2, y, y - H)
    d = np.hypot(x, ky)
    return _taper(d, r_outer, s_outer) * (1.0 - _taper(d, r_inner, s_inner))


def filter2d(x: np.ndarray, r_inner: int, r_outer: int, s_inner: int, s_outer: int) -> np.ndarray:
    """Filter2D over the last two axes of a complex or real array."""
    H, W = x.shape[-2:]
    if np.iscomplexobj(x):
        F = filter2d_table(H, W, r_inner, r_outer, s_inner, s_outer, True)
        return np.fft.ifft2(np.fft.fft2(x.astype(np.complex128)) * F) * (H * W)
    F = filter2d_table(H, W, r_inner, r_outer, s_inner, s_outer, False)
    return np.fft.irfft2(np.fft.rfft2(x.astype(np.float64)) * F, s=(H, W)) * (H * W)


def gaussian_taps(sigma: float) -> np.ndarray:
    """Normalised FP32 taps of radius ceil(3 sigma), summed in ascending order."""
    sigma = np.float32(sigma)
    radius = int(math.ceil(np.float32(3.0) * sigma))
    if radius <= 0:
        raise ValueError("sigma is too small to form a Gaussian kernel")
    inv = np.float32(1.0) / (np.float32(2.0) * sigma * sigma)
    k = np.array([np.exp(-np.float32(i * i) * inv) for i in range(-radius, radius + 1)], np.float32)
    s = np.float32(0.0)
    for v in k:
        s = np.float32(s + v)
    return (k / s).astype(np.float32)


def flatfield_uses_fft(H: int, W: int, sigma_y: float, sigma_x: float) -> bool:
    return H > 1 and W > 1 and max(len(gaussian_taps(sigma_y)), len(gaussian_taps(sigma_x))) >= 65


def six_decimals(v: float) -> float:
    """A value as "%f" prints it, read back as FP32."""
    return float(np.float32(float("%f" % float(np.float32(v)))))


def flatfield(x: np.ndarray, sigma_y: float, sigma_x: float) -> np.ndarray:
    """Flatfield over the last two axes of a real array."""
    H, W = x.shape[-2:]
    x = x.astype(np.float64)
    if flatfield_uses_fft(H, W, sigma_y, sigma_x):
        sy, sx = six_decimals(sigma_y), six_decimals(sigma_x)
        scale = six_decimals(np.float32(1.0) / np.float32(W * H))
        FW = W // 2 + 1
        y, k = np.mgrid[0:H, 0:FW].astype(np.float64)
        fx = k / W
        fy = np.where(y <= H // 2, y, y - H) / H
        gain = (1.0 - np.exp(-2 * math.pi ** 2 * (sx * sx * fx * fx + sy * sy * fy * fy))) * scale
        return np.fft.irfft2(np.fft.rfft2(x) * gain, s=(H, W)) * (H * W)
    ky, kx = gaussian_taps(sigma_y).astype(np.float64), gaussian_taps(sigma_x).astype(np.float64)
    ry, rx = len(ky) // 2, len(kx) // 2
    lead = [(0, 0)] * (x.ndim - 2)
    xp = np.pad(x, lead + [(0, 0), (rx, rx)], mode="edge")
    hor = sum(kx[i] * xp[..., :, i:i + W] for i in range(2 * rx + 1))
    hp = np.pad(hor, lead + [(ry, ry), (0, 0)], mode="edge")
    blur = sum(ky[i] * hp[..., i:i + H, :] for i in range(2 * ry + 1))
    return x - blur
