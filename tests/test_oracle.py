"""Pins the oracle (CPU restatement) before it is trusted as the parity checker.

Each check is chosen so that a plausible mistake (dropped term, wrong sign, wrong index,
transposed operand, wrong scale) fails: reference known-answer tests, brute-force DFTs,
closed forms of the Fresnel lens (discrete Gauss sum) and of the angular spectrum
(plane-wave eigenfunctions, z = 0, reciprocity), tone inputs for the moment weights, and
Random123's Philox known-answer vectors.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle.synth import SynthParams, philox4x32_10, synth_frames
from tests.conftest import rel_l2


def brute_dft(x, axis=0, sign=-1):
    x = np.moveaxis(np.asarray(x, dtype=np.complex128), axis, 0)
    n = x.shape[0]
    out = np.zeros_like(x)
    for k in range(n):
        for j in range(n):
            out[k] += x[j] * complex(math.cos(2 * math.pi * k * j / n), sign * math.sin(2 * math.pi * k * j / n))
    return np.moveaxis(out, 0, axis)


# ----------------------------------------------------------------------- time DFT
def test_fft_reference_kat():
    # fft_test.cu:86 inputs; closed forms of the 4-point DFT
    x = np.array([[1, 2, 3, 4], [0, -1, 2, -3]], dtype=np.float64)
    X = torch.fft.fft(torch.as_tensor(x), dim=1).numpy()
    np.testing.assert_allclose(X[0], [10, -2 + 2j, -2, -2 - 2j], atol=1e-12)
    np.testing.assert_allclose(X[1], [-2, -2 - 2j, 6, -2 + 2j], atol=1e-12)


def test_rfft_reference_kat_vs_bruteforce():
    # rfft_test.cu:86 inputs [2, 6] along axis 1
    x = np.array([[1, 2, 3, 4, 5, 6], [0, -1, 2, -3, 1, 4]], dtype=np.float64)
    p = O.PathParams(1, 2, 6, 6, 0, 4, both_sidebands=False)
    got = O.time_dft(x.T.reshape(6, 1, 2), p, "f64").numpy().reshape(4, 2).T
    ref = brute_dft(x, axis=1)[:, :4]
    np.testing.assert_allclose(got, ref, atol=1e-10)
    np.testing.assert_allclose(ref[0, 0], 21)          # DC = sum


@pytest.mark.parametrize("both", [False, True])
def test_time_dft_band_vs_bruteforce(both):
    rng = np.random.default_rng(1)
    T, H, W = 16, 3, 4
    win = rng.integers(0, 4096, size=(T, H, W)).astype(np.uint16)
    p = O.PathParams(H, W, T, T, 2, 6, both_sidebands=both)
    got = O.time_dft(win, p, "f64").numpy()
    full = brute_dft(win.astype(np.float64), axis=0)
    exp = full[2:6]
    if both:
        exp = np.concatenate([exp, full[T - 6:T - 2]])
        # negative band of real input = conjugate of bins (kb, ke]
        np.testing.assert_allclose(got[4:], np.conj(full[6:2:-1]), atol=1e-7)
    np.testing.assert_allclose(got, exp, atol=1e-7)


def test_time_dft_pure_tone():
    T, k0, A = 64, 13, 700.0
    n = np.arange(T)
    x = 2048 + A * np.cos(2 * np.pi * k0 * n / T)
    win = np.broadcast_to(x[:, None, None], (T, 2, 4)).astype(np.float64)
    p = O.PathParams(2, 4, T, T, 11, 26)
    S = O.time_dft(win, p, "f64").numpy()
    mag = np.abs(S[:, 0, 0])
    assert abs(mag[k0 - 11] - T / 2 * A) < 1e-8 * T * A
    assert np.all(np.delete(mag, k0 - 11) < 1e-8 * T * A)


# ----------------------------------------------------------------------- Fresnel
def test_fresnel_lens_gauss_sum_closed_form():
    # lambda*z = N*dx^2 makes the lens the discrete chirp exp(i*pi*(c-N/2)^2/N); for even N
    # its DFT has magnitude sqrt(N) at every bin (quadratic Gauss sum), so a constant field
    # propagates to |FFT2| == N everywhere.  A wrong scale of the phase breaks flatness.
    N, dx, lam = 16, 10e-6, 500e-9
    z = N * dx * dx / lam
    p = O.PathParams(N, N, 4, 4, 0, 1, "fresnel", lam, dx, dx, z)
    S = torch.ones((1, N, N), dtype=torch.complex128)
    P = O.fresnel(S, p, "f64").numpy()
    np.testing.assert_allclose(P, N, rtol=1e-10)
    # a doubled phase would not be flat
    p2 = O.PathParams(N, N, 4, 4, 0, 1, "fresnel", lam, dx, dx, z / 2)
    assert np.ptp(O.fresnel(S, p2, "f64").numpy()) > 1.0


def test_fresnel_far_field_and_delta():
    H, W = 8, 12
    rng = np.random.default_rng(2)
    S = torch.as_tensor(rng.normal(size=(2, H, W)) + 1j * rng.normal(size=(2, H, W)))
    p = O.PathParams(H, W, 4, 4, 0, 1, "fresnel", 852e-9, 12e-6, 12e-6, 1e12)
    np.testing.assert_allclose(O.fresnel(S, p, "f64").numpy(), np.abs(np.fft.fft2(S.numpy())), rtol=1e-8)
    # a delta anywhere propagates to a flat unit magnitude
    d = torch.zeros((1, H, W), dtype=torch.complex128)
    d[0, 3, 5] = 1.0
    p = O.PathParams(H, W, 4, 4, 0, 1, "fresnel", 852e-9, 12e-6, 12e-6, 0.2)
    np.testing.assert_allclose(O.fresnel(d, p, "f64").numpy(), 1.0, rtol=1e-12)


def _converging_wave(H, W, lam, dx, dy, z, shift=(0, 0)):
    """exp(-i pi (x^2 + y^2) / (lambda z)) on the sensor with the optical axis through its
    centre (x = 0 at column W/2, y = 0 at row H/2), optionally displaced by `shift` pixels."""
    x = (np.arange(W) - W / 2.0 - shift[1]) * dx
    y = (np.arange(H) - H / 2.0 - shift[0]) * dy
    return np.exp(-1j * math.pi * (y[:, None] ** 2 + x[None, :] ** 2) / (lam * z))


@pytest.mark.parametrize("H,W", [(16, 16), (12, 20), (64, 48)])
def test_fresnel_focuses_a_converging_wave(H, W):
    # Goodman, Introduction to Fourier Optics, eq. (4-17): the Fresnel integral is the
    # forward Fourier transform of U(x, y) exp(+i pi (x^2 + y^2) / (lambda z)), the phase the
    # reference puts on its input before a CUFFT_FORWARD transform
    # (fresnel_diffraction.cu:326-341, 427).  A spherical wave converging to a point at
    # distance z on the axis, exp(-i pi (x^2 + y^2) / (lambda z)), cancels that phase
    # exactly: the propagated field is a single bin at zero frequency (the unshifted output's
    # (0, 0)) holding H*W.  A conjugated lens leaves the doubled chirp (spread over many
    # bins); an optical axis one pixel off the sensor centre leaves a linear phase, which with
    # lambda z = S dx^2 (S = max(H, W)) moves the focus by exactly one bin along that axis.
    S = max(H, W)
    dx = 10e-6
    lam = 500e-9
    z = S * dx * dx / lam
    p = O.PathParams(H, W, 4, 4, 0, 1, "fresnel", lam, dx, dx, z)
    u = torch.as_tensor(_converging_wave(H, W, lam, dx, dx, z)[None])
    P = O.fresnel(u, p, "f64").numpy()[0]
    assert np.unravel_index(np.argmax(P), P.shape) == (0, 0)
    np.testing.assert_allclose(P[0, 0], H * W, rtol=1e-10)
    assert P.max() / P.sum() > 1 - 1e-9
    # the conjugated lens (a planted sign slip): the focus is gone
    L = O.fresnel_lens(H, W, lam, dx, dx, z)
    Pc = np.abs(np.fft.fft2(u.numpy()[0] * np.conj(L)))
    assert Pc.max() / Pc.sum() < 0.2
    # a converging wave one pixel off axis (along an axis of length S) leaves the linear phase
    # exp(+2 pi i r / S): it focuses exactly one bin further, so the centre is not free either
    for axis, n in ((0, H), (1, W)):
        if n != S:
            continue
        sh = (1, 0) if axis == 0 else (0, 1)
        v = torch.as_tensor(_converging_wave(H, W, lam, dx, dx, z, sh)[None])
        Q = O.fresnel(v, p, "f64").numpy()[0]
        assert np.unravel_index(np.argmax(Q), Q.shape) == sh
        np.testing.assert_allclose(Q[sh], H * W, rtol=1e-10)


def test_fresnel_bruteforce_and_complex_output():
    H, W = 6, 8
    rng = np.random.default_rng(3)
    S = rng.normal(size=(H, W)) + 1j * rng.normal(size=(H, W))
    p = O.PathParams(H, W, 4, 4, 0, 1, "fresnel", 852e-9, 20e-6, 20e-6, 0.05)
    L = O.fresnel_lens(H, W, p.lambda_m, p.dx_m, p.dy_m, p.z_m)
    F = brute_dft(brute_dft(S * L, axis=0), axis=1)
    got = O.fresnel(torch.as_tensor(S[None]), p, "f64", output_magnitude=False).numpy()[0]
    np.testing.assert_allclose(got, F * L, atol=1e-9)
    np.testing.assert_allclose(np.abs(got), O.fresnel(torch.as_tensor(S[None]), p, "f64").numpy()[0], rtol=1e-12)
    # the lens is unit-modulus and 1 at the optical axis (x = y = 0 at c = W/2, r = H/2)
    np.testing.assert_allclose(np.abs(L), 1.0, rtol=1e-14)
    np.testing.assert_allclose(L[H // 2, W // 2], 1.0, atol=1e-14)


# ----------------------------------------------------------------------- angular spectrum
def _asm(S, H, W, z, pad=(0, 0), filt=None, lam=852e-9, dx=20e-6):
    p = O.PathParams(H, W, 4, 4, 0, 1, "asm", lam, dx, dx, z, pad=pad, filt=filt)
    return O.angular_spectrum(torch.as_tensor(S), p, "f64").numpy()


def test_asm_reference_kats():
    # angular_spectrum_test.cu:146-252: z = 0 returns input * (H*W), padding is centred
    np.testing.assert_allclose(_asm(np.array([[[1.0, 2.0], [3.0, 4.0]]]), 2, 2, 0.0)[0],
                               [[4, 8], [12, 16]], atol=1e-12)
    exp = np.array([[0, 0, 0, 0], [0, 16, 32, 0], [0, 48, 64, 0], [0, 0, 0, 0]])
    np.testing.assert_allclose(_asm(np.array([[[1.0, 2.0], [3.0, 4.0]]]), 2, 2, 0.0, pad=(4, 4))[0], exp,
                               atol=1e-12)
    x = np.array([(k + 1) * (1 + 1j) for k in range(8)]).reshape(2, 2, 2)
    got = _asm(x, 2, 2, 0.0, pad=(4, 4))
    for b in range(2):
        np.testing.assert_allclose(got[b, 1:3, 1:3], 16 * x[b], atol=1e-12)
        got[b, 1:3, 1:3] = 0
    np.testing.assert_allclose(got, 0, atol=1e-12)


def test_asm_all_pass_filter():
    d = np.zeros((1, 4, 4), complex)
    d[0, 1, 1] = 1.0
    np.testing.assert_allclose(_asm(d, 4, 4, 0.01, filt=(0, 100, 0, 0)), _asm(d, 4, 4, 0.01), atol=1e-12)


@pytest.mark.parametrize("m,q", [(3, 0), (-5, 2), (7, -6), (0, 0)])
def test_asm_plane_wave_eigenfunction(m, q):
    # exp(2 pi i (m c / W + q r / H)) is an eigenfunction: output = H*W * Hc(u_m, v_q) * input,
    # Hc = exp(i 2 pi z / lambda sqrt(1 - lambda^2 (u^2 + v^2)))  (Goodman, eq. 3-74)
    H, W, dx, lam, z = 16, 20, 12e-6, 852e-9, 0.3
    r, c = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    u_in = np.exp(2j * np.pi * (m * c / W + q * r / H))
    u, v = m / (W * dx), q / (H * dx)
    Hc = np.exp(1j * 2 * np.pi * z / lam * np.sqrt(1 - lam ** 2 * (u ** 2 + v ** 2)))
    got = _asm(u_in[None], H, W, z, lam=lam, dx=dx)[0]
    np.testing.assert_allclose(got, H * W * Hc * u_in, rtol=1e-9, atol=1e-9 * H * W)


def test_asm_reciprocity_and_unitarity():
    H, W = 12, 16
    rng = np.random.default_rng(4)
    S = rng.normal(size=(1, H, W)) + 1j * rng.normal(size=(1, H, W))
    fwd = _asm(S, H, W, 0.25, dx=12e-6)
    back = _asm(fwd, H, W, -0.25, dx=12e-6)
    np.testing.assert_allclose(back, (H * W) ** 2 * S, rtol=1e-9, atol=1e-9 * (H * W) ** 2)
    Hs = O.asm_transfer(H, W, 852e-9, 12e-6, 12e-6, 0.25)
    np.testing.assert_allclose(np.abs(Hs), 1.0, rtol=1e-13)
    # swap_corners is ifftshift for even sizes; the DC entry holds exp(i 2 pi z / lambda)
    np.testing.assert_allclose(O.swap_corners(np.arange(H * W).reshape(H, W)),
                               np.fft.ifftshift(np.arange(H * W).reshape(H, W)))
    np.testing.assert_allclose(Hs[0, 0], np.exp(1j * 2 * np.pi * 0.25 / 852e-9), rtol=1e-12)


# ----------------------------------------------------------------------- moments & shift
def test_mean_reference_kats():
    # mean_test.cu:109-141
    a = torch.tensor([[3.0, 1.0, 2.0], [8.0, 5.0, 7.0]], dtype=torch.float64)
    np.testing.assert_allclose(a.mean(dim=1).numpy(), [2.0, 20.0 / 3.0])
    np.testing.assert_allclose(torch.tensor([2.0, 4.0, 6.0, 8.0]).mean().item(), 5.0)


def test_fftshift_reference_kat_and_odd_length():
    # fftshift_test.cu:105-121: [3, 4] along axis 1
    x = torch.arange(1, 13, dtype=torch.float32).reshape(3, 4)
    np.testing.assert_array_equal(O.fftshift_ref(x, (1,)).numpy(), [[3, 4, 1, 2], [7, 8, 5, 6], [11, 12, 9, 10]])
    # odd n: out[i] = in[(i + n//2) mod n] is numpy's ifftshift
    y = torch.arange(5)
    np.testing.assert_array_equal(O.fftshift_ref(y, (0,)).numpy(), np.fft.ifftshift(np.arange(5)))
    np.testing.assert_array_equal(O.fftshift_ref(torch.arange(6), (0,)).numpy(), np.fft.fftshift(np.arange(6)))


def test_arange_weights_and_band_bins():
    assert O.arange_len(0.0, 1.0, 0.25) == 4
    assert O.arange_len(0.0, 1.1, 0.25) == 5
    assert O.arange_len(1.0, 0.0, 0.5) == 0
    # the workloads' bands
    assert O.band_bins(6e3, 15e3, 37e3, 32) == (6, 13)
    assert O.band_bins(6e3, 15e3, 37e3, 64) == (11, 26)
    assert O.band_bins(1e3, 5e3, 10e3, 128) == (13, 65)
    assert O.band_bins(4e3, 5e3, 10e3, 128)[0] == 52
    p = O.config_params("cfg1_1024_fresnel")
    f = O.freq_weights(p)
    assert len(f) == O.arange_len(11 * 37e3 / 64, 26 * 37e3 / 64, 37e3 / 64) == 15
    np.testing.assert_array_equal(f, (np.arange(11, 26) * 578.125).astype(np.float32))
    pb = O.PathParams(8, 8, 64, 64, 11, 26, both_sidebands=True)
    np.testing.assert_allclose(O.freq_weights(pb)[15:], np.arange(11, 26) * 578.125 - 37e3)


def test_moment_weights_on_a_tone():
    # a spatially uniform tone at bin k0 concentrates |S| at the DC pixel of one plane:
    # M1/M0 = f_k0 and M2/M0 = f_k0^2 there, and zero elsewhere (Fresnel at far field)
    T, k0, H, W = 64, 17, 4, 4
    n = np.arange(T)
    x = 2048 + 900 * np.cos(2 * np.pi * k0 * n / T)
    win = np.broadcast_to(x[:, None, None], (T, H, W)).copy()
    p = O.PathParams(H, W, T, T, 11, 26, "fresnel", 852e-9, 12e-6, 12e-6, 1e15)
    M = O.doppler_moments(win, p, "f64")
    fk = 17 * 37e3 / 64
    assert abs(M[1, 0, 0] / M[0, 0, 0] - fk) < 1e-9 * fk
    # the reference squares the F32 weights in F32 (Multiply of two F32 tensors):
    # 9828.125^2 = 96591041.015625 rounds to 96591040 in F32
    fk2 = float(np.float32(fk) * np.float32(fk))
    assert fk2 != fk * fk
    assert abs(M[2, 0, 0] / M[0, 0, 0] - fk2) < 1e-9 * fk2
    np.testing.assert_allclose(M[0, 0, 0], 900 * T / 2 * H * W / 15, rtol=1e-9)
    assert np.abs(M[0]).sum() - M[0, 0, 0] < 1e-6 * M[0, 0, 0]


# ----------------------------------------------------------------------- generator
def test_philox_random123_kat():
    kat = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
           ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
           ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
            (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for ctr, key, exp in kat:
        assert tuple(int(v) for v in philox4x32_10(ctr, key)) == exp


def test_synth_statistics_and_continuation():
    sp = SynthParams(32, 32, 12.0)
    fr, phi = synth_frames(sp, 0, 8)
    assert fr.dtype == np.uint16 and fr.shape == (8, 32, 32) and fr.max() <= 4095
    assert 900 < fr[:, 0, 0].mean() < 1100          # reference beam only outside the disk
    a, pa = synth_frames(sp, 0, 4)
    b, _ = synth_frames(sp, 4, 4, pa)
    np.testing.assert_array_equal(np.concatenate([a, b]), fr)


def test_fp32_restatement_within_parity_tolerance_of_fp64():
    sp = SynthParams(64, 64, 24.0)
    win, _ = synth_frames(sp, 0, 32)
    for prop in ("fresnel", "asm"):
        p = O.PathParams(64, 64, 32, 32, 6, 13, prop, 852e-9, 12e-6, 12e-6, 0.3, fftshift=prop == "fresnel")
        m32 = O.doppler_moments(win, p, "f32")
        m64 = O.doppler_moments(win, p, "f64")
        for i in range(3):
            assert rel_l2(m32[i], m64[i]) < 2e-6, (prop, i)
