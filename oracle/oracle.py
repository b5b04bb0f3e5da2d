"""Oracle: CPU restatement of the Holoflow Doppler-moment reconstruction path.

TEST INFRASTRUCTURE ONLY.  This module is imported by tests/, by
__graft_entry__.smoke() (as the checker) and by bench.py's cpu_baseline leg (the timed
CPU baseline).  The product path (holoflow_amd) never imports it and has no CPU
fallback.

It restates, stage by stage, the node chain the reference's GraphBuilder wires
(src/holovibes/src/pipeline/graph_builder.cc:232-305, 522-566, 578-647):

    Conversion(U16 -> F32)                      conversion.cu:156-202
    RFFT along time, norm "backward"            rfft.cu:148-171, 264-297
      | FFT along time of real input            fft.cu:163-194, 331-386
    Slice [z_begin, z_end) (+ [T-z_end, T-z_begin) and Concatenate for FFT)
                                                graph_builder.cc:259-287
    FresnelDiffraction, output_magnitude        fresnel_diffraction.cu:262-341, 410-441
      | AngularSpectrum                          angular_spectrum.cu:189-387, 408-431
    MeanAbs / Mean(-3)            -> M0           mean_abs.cu:239-276, mean.cu:268-306
    Abs * Arange weights, Mean(-3) -> M1, M2     graph_builder.cc:578-631, arange.cu:61-68
    FFTShift(-2, -1) (optional)                  fftshift.cu:110-145, 254-260

The same code runs in two precisions:
  * "f32": float32 / complex64 on torch-CPU, whose FFTs are MKL's -- the reference
    FP32 semantics, and the CPU baseline that is timed;
  * "f64": float64 / complex128 -- the FP64 oracle every moment image is scored against.
All propagation phases (Fresnel lens, angular-spectrum transfer function) are generated
in float64 in both precisions and rounded once to complex64 for "f32".

Optional stages (oracle/stages.py): Filter2D on the band planes before Fresnel
(graph_builder.cc:124-126) and Flatfield on each moment image after the fftshift
(graph_builder.cc:637-641), when PathParams.filt (Fresnel) / PathParams.flatfield are set.

Pinning (tests/test_oracle.py): reference KATs (fft_test.cu:86, rfft_test.cu:86,
fftshift_test.cu:105-121, mean_test.cu:109-141, angular_spectrum_test.cu:146-283),
brute-force O(n^2) DFTs, the discrete Gauss-sum identity for the Fresnel lens, plane-wave
eigenfunctions of the angular spectrum, z=0 / reciprocity closed forms, tone inputs for
the moment weights, and Random123 known-answer vectors for Philox.  The sign and the
S = max(H, W) centring of the Fresnel lens are taken from the reference source text
(fresnel_diffraction.cu:326-341); magnitudes do not constrain them.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

_DT = {"f32": (torch.float32, torch.complex64), "f64": (torch.float64, torch.complex128)}


# ------------------------------------------------------------------------------- settings
@dataclass
class PathParams:
    """Settings of one reconstruction (the subset of Settings the path reads)."""

    height: int
    width: int
    window: int                 # time_window T
    stride: int                 # time stride S
    bin_begin: int              # time_z_begin
    bin_end: int                # time_z_end
    prop: str = "fresnel"       # "fresnel" | "asm"
    lambda_m: float = 852e-9
    dx_m: float = 12e-6
    dy_m: float = 12e-6
    z_m: float = 0.3
    weight_fs_hz: float = 37e3  # graph_builder.cc:580 hard-codes 37 kHz
    both_sidebands: bool = False
    fftshift: bool = False
    power2: bool = False
    pad: tuple = (0, 0)         # (pad_height, pad_width), 0 = none
    filt: Optional[tuple] = None  # (r_inner, r_outer, s_inner, s_outer): ASM filter, or Filter2D before Fresnel
    split_bin: int = 0
    flatfield: Optional[tuple] = None  # (sigma_y, sigma_x): Flatfield on every moment image

    @property
    def n_bins(self) -> int:
        return self.bin_end - self.bin_begin

    @property
    def out_shape(self) -> tuple:
        oh = self.pad[0] or self.height
        ow = self.pad[1] or self.width
        return oh, ow


def band_bins(f_lo_hz: float, f_hi_hz: float, fs_hz: float, window: int) -> tuple:
    """Hz band -> [z_begin, z_end): ceil(f_lo/df), floor(f_hi/df)+1, clipped to T/2+1
    (the RFFT limit of validation.cc:47-52)."""
    df = fs_hz / window
    kb = int(math.ceil(f_lo_hz / df - 1e-12))
    ke = int(math.floor(f_hi_hz / df + 1e-12)) + 1
    return kb, min(ke, window // 2 + 1)


def arange_len(start: float, stop: float, step: float) -> int:
    """Length rule of holonp::Arange (arange.cu:78-99): ceil((stop-start)/step)."""
    forward = step > 0.0
    if (forward and stop <= start) or ((not forward) and stop >= start):
        return 0
    n = math.ceil((stop - start) / step)
    if not math.isfinite(n) or n <= 0:
        raise ValueError("Arange: invalid resulting length")
    return int(n)


def freq_weights(p: PathParams) -> np.ndarray:
    """F32 frequency weights f_k (graph_builder.cc:578-601): Arange(f0, f1, df) values
    start + step*i computed in double and cast to F32 (arange.cu:61-68); the FFT path
    appends Arange(f0 - fs, f1 - fs, df)."""
    fs = p.weight_fs_hz
    df = fs / p.window
    f0 = p.bin_begin * df
    n = p.n_bins
    w = [np.float32(f0 + df * i) for i in range(n)]
    if p.both_sidebands:
        w += [np.float32((f0 - fs) + df * i) for i in range(n)]
    return np.array(w, dtype=np.float32)


# ------------------------------------------------------------------- FP64 phase tables
def fresnel_lens(H: int, W: int, lambda_m: float, dx: float, dy: float, z: float) -> np.ndarray:
    """Quadratic lens of fresnel_diffraction.cu:326-341 in FP64:
    L[r,c] = exp(i*pi/(lambda*z)*(x^2+y^2)), x = ((c+(S-W)//2) - S/2)*dx,
    y = ((r+(S-H)//2) - S/2)*dy, S = max(H, W) (S/2 a real division)."""
    S = max(H, W)
    c = np.arange(W, dtype=np.float64)
    r = np.arange(H, dtype=np.float64)
    x = ((c + (S - W) // 2) - S / 2.0) * dx
    y = ((r + (S - H) // 2) - S / 2.0) * dy
    a = math.pi / (lambda_m * z)
    return np.exp(1j * a * (y[:, None] ** 2 + x[None, :] ** 2))


def _filter_weight(H: int, W: int, filt: tuple) -> np.ndarray:
    """Annular filter of angular_spectrum.cu:280-330 (value squared), centred grid."""
    r_in, r_out, s_in, s_out = (float(v) for v in filt)
    xx = np.arange(W, dtype=np.float64)[None, :] - W / 2.0
    yy = np.arange(H, dtype=np.float64)[:, None] - H / 2.0
    dist = np.hypot(xx, yy)
    i0, i1, o0, o1 = r_in - s_in, r_in + s_in, r_out - s_out, r_out + s_out
    val = np.zeros_like(dist)
    with np.errstate(divide="ignore", invalid="ignore"):
        fade_in = 0.5 * (1.0 + np.sin(np.pi / (2.0 * s_in) * (dist - i0))) if s_in else 0.0
        fade_out = 0.5 * (1.0 + np.sin(np.pi / (2.0 * s_out) * (o1 - dist))) if s_out else 0.0
    val = np.where((dist >= i0) & (dist < i1), fade_in, val)
    val = np.where((dist >= i1) & (dist < o0), 1.0, val)
    val = np.where((dist >= o0) & (dist < o1), fade_out, val)
    return val * val


def swap_corners(a: np.ndarray) -> np.ndarray:
    """Quadrant exchange of angular_spectrum.cu:332-363 on the last two axes (equals
    numpy.fft.ifftshift for even sizes; odd sizes leave the last row/column alone)."""
    H, W = a.shape[-2:]
    hh, hw = H // 2, W // 2
    out = a.copy()
    tl = a[..., :hh, :hw].copy()
    br = a[..., hh:2 * hh, hw:2 * hw].copy()
    tr = a[..., :hh, hw:2 * hw].copy()
    bl = a[..., hh:2 * hh, :hw].copy()
    out[..., :hh, :hw] = br
    out[..., hh:2 * hh, hw:2 * hw] = tl
    out[..., :hh, hw:2 * hw] = bl
    out[..., hh:2 * hh, :hw] = tr
    return out


def asm_transfer(H: int, W: int, lambda_m: float, dx: float, dy: float, z: float,
                 filt: Optional[tuple] = None) -> np.ndarray:
    """Angular-spectrum transfer function of angular_spectrum.cu:261-387 in FP64:
    centred grid u = (c - W//2)/(W dx), v = (r - H//2)/(H dy),
    Hc = exp(i 2 pi z/lambda sqrt(max(0, 1 - lambda^2 (u^2+v^2)))), times the filter
    weight, then swap_corners."""
    u = (np.arange(W) - W // 2) / (W * dx)
    v = (np.arange(H) - H // 2) / (H * dy)
    tmp = 1.0 - lambda_m ** 2 * (v[:, None] ** 2 + u[None, :] ** 2)
    tmp = np.maximum(tmp, 0.0)
    Hc = np.exp(1j * (2.0 * math.pi * z / lambda_m) * np.sqrt(tmp))
    if filt is not None:
        Hc = Hc * _filter_weight(H, W, filt)
    return swap_corners(Hc)


# ------------------------------------------------------------------------------ stages
def _t(a, dtype):
    return torch.as_tensor(a).to(dtype)


def time_dft(window: np.ndarray, p: PathParams, prec: str = "f64") -> torch.Tensor:
    """Conversion -> RFFT (or FFT) along axis 0 -> band slice(s): [planes][H][W]."""
    rdt, cdt = _DT[prec]
    x = _t(window, rdt)                                  # exact for 12-bit data
    if not p.both_sidebands:
        X = torch.fft.rfft(x, dim=0)                     # norm "backward": unscaled
        return X[p.bin_begin:p.bin_end].contiguous()
    X = torch.fft.fft(x.to(cdt), dim=0)
    T = p.window
    pos = X[p.bin_begin:p.bin_end]
    neg = X[T - p.bin_end:T - p.bin_begin]
    return torch.cat([pos, neg], dim=0).contiguous()


def fresnel(S: torch.Tensor, p: PathParams, prec: str = "f64", output_magnitude: bool = True,
            skip_phase_shift: bool = False) -> torch.Tensor:
    """|FFT2(S * L)| (output_magnitude) or FFT2(S * L) * L (fresnel_diffraction.cu:359-380)."""
    _, cdt = _DT[prec]
    L = _t(fresnel_lens(p.height, p.width, p.lambda_m, p.dx_m, p.dy_m, p.z_m), cdt)
    F = torch.fft.fft2(S.to(cdt) * L)
    if output_magnitude:
        return torch.abs(F)
    return F if skip_phase_shift else F * L


def angular_spectrum(S: torch.Tensor, p: PathParams, prec: str = "f64") -> torch.Tensor:
    """IFFT2_unnormalised(FFT2(pad(S)) * swap_corners(H)) (angular_spectrum.cu:408-431)."""
    _, cdt = _DT[prec]
    OH, OW = p.out_shape
    S = S.to(cdt)
    if (OH, OW) != (p.height, p.width):
        oy, ox = (OH - p.height) // 2, (OW - p.width) // 2
        Z = torch.zeros(S.shape[:-2] + (OH, OW), dtype=cdt)
        Z[..., oy:oy + p.height, ox:ox + p.width] = S
        S = Z
    Hs = _t(asm_transfer(OH, OW, p.lambda_m, p.dx_m, p.dy_m, p.z_m, p.filt), cdt)
    return torch.fft.ifft2(torch.fft.fft2(S) * Hs, norm="forward")


def fftshift_ref(x: torch.Tensor, axes: Sequence[int]) -> torch.Tensor:
    """holonp::FFTShift: out[i] = in[(i + n//2) mod n] per axis (fftshift.cu:131-135)."""
    for ax in axes:
        n = x.shape[ax]
        x = torch.roll(x, shifts=-(n // 2), dims=ax)
    return x


def moments(P: torch.Tensor, p: PathParams, prec: str = "f64") -> torch.Tensor:
    """M0 = mean_k a_k, M1 = mean_k f_k a_k, M2 = mean_k (f_k f_k) a_k over the plane
    axis, a_k = |P_k| (or |P_k|^2), weights and their squares in F32 (graph_builder.cc:
    604-631); band-split M0 over [begin, split) and [split, end) when requested."""
    rdt, _ = _DT[prec]
    a = torch.abs(P).to(rdt) if P.is_complex() else P.to(rdt)
    if p.power2:
        a = a * a
    f32 = freq_weights(p)
    f = _t(f32, rdt)[:, None, None]
    f2 = _t(f32 * f32, rdt)[:, None, None]               # F32 product as the reference
    outs = [a.mean(dim=0), (f * a).mean(dim=0), (f2 * a).mean(dim=0)]
    if p.split_bin:
        s = p.split_bin - p.bin_begin
        outs += [a[:s].mean(dim=0), a[s:].mean(dim=0)]
    out = torch.stack(outs)
    if p.fftshift:
        out = fftshift_ref(out, (-2, -1))
    return out


def doppler_moments(window: np.ndarray, p: PathParams, prec: str = "f64") -> np.ndarray:
    """Full path for one [T][H][W] uint16 window -> [n_out][OH][OW] moment images."""
    S = time_dft(window, p, prec)
    if p.prop == "fresnel":
        if p.filt is not None:      # Filter2D on the band planes (graph_builder.cc:124-126)
            from . import stages
            S = torch.from_numpy(stages.filter2d(S.numpy(), *(int(v) for v in p.filt))).to(S.dtype)
        P = fresnel(S, p, prec, output_magnitude=True)
    elif p.prop == "asm":
        P = angular_spectrum(S, p, prec)
    else:
        raise ValueError(f"unknown propagation {p.prop!r}")
    M = moments(P, p, prec).numpy()
    if p.flatfield is not None:     # after the fftshift (graph_builder.cc:633-641)
        from . import stages
        M = stages.flatfield(M, *p.flatfield).astype(M.dtype)
    return M


def rel_l2(x, ref) -> float:
    """||x - ref|| / ||ref|| in float64 (complex128 when either side is complex)."""
    x = np.asarray(x)
    ref = np.asarray(ref)
    dt = np.complex128 if (np.iscomplexobj(x) or np.iscomplexobj(ref)) else np.float64
    x = x.astype(dt)
    ref = ref.astype(dt)
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / (den if den > 0 else 1.0))


# --------------------------------------------------------------------- named configs
def config_params(name: str) -> PathParams:
    """The workloads of the benchmark / parity configurations, by name."""
    if name == "cfg0_512_asm":
        kb, ke = band_bins(6e3, 15e3, 37e3, 32)
        return PathParams(512, 512, 32, 32, kb, ke, "asm", 852e-9, 12e-6, 12e-6, 0.3)
    if name == "cfg1_1024_fresnel":
        kb, ke = band_bins(6e3, 15e3, 37e3, 64)
        return PathParams(1024, 1024, 64, 16, kb, ke, "fresnel", 852e-9, 12e-6, 12e-6, 0.45,
                          fftshift=True)
    if name == "cfg4_2048_asm":
        kb, ke = band_bins(1e3, 5e3, 10e3, 128)
        split, _ = band_bins(4e3, 5e3, 10e3, 128)
        return PathParams(2048, 2048, 128, 32, kb, ke, "asm", 852e-9, 12e-6, 12e-6, 0.35,
                          weight_fs_hz=10e3, split_bin=split)
    raise KeyError(name)
