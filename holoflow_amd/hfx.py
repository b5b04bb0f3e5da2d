"""Python host API over libhfx's C ABI (device memory and streams come from torch).

Every call goes through the HIP library; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import torch

from . import _lib
from ._lib import check, lib


def _stream(stream) -> C.c_void_p:
    if stream is None:
        stream = torch.cuda.current_stream()
    if isinstance(stream, torch.cuda.Stream):
        return C.c_void_p(stream.cuda_stream)
    return C.c_void_p(int(stream))


def _ptr(t: torch.Tensor) -> C.c_void_p:
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    return C.c_void_p(t.data_ptr())


_U16_STORAGE = (torch.int16, torch.uint16)


def check_ring(ring: torch.Tensor, height: int, width: int, window: int) -> None:
    """The C ABI receives only a pointer and ring.shape[0]: the tensor must be a contiguous
    [F][H][W] 16-bit frame ring of at least `window` frames on the device, or the kernels would
    read past it."""
    if ring.dtype not in _U16_STORAGE:
        raise ValueError(f"ring must hold uint16 frames (int16/uint16 storage), got {ring.dtype}")
    if ring.dim() != 3 or tuple(ring.shape[1:]) != (height, width):
        raise ValueError(f"ring must have shape [frames, {height}, {width}], got {tuple(ring.shape)}")
    if ring.shape[0] < window:
        raise ValueError(f"ring holds {ring.shape[0]} frames, fewer than the window {window}")
    if not ring.is_contiguous():
        raise ValueError("ring must be contiguous")
    if not ring.is_cuda:
        raise ValueError("ring must be a device tensor")


def check_out(out: torch.Tensor, n_windows: int, shape: Sequence[int]) -> None:
    """`out` must be a contiguous float32 device tensor holding n_windows (0: one window, no
    leading axis) moment stacks of `shape` = (n_out, OH, OW)."""
    per = 1
    for v in shape:
        per *= int(v)
    if out.dtype != torch.float32:
        raise ValueError(f"out must be float32, got {out.dtype}")
    if not out.is_contiguous():
        raise ValueError("out must be contiguous")
    if n_windows == 0:
        if out.numel() < per:
            raise ValueError(f"out holds {out.numel()} floats, fewer than one window's {per}")
    else:
        if out.dim() < 1 or out.shape[0] < n_windows or out[0].numel() != per:
            raise ValueError(f"out must have shape [>= {n_windows}, {', '.join(str(int(v)) for v in shape)}], "
                             f"got {tuple(out.shape)}")
    if not out.is_cuda:
        raise ValueError("out must be a device tensor")


@dataclass
class MomentConfig:
    """Settings of the fused Doppler-moment plan (mirror of hfx_plan_desc)."""

    height: int
    width: int
    window: int
    stride: int
    bin_begin: int
    bin_end: int
    prop: str = "fresnel"          # "fresnel" | "asm"
    lambda_m: float = 852e-9
    dx_m: float = 12e-6
    dy_m: float = 12e-6
    z_m: float = 0.3
    fs_hz: float = 37e3
    weight_fs_hz: float = 37e3
    precision: str = "fp32"        # fp32 | fp16 | bf16 | fp16_split
    fftshift: bool = False
    power2: bool = False
    both_sidebands: bool = False
    pad: tuple = (0, 0)
    filt: Optional[tuple] = None           # (r_inner, r_outer, s_inner, s_outer)
    split_bin: int = 0
    flatfield: Optional[tuple] = None      # (sigma_y, sigma_x)
    batch_windows: int = 1                 # windows per front launch in moments_batch
    batch_lanes: int = 0                   # streams moments_batch alternates over (0 = library default)

    def to_desc(self) -> _lib.PlanDesc:
        d = _lib.PlanDesc()
        lib().hfx_plan_desc_init(C.byref(d))
        d.height, d.width = self.height, self.width
        d.window, d.stride = self.window, self.stride
        d.fs_hz = self.fs_hz
        d.lambda_m, d.dx_m, d.dy_m, d.z_m = self.lambda_m, self.dx_m, self.dy_m, self.z_m
        d.bin_begin, d.bin_end = self.bin_begin, self.bin_end
        d.prop = {"fresnel": _lib.PROP_FRESNEL, "asm": _lib.PROP_ANGULAR_SPECTRUM}[self.prop]
        d.precision = _lib.PRECISIONS[self.precision]
        d.flags = ((_lib.FLAG_FFTSHIFT if self.fftshift else 0)
                   | (_lib.FLAG_POWER2 if self.power2 else 0)
                   | (_lib.FLAG_BOTH_SIDEBANDS if self.both_sidebands else 0))
        d.pad_height, d.pad_width = int(self.pad[0]), int(self.pad[1])
        if self.filt is not None:
            d.filter_enabled = 1
            d.r_inner, d.r_outer, d.s_inner, d.s_outer = (float(v) for v in self.filt)
        d.split_bin = self.split_bin
        d.weight_fs_hz = self.weight_fs_hz
        d.batch_windows = self.batch_windows
        d.batch_lanes = self.batch_lanes
        if self.flatfield is not None:
            d.flatfield_enabled = 1
            d.flatfield_sigma_y, d.flatfield_sigma_x = (float(v) for v in self.flatfield)
        return d


class MomentPlan:
    """Resident plan of the fused path: tables and workspaces live on the device."""

    def __init__(self, cfg: MomentConfig, stream=None):
        self.cfg = cfg
        self._h = C.c_void_p()
        d = cfg.to_desc()
        check(lib().hfx_plan_create(C.byref(d), _stream(stream), C.byref(self._h)))
        info = _lib.PlanInfo()
        check(lib().hfx_plan_get_info(self._h, C.byref(info)))
        self.out_shape = (info.n_outputs, info.out_height, info.out_width)
        self.n_planes = info.n_planes
        self.workspace_bytes = info.workspace_bytes
        self.batch_lanes = info.batch_lanes

    def close(self):
        if self._h and lib is not None:      # module globals are gone at interpreter exit
            lib().hfx_plan_destroy(self._h)
            self._h = C.c_void_p()

    __del__ = close

    def update_z(self, z_m: float, stream=None):
        check(lib().hfx_plan_update_z(self._h, z_m, _stream(stream)))
        self.cfg.z_m = z_m

    def new_output(self, n_windows: int = 0) -> torch.Tensor:
        shape = self.out_shape if n_windows == 0 else (n_windows,) + self.out_shape
        return torch.empty(shape, dtype=torch.float32, device="cuda")

    def moments(self, ring: torch.Tensor, first_frame: int = 0, out: Optional[torch.Tensor] = None,
                stream=None) -> torch.Tensor:
        """ring: device uint16 [F][H][W] (int16 storage is accepted bit-for-bit)."""
        check_ring(ring, self.cfg.height, self.cfg.width, self.cfg.window)
        out = self.new_output() if out is None else out
        check_out(out, 0, self.out_shape)
        check(lib().hfx_moments(self._h, _ptr(ring), ring.shape[0], first_frame, _ptr(out),
                                _stream(stream)))
        return out

    def moments_batch(self, ring: torch.Tensor, first_frame: int, n_windows: int,
                      out: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
        check_ring(ring, self.cfg.height, self.cfg.width, self.cfg.window)
        if n_windows < 0:
            raise ValueError("n_windows must be >= 0")
        out = self.new_output(n_windows) if out is None else out
        if n_windows:
            check_out(out, n_windows, self.out_shape)
        stride = out[0].numel() if n_windows else 0
        check(lib().hfx_moments_batch(self._h, _ptr(ring), ring.shape[0], first_frame, n_windows,
                                      stride, _ptr(out), _stream(stream)))
        return out

    @staticmethod
    def moments_group(plans, rings, first_frames, outs=None, stream=None):
        """One window of each plan in one front launch and one launch per back pass
        (hfx_moments_group): plans of one shape whose z, wavelength or pitch may differ.
        Returns the list of outputs ([n_out][OH][OW] each)."""
        n = len(plans)
        if not (len(rings) == len(first_frames) == n):
            raise ValueError("plans, rings and first_frames must have the same length")
        outs = [p.new_output() for p in plans] if outs is None else list(outs)
        for p, r, o in zip(plans, rings, outs):
            check_ring(r, p.cfg.height, p.cfg.width, p.cfg.window)
            check_out(o, 0, p.out_shape)
        arr = lambda T, xs: (T * n)(*xs)  # noqa: E731
        check(lib().hfx_moments_group(arr(C.c_void_p, [p._h for p in plans]),
                                      arr(C.c_void_p, [_ptr(r) for r in rings]),
                                      arr(C.c_int64, [r.shape[0] for r in rings]),
                                      arr(C.c_int64, [int(f) for f in first_frames]),
                                      arr(C.c_void_p, [_ptr(o) for o in outs]), n, _stream(stream)))
        return outs

    def run_stage(self, stage: int, ring: Optional[torch.Tensor], first_frame: int = 0,
                  out: Optional[torch.Tensor] = None, stream=None) -> None:
        """Stage 0 (time DFT + row FFT into the workspace) or stage 1 (rest -> out)."""
        if ring is not None:
            check_ring(ring, self.cfg.height, self.cfg.width, self.cfg.window)
        if out is not None:
            check_out(out, 0, self.out_shape)
        rp = _ptr(ring) if ring is not None else None
        op = _ptr(out) if out is not None else None
        n = ring.shape[0] if ring is not None else 0
        check(lib().hfx_plan_run_stage(self._h, stage, rp, n, first_frame, op, _stream(stream)))

    def time_dft(self, ring: torch.Tensor, first_frame: int = 0, stream=None) -> torch.Tensor:
        """Runs the time-DFT stage only: [n_planes][H][W] complex64 (lens applied for Fresnel)."""
        check_ring(ring, self.cfg.height, self.cfg.width, self.cfg.window)
        out = torch.empty((self.n_planes, self.cfg.height, self.cfg.width), dtype=torch.complex64,
                          device="cuda")
        check(lib().hfx_plan_time_dft(self._h, _ptr(ring), ring.shape[0], first_frame, _ptr(out),
                                      _stream(stream)))
        return out


# ----------------------------------------------------------------------- per-node kinds
def dft_axis(x: torch.Tensor, axis: int, k_begin: int = 0, k_end: Optional[int] = None,
             scale: float = 1.0, stream=None) -> torch.Tensor:
    """RFFT/FFT along `axis` (times `scale`), bins [k_begin, k_end) -> complex64."""
    axis = axis % x.dim()
    n = x.shape[axis]
    k_end = n if k_end is None else k_end
    outer = 1
    for s in x.shape[:axis]:
        outer *= s
    inner = 1
    for s in x.shape[axis + 1:]:
        inner *= s
    dt = {torch.int16: _lib.DTYPE_U16, torch.uint16: _lib.DTYPE_U16, torch.float32: _lib.DTYPE_F32,
          torch.complex64: _lib.DTYPE_CF32}[x.dtype]
    x = x.contiguous()
    shape = list(x.shape)
    shape[axis] = k_end - k_begin
    out = torch.empty(shape, dtype=torch.complex64, device="cuda")
    check(lib().hfx_dft_axis(_ptr(x), dt, outer, n, inner, k_begin, k_end, scale, _ptr(out),
                             _stream(stream)))
    return out


class Fresnel:
    def __init__(self, batch, height, width, lambda_m, dx_m, dy_m, z_m, in_real=False,
                 output_magnitude=False, skip_phase_shift=False, stream=None):
        self._h = C.c_void_p()
        self.shape = (batch, height, width)
        self.output_magnitude = output_magnitude
        check(lib().hfx_fresnel_create(batch, height, width, lambda_m, dx_m, dy_m, z_m, int(in_real),
                                       int(output_magnitude), int(skip_phase_shift), _stream(stream),
                                       C.byref(self._h)))

    def update_z(self, z_m, stream=None):
        check(lib().hfx_fresnel_update_z(self._h, z_m, _stream(stream)))

    def __call__(self, x: torch.Tensor, out: Optional[torch.Tensor] = None, stream=None):
        if out is None:
            dt = torch.float32 if self.output_magnitude else torch.complex64
            out = torch.empty(x.shape, dtype=dt, device="cuda")
        check(lib().hfx_fresnel_execute(self._h, _ptr(x), _ptr(out), _stream(stream)))
        return out

    def close(self):
        if self._h and lib is not None:      # module globals are gone at interpreter exit
            lib().hfx_fresnel_destroy(self._h)
            self._h = C.c_void_p()

    __del__ = close


class AngularSpectrum:
    def __init__(self, batch, height, width, lambda_m, dx_m, dy_m, z_m, in_real=False, pad=(0, 0),
                 filt=None, stream=None):
        d = _lib.AsmDesc()
        d.batch, d.height, d.width = batch, height, width
        d.lambda_m, d.dx_m, d.dy_m, d.z_m = lambda_m, dx_m, dy_m, z_m
        d.in_real = int(in_real)
        d.pad_height, d.pad_width = int(pad[0]), int(pad[1])
        if filt is not None:
            d.filter_enabled = 1
            d.r_inner, d.r_outer, d.s_inner, d.s_outer = (float(v) for v in filt)
        self.out_shape = (batch, pad[0] or height, pad[1] or width)
        self._h = C.c_void_p()
        check(lib().hfx_asm_create(C.byref(d), _stream(stream), C.byref(self._h)))

    def __call__(self, x: torch.Tensor, out: Optional[torch.Tensor] = None, stream=None):
        if out is None:
            out = torch.empty(self.out_shape, dtype=torch.complex64, device="cuda")
        check(lib().hfx_asm_execute(self._h, _ptr(x), _ptr(out), _stream(stream)))
        return out

    def close(self):
        if self._h and lib is not None:      # module globals are gone at interpreter exit
            lib().hfx_asm_destroy(self._h)
            self._h = C.c_void_p()

    __del__ = close


def weighted_mean(x: torch.Tensor, weights: Optional[torch.Tensor] = None, power2=False,
                  absval=False, stream=None) -> torch.Tensor:
    """Mean over axis 0 of |x| (complex) or x (real; |x| with absval), optionally weighted
    per plane."""
    planes = x.shape[0]
    npix = x[0].numel()
    out = torch.empty(x.shape[1:], dtype=torch.float32, device="cuda")
    w = C.c_void_p(weights.data_ptr()) if weights is not None else None
    flags = (_lib.REDUCE_POWER2 if power2 else 0) | (_lib.REDUCE_ABS if absval else 0)
    check(lib().hfx_weighted_mean(_ptr(x.contiguous()), int(x.is_complex()), planes, npix, w,
                                  flags, _ptr(out), _stream(stream)))
    return out


def fftshift(x: torch.Tensor, axes: Sequence[int], stream=None) -> torch.Tensor:
    x = x.contiguous()
    nd = x.dim()
    shape = (C.c_int64 * nd)(*x.shape)
    mask = (C.c_int32 * nd)(*[1 if (i in [a % nd for a in axes]) else 0 for i in range(nd)])
    out = torch.empty_like(x)
    check(lib().hfx_fftshift(_ptr(x), _ptr(out), x.element_size(), nd, shape, mask, _stream(stream)))
    return out


def copy_contiguous(src: torch.Tensor, dst: torch.Tensor, stream=None) -> torch.Tensor:
    """Device copy of a contiguous tensor through the library's gather kernel."""
    if not (src.is_contiguous() and dst.is_contiguous()) or src.nbytes != dst.nbytes:
        raise ValueError("copy_contiguous needs two contiguous tensors of the same size")
    e = 4 if src.nbytes % 4 == 0 else 1
    n = src.nbytes // e
    shape = (C.c_int64 * 1)(n)
    strides = (C.c_int64 * 1)(e)
    check(lib().hfx_strided_copy(_ptr(src), shape, strides, 0, 1, e, _ptr(dst), 0, 0, _stream(stream)))
    return dst


def convert_u16(x: torch.Tensor, to_complex=False, stream=None) -> torch.Tensor:
    out = torch.empty(x.shape, dtype=torch.complex64 if to_complex else torch.float32, device="cuda")
    check(lib().hfx_convert_u16(_ptr(x.contiguous()), _ptr(out), x.numel(), int(to_complex),
                                _stream(stream)))
    return out


def synth_desc(height, width, **kw) -> _lib.SynthDesc:
    d = _lib.SynthDesc()
    lib().hfx_synth_desc_init(C.byref(d), height, width)
    for k, v in kw.items():
        setattr(d, k, v)
    return d


def synth_frames(desc: _lib.SynthDesc, first: int, n: int, out: Optional[torch.Tensor] = None,
                 phase_state: Optional[torch.Tensor] = None, stream=None):
    """Generates frames [first, first+n) into out ([n][H][W] int16 storage of uint16)."""
    H, W = desc.height, desc.width
    if out is None:
        out = torch.empty((n, H, W), dtype=torch.int16, device="cuda")
    if phase_state is None:
        phase_state = torch.zeros(H * W, dtype=torch.float64, device="cuda")
    check(lib().hfx_synth_frames(C.byref(desc), _ptr(out), first, n, _ptr(phase_state),
                                 _stream(stream)))
    return out, phase_state
