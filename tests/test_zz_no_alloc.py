"""The allocation-free Scaled conversion (hfx_convert_noalloc) against hfx_convert with a
caller-supplied scratch buffer, through the C ABI: identical outputs on two calls in a row,
the second with a smaller input (the library-held partials keep the first call's contents,
which the second must not read), for U8 and U16 targets; plus the reference formula
int((x - min) * (top / (max - min)) + 0.5), clamped (conversion.cu:177-195), within one code."""
import ctypes

import numpy as np
import pytest
import torch

from holoflow_amd import _lib

pytestmark = pytest.mark.gpu

U8, U16, F32 = 4, 1, 2
SCALED = 1


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _with_scratch(x, out_dt):
    L = _lib.lib()
    n = x.numel()
    out = torch.empty(n, dtype=torch.uint8 if out_dt == U8 else torch.int16, device=x.device)
    scratch = torch.empty(int(L.hfx_convert_scratch_bytes(n)), dtype=torch.uint8, device=x.device)
    rc = L.hfx_convert(_ptr(x), F32, _ptr(out), out_dt, SCALED, n, _ptr(scratch), _stream())
    assert rc == 0, L.hfx_last_error()
    torch.cuda.synchronize()
    return out


def _no_alloc(x, out_dt):
    L = _lib.lib()
    n = x.numel()
    out = torch.empty(n, dtype=torch.uint8 if out_dt == U8 else torch.int16, device=x.device)
    rc = L.hfx_convert_noalloc(_ptr(x), F32, _ptr(out), out_dt, SCALED, n, _stream())
    assert rc == 0, L.hfx_last_error()
    torch.cuda.synchronize()
    return out


def _formula(x, top):
    x = x.astype(np.float32)
    lo, hi = np.float32(x.min()), np.float32(x.max())
    scale = np.float32(top) / (hi - lo)
    v = np.floor((x - lo) * scale + np.float32(0.5)).astype(np.int64)
    return np.clip(v, 0, top)


@pytest.mark.parametrize("out_dt,top", [(U8, 255), (U16, 65535)])
def test_no_alloc_scaled_matches_scratch_variant_twice(gpu, out_dt, top):
    assert _lib.lib().hfx_convert_noalloc_scratch_bytes() >= _lib.lib().hfx_convert_scratch_bytes(1 << 40)
    rng = np.random.default_rng(5)
    # first call: 3 M elements (1024 partial blocks); second: 1,000 elements (4 blocks) whose
    # range lies inside the first's, so stale partials from the first call would widen it
    big = rng.normal(0.0, 100.0, 3 << 20).astype(np.float32)
    small = rng.uniform(-1.0, 2.0, 1000).astype(np.float32)
    for host in (big, small):
        x = torch.from_numpy(host).to(gpu)
        a = _no_alloc(x, out_dt)
        b = _with_scratch(x, out_dt)
        assert torch.equal(a, b)
        got = a.cpu().numpy().astype(np.int64) & (0xFF if out_dt == U8 else 0xFFFF)
        assert np.abs(got - _formula(host, top)).max() <= 1
        assert got.min() == 0 and got.max() == top


def test_no_alloc_other_strategies_and_empty(gpu):
    L = _lib.lib()
    x = torch.arange(7, dtype=torch.float32, device=gpu)
    out = torch.empty(7, dtype=torch.complex64, device=gpu)
    assert L.hfx_convert_noalloc(_ptr(x), F32, _ptr(out), 3, 0, 7, _stream()) == 0   # F32 -> CF32 Real
    torch.cuda.synchronize()
    assert torch.equal(out.real.cpu(), x.cpu()) and not out.imag.any()
    assert L.hfx_convert_noalloc(_ptr(x), F32, _ptr(out), U8, SCALED, 0, _stream()) == 0
    assert L.hfx_convert_noalloc(_ptr(x), F32, _ptr(out), U8, SCALED, -1, _stream()) == _lib.HFX_E_INVALID
