"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle still reproduces every committed vector (FP64 bit-for-bit up to FFT
rounding; its FP32 restatement within the parity tolerance).  GPU: the fused HIP path and
its time-DFT stage reproduce them to the FP32 parity bar (rel-L2 <= 1e-5)."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_l2

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLDEN, "index.json")) as f:
    INDEX = json.load(f)
NAMES = sorted(INDEX)


def _load(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    pk = dict(INDEX[name]["params"])
    for k in ("pad", "filt"):
        if k in pk and pk[k] is not None:
            pk[k] = tuple(pk[k])
    return z["window"], z["planes"], z["moments"], O.PathParams(**pk)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    window, planes, moments, p = _load(name)
    assert rel_l2(O.time_dft(window, p, "f64").numpy(), planes) < 1e-13
    assert rel_l2(O.doppler_moments(window, p, "f64"), moments) < 1e-12
    assert rel_l2(O.doppler_moments(window, p, "f32"), moments) < 1e-5


def test_fixture_set_covers_the_options():
    ps = [_load(n)[3] for n in NAMES]
    assert {p.prop for p in ps} == {"fresnel", "asm"}
    assert any(p.both_sidebands for p in ps) and any(p.power2 for p in ps)
    assert any(p.split_bin for p in ps) and any(p.pad != (0, 0) for p in ps)
    assert any(p.filt for p in ps) and any(p.fftshift for p in ps) and any(p.bin_begin == 0 for p in ps)


def _config(p):
    from holoflow_amd.hfx import MomentConfig
    return MomentConfig(p.height, p.width, p.window, p.stride, p.bin_begin, p.bin_end, p.prop,
                        lambda_m=p.lambda_m, dx_m=p.dx_m, dy_m=p.dy_m, z_m=p.z_m,
                        weight_fs_hz=p.weight_fs_hz, fftshift=p.fftshift, power2=p.power2,
                        both_sidebands=p.both_sidebands, pad=p.pad, filt=p.filt, split_bin=p.split_bin)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_path_matches_fixture(gpu, name):
    import torch
    from holoflow_amd.hfx import MomentPlan

    window, planes, moments, p = _load(name)
    plan = MomentPlan(_config(p))
    ring = torch.from_numpy(np.ascontiguousarray(window).view(np.int16)).cuda()
    out = plan.moments(ring, 0)
    S = plan.time_dft(ring, 0)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), moments) < 1e-5
    ref = planes
    if p.prop == "fresnel":         # the fused stage applies the chirp to the band planes
        ref = planes * O.fresnel_lens(p.height, p.width, p.lambda_m, p.dx_m, p.dy_m, p.z_m)
    assert rel_l2(S.cpu().numpy(), ref) < 1e-6
