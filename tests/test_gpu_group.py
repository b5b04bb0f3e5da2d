"""Windows of different acquisitions batched into one launch (hfx_moments_group): plans of
one shape (frame size, window, band, weights, propagation kind) whose z, wavelength and pitch
differ run one window each in a single front launch and one launch per back pass.  Each
window's moments must equal hfx_moments on its plan alone, bit for bit: the grouped launch
runs the same kernels on the same frames and tables, only the window a block works on comes
from a per-window table.  Group validation (mismatched shapes, a plan twice, too many
windows, sizes without grouped kernels) is checked through the same ABI."""
import numpy as np
import pytest
import torch

from holoflow_amd import hfx
from holoflow_amd._lib import HfxError, HfxInvalidArgument

pytestmark = pytest.mark.gpu


def _ring(H, W, n, seed, radius):
    d = hfx.synth_desc(H, W, disk_radius_px=radius, seed=seed)
    ring, _ = hfx.synth_frames(d, 0, n)
    return ring


def _cfg(H, W, T, kb, ke, prop, z, lam=852e-9, dx=12e-6, **kw):
    return hfx.MomentConfig(H, W, T, T // 4, kb, ke, prop, lam, dx, dx, z, batch_lanes=1, **kw)


@pytest.mark.parametrize("H,W,T,kb,ke,prop,extra", [
    (512, 512, 32, 6, 13, "fresnel", dict(fftshift=True)),           # up to 8 bins: interleaved DFT
    (512, 768, 64, 11, 26, "fresnel", dict(fftshift=True)),          # 9..16 bins: folded DFT
    (768, 768, 32, 5, 14, "asm", {}),                                 # angular spectrum, 3 passes
    (1024, 1024, 64, 11, 26, "asm", dict(split_bin=20)),             # band split (5 images)
])
def test_group_equals_each_plan_alone(gpu, H, W, T, kb, ke, prop, extra):
    zs = (0.25, 0.45, 0.6, 0.35)[: 4 if H < 1024 else 3]
    lams = (852e-9, 852e-9, 633e-9, 852e-9)
    plans = [hfx.MomentPlan(_cfg(H, W, T, kb, ke, prop, z, lam=lams[i], **extra)) for i, z in enumerate(zs)]
    rings = [_ring(H, W, T + 3 * (T // 4), seed=40 + i, radius=0.4 * H) for i in range(len(plans))]
    # the ring of window 1 wraps: its window starts 2 strides before the ring's end
    firsts = [0, rings[1].shape[0] - 2 * (T // 4)] + [T // 4 * (i + 1) for i in range(2, len(plans))]
    outs = hfx.MomentPlan.moments_group(plans, rings, firsts)
    for p, r, f, o in zip(plans, rings, firsts, outs):
        alone = p.moments(r, f)
        torch.cuda.synchronize()
        assert torch.equal(o, alone)
    # the windows really differ (different z, frames): the group did not reuse one window
    assert not torch.equal(outs[0], outs[1])
    for p in plans:
        p.close()


def test_group_of_one_and_repeated_calls(gpu):
    plans = [hfx.MomentPlan(_cfg(512, 512, 32, 6, 13, "fresnel", z)) for z in (0.3, 0.5)]
    rings = [_ring(512, 512, 40, seed=s, radius=200.0) for s in (1, 2)]
    one = hfx.MomentPlan.moments_group(plans[:1], rings[:1], [4])
    assert torch.equal(one[0], plans[0].moments(rings[0], 4))
    a = hfx.MomentPlan.moments_group(plans, rings, [0, 8])
    b = hfx.MomentPlan.moments_group(plans[::-1], rings[::-1], [8, 0])   # order within the group is free
    torch.cuda.synchronize()
    assert torch.equal(a[0], b[1]) and torch.equal(a[1], b[0])


def test_group_validation(gpu):
    base = _cfg(512, 512, 32, 6, 13, "fresnel", 0.3)
    p0 = hfx.MomentPlan(base)
    other_band = hfx.MomentPlan(_cfg(512, 512, 32, 6, 12, "fresnel", 0.3))
    other_kind = hfx.MomentPlan(_cfg(512, 512, 32, 6, 13, "asm", 0.3))
    ring = _ring(512, 512, 40, seed=3, radius=200.0)
    for bad in (other_band, other_kind):
        with pytest.raises(HfxInvalidArgument):
            hfx.MomentPlan.moments_group([p0, bad], [ring, ring], [0, 0])
    with pytest.raises(HfxInvalidArgument):                     # one workspace per plan
        hfx.MomentPlan.moments_group([p0, p0], [ring, ring], [0, 8])
    extra = [hfx.MomentPlan(_cfg(512, 512, 32, 6, 13, "fresnel", 0.3 + 0.01 * i)) for i in range(4)]
    with pytest.raises(HfxInvalidArgument):                     # more than HFX_GROUP_MAX windows
        hfx.MomentPlan.moments_group([p0] + extra, [ring] * 5, [0] * 5)
    wide = [hfx.MomentPlan(_cfg(64, 2048, 32, 6, 13, "asm", z)) for z in (0.3, 0.4)]
    r2 = _ring(64, 2048, 40, seed=4, radius=30.0)
    with pytest.raises(HfxError) as e:                          # no grouped kernels at 2048 wide
        hfx.MomentPlan.moments_group(wide, [r2, r2], [0, 8])
    assert not isinstance(e.value, HfxInvalidArgument)
    for p in [p0, other_band, other_kind] + extra + wide:
        p.close()
