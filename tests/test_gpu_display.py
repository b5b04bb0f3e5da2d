"""GPU parity of the kinds right after the path (the post-moment display chain,
graph_builder.cc:647-670): Conversion in all its configurations, SlidingAverage and
DualReaderBatchQueue, through the node ABI.

Conversion Real / Modulus / Argument compare to NumPy (float32, rel 1e-6; Real bit-exact);
Scaled to the reference formula int((x - min) * (top / (max - min)) + 0.5), clamped, within one
code (the reference's CUDA build may contract the multiply-add).  SlidingAverage follows the
reference's FP32 update order (running += x / ws; running -= old / ws) emulated in NumPy
float32: bit-exact.  The queue outputs are data movement: bit-exact."""
import numpy as np
import pytest
import torch

from holoflow_amd.nodes import AsyncNode, Node, desc, desc_of

pytestmark = pytest.mark.gpu


def _run(kind, settings, *tensors):
    n = Node(kind, settings, [desc_of(t) for t in tensors])
    out = n.execute(*tensors)[0]
    torch.cuda.synchronize()
    return out


def test_conversion_real_modulus_argument(gpu):
    rng = np.random.default_rng(3)
    u8 = rng.integers(0, 256, (5, 7)).astype(np.uint8)
    out = _run("Conversion", {"target": "F32", "strategy": "Real"}, torch.from_numpy(u8).cuda())
    assert np.array_equal(out.cpu().numpy(), u8.astype(np.float32))
    out = _run("Conversion", {"target": "CF32", "strategy": "Real"}, torch.from_numpy(u8).cuda())
    assert np.array_equal(out.cpu().numpy(), u8.astype(np.complex64))
    f = rng.normal(size=(5, 7)).astype(np.float32)
    out = _run("Conversion", {"target": "CF32", "strategy": "Real"}, torch.from_numpy(f).cuda())
    assert np.array_equal(out.cpu().numpy(), f.astype(np.complex64))
    z = (rng.normal(size=(64, 33)) + 1j * rng.normal(size=(64, 33))).astype(np.complex64)
    mod = _run("Conversion", {"target": "F32", "strategy": "Modulus"}, torch.from_numpy(z).cuda()).cpu().numpy()
    arg = _run("Conversion", {"target": "F32", "strategy": "Argument"}, torch.from_numpy(z).cuda()).cpu().numpy()
    np.testing.assert_allclose(mod, np.abs(z.astype(np.complex128)), rtol=1e-6)
    np.testing.assert_allclose(arg, np.angle(z.astype(np.complex128)), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("target,top", [("U8", 255), ("U16", 65535)])
def test_conversion_scaled(gpu, target, top):
    rng = np.random.default_rng(7)
    x = (rng.normal(size=(301, 257)) * 50 + 10).astype(np.float32)   # spans several blocks
    out = _run("Conversion", {"target": target, "strategy": "Scaled"}, torch.from_numpy(x).cuda())
    got = out.cpu().numpy().view(np.uint16 if target == "U16" else np.uint8).astype(np.int64)
    lo, hi = np.float32(x.min()), np.float32(x.max())
    scale = np.float32(top) / np.float32(hi - lo)
    ref = np.clip(((x - lo) * scale + np.float32(0.5)).astype(np.int64), 0, top)
    assert np.abs(got - ref).max() <= 1
    assert got.min() == 0 and got.max() == top


def test_sliding_average_matches_fp32_running_mean(gpu):
    ws, cap, H, W = 4, 3, 16, 24
    q = AsyncNode("SlidingAverage", {"target_capacity": cap, "window_size": ws}, [desc([1, H, W], "F32")])
    rng = np.random.default_rng(11)
    frames = rng.normal(size=(12, 1, H, W)).astype(np.float32)
    ring = [np.zeros((1, H, W), np.float32) for _ in range(cap + ws)]
    running = np.zeros((1, H, W), np.float32)
    wsf = np.float32(ws)
    expect, got = [], []
    w_idx, a_idx = 0, cap
    for f in frames:
        while not q.push(torch.from_numpy(f).cuda()):
            got.append(q.pop().cpu().numpy())
        ring[w_idx] = f.copy()
        running = (running + ring[w_idx] / wsf).astype(np.float32)
        running = (running - ring[a_idx] / wsf).astype(np.float32)
        ring[a_idx] = running.copy()
        expect.append(running.copy())
        w_idx, a_idx = (w_idx + 1) % (cap + ws), (a_idx + 1) % (cap + ws)
    while True:
        o = q.pop()
        if o is None:
            break
        got.append(o.cpu().numpy())
    # the first pop is the first full window (after ws pushes)
    assert len(got) == len(frames) - ws + 1
    for g, e in zip(got, expect[ws - 1:]):
        assert np.array_equal(g, e)
    mean_last = frames[-ws:].mean(axis=0)
    np.testing.assert_allclose(got[-1], mean_last, rtol=1e-5, atol=1e-5)


def test_sliding_average_validity_and_discard(gpu):
    q = AsyncNode("SlidingAverage", {"target_capacity": 4, "window_size": 2, "discard_first": 1},
                  [desc([1, 2, 2], "F32"), desc([1], "U8", mem_loc="Host")])
    vals = [5.0, 1.0, 100.0, 3.0]            # 5 discarded (first), 100 flagged invalid
    for v, ok in zip(vals, [1, 1, 0, 1]):
        assert q.push(torch.full((1, 2, 2), v, device="cuda"), ok)
    out = q.pop()
    assert out is not None and np.allclose(out.cpu().numpy(), 2.0)   # mean of 1 and 3
    assert q.pop() is None


def test_dual_reader_batch_queue_current_delayed_valid(gpu):
    ws = 5                                     # delay h = 2
    q = AsyncNode("DualReaderBatchQueue", {"target_capacity": 4, "window_size": ws}, [desc([2, 3], "F32")])
    seq = []
    n = 0
    for batch in range(4):
        x = torch.tensor([[n, n, n], [n + 1, n + 1, n + 1]], dtype=torch.float32, device="cuda")
        n += 2
        assert q.push(x)
        while True:
            o = q.pop()
            if o is None:
                break
            cur, dly, valid = o
            seq.append((float(cur[0, 0]), float(dly[0, 0]), valid))
    assert [c for c, _, _ in seq] == list(range(8))
    # delayed reader: zero scratch until sequence >= 2, then I[n-2]
    assert [d for _, d, _ in seq] == [0.0, 0.0, 0.0, 1.0, 2.0, 3.0, 4.0, 5.0]
    assert [v for _, _, v in seq] == [0, 0, 0, 0, 1, 1, 1, 1]
