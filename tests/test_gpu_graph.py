"""GPU parity of the node layer: each kind through the node ABI, and the reference's own
Doppler-moment graph (graph_builder.cc:232-305, 522-566, 578-647) wired node by node,
against the oracle and against the fused DopplerMoments node.

Data movement kinds (Copy, Slice, Reshape, Concatenate, BatchQueue, FFTShiftNp, Arange)
are checked bit-exactly; arithmetic kinds to FP32 tolerance against float64 references."""
import numpy as np
import pytest
import torch

from holoflow_amd import nodes
from holoflow_amd.nodes import Node, AsyncNode, desc, desc_of
from oracle import oracle as O
from oracle.synth import SynthParams, synth_frames
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _u16(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int16)).cuda()


def _run(kind, settings, *tensors):
    n = Node(kind, settings, [desc_of(t) for t in tensors])
    outs = n.execute(*tensors)
    torch.cuda.synchronize()
    return outs if len(outs) > 1 else outs[0]


# ------------------------------------------------------------------ data movement
def test_copy_and_slice_views_bit_exact(gpu):
    x = torch.randn(3, 7, 10, dtype=torch.complex64, device="cuda")
    v = _run("Slice", {"slices": [1, {"start": 2, "stop": 7, "step": 2}, {"start": -6}]}, x)
    assert v.data_ptr() == x[1, 2::2, -6:].data_ptr()
    c = _run("Copy", {}, v)
    assert c.is_contiguous() and torch.equal(c, x[1, 2::2, -6:])


def test_reshape_view_and_copy(gpu):
    x = torch.arange(24, dtype=torch.float32, device="cuda").reshape(4, 6)
    v = _run("Reshape", {"shape": [2, -1, 3], "copy": False}, x)
    assert v.data_ptr() == x.data_ptr() and torch.equal(v, x.reshape(2, 4, 3))
    s = x[:, :3]
    c = _run("Reshape", {"shape": [12]}, s)
    assert c.data_ptr() != x.data_ptr() and torch.equal(c, s.reshape(12))


@pytest.mark.parametrize("axis", [0, 1, -1, None])
def test_concatenate_strided_inputs(gpu, axis):
    base = torch.randn(4, 9, 6, dtype=torch.complex64, device="cuda")
    a = base[:, 1:4]           # strided, offset
    b = base[:, 5:9:2]         # strided with step
    st = {"axis": axis}
    if axis == 0:
        a, b = base[1:3, 1:4], base[:1, 5:8]
    if axis == -1:
        a, b = base[:, :, 1:3], base[:, :, 3:6]
    out = _run("Concatenate", st, a, b)
    ref = torch.cat([a.reshape(-1), b.reshape(-1)]) if axis is None else torch.cat([a, b], dim=axis)
    assert torch.equal(out, ref)


def test_arange_values_match_reference_rule(gpu):
    # f0 + df*i in double, cast (arange.cu:61-76): the frequency weights of the graph
    T, kb, ke, fs = 64, 11, 26, 37e3
    df = fs / T
    out = _run("Arange", {"start": kb * df, "stop": ke * df, "step": df, "dtype": "F32"})
    ref = np.array([np.float32(kb * df + df * i) for i in range(O.arange_len(kb * df, ke * df, df))])
    assert np.array_equal(out.cpu().numpy(), ref)
    n = Node("Arange", {"start": 0, "stop": 3, "step": 1, "dtype": "CF32"}, [])
    c = n.execute()[0]
    assert torch.equal(c.cpu(), torch.tensor([0, 1, 2], dtype=torch.complex64))
    u = Node("Arange", {"start": 2, "stop": 9, "step": 3, "dtype": "U8"}, []).execute()[0]
    assert u.cpu().tolist() == [2, 5, 8]


def test_multiply_broadcast(gpu):
    f = torch.randn(1, 5, 1, 1, device="cuda")
    a = torch.randn(3, 5, 4, 6, device="cuda")
    assert torch.allclose(_run("Multiply", {}, f, a), f * a, rtol=0, atol=0)
    z = torch.randn(4, 6, dtype=torch.complex64, device="cuda")
    r = torch.randn(6, device="cuda")
    out = _run("Multiply", {}, z, r)
    assert torch.allclose(out, z * r, rtol=1e-6, atol=1e-6)


def test_fftshift_axes_bit_exact(gpu):
    x = torch.randn(2, 1, 5, 8, device="cuda")
    out = _run("FFTShiftNp", {"axes": [-2, -1]}, x)
    assert torch.equal(out.cpu(), O.fftshift_ref(x.cpu(), (-2, -1)))


def test_batch_queue_order_and_backpressure(gpu):
    # batch_queue.cc:54-181: [1,...] pushes, [N,...] pops, capacity target 2N
    q = AsyncNode("BatchQueue", {"target_capacity": 4, "output_size": 2, "output_stride": 2},
                  [desc([1, 3, 4], "F32")])
    frames = [torch.full((1, 3, 4), float(i), device="cuda") for i in range(8)]
    pushed = 0
    while pushed < len(frames) and q.push(frames[pushed]):
        pushed += 1
    assert pushed == 5        # ring of lcm_above(1, 2, 4 + 1) = 6 slots keeps one free
    a = q.pop()
    assert a is not None and a[:, 0, 0].tolist() == [0.0, 1.0]
    b = q.pop()
    assert b[:, 0, 0].tolist() == [2.0, 3.0]
    assert q.pop() is None    # one element left, stride 2 not reached
    assert q.push(frames[5]) and q.pop()[:, 0, 0].tolist() == [4.0, 5.0]


# ------------------------------------------------------------------ arithmetic kinds
def test_conversion_u16(gpu):
    a = np.random.default_rng(1).integers(0, 4096, (4, 8, 8)).astype(np.uint16)
    out = _run("Conversion", {"target": "F32", "strategy": "Real"}, _u16(a))
    assert np.array_equal(out.cpu().numpy(), a.astype(np.float32))


@pytest.mark.parametrize("kind,norm", [("RFFT", "backward"), ("FFT", "forward"), ("FFT", "ortho")])
def test_time_axis_transforms(gpu, kind, norm):
    x = torch.randn(2, 32, 6, 5, device="cuda")
    out = _run(kind, {"axis": 1, "norm": norm}, x)
    xd = x.cpu().double()
    ref = torch.fft.rfft(xd, dim=1, norm=norm) if kind == "RFFT" else torch.fft.fft(xd, dim=1, norm=norm)
    assert rel_l2(out.cpu().numpy(), ref.numpy()) < 1e-6


def test_fresnel_and_asm_nodes_vs_oracle(gpu):
    p = O.PathParams(48, 64, 32, 32, 6, 13, "fresnel", z_m=0.3)
    S = torch.randn(2, 3, 48, 64, dtype=torch.complex64)
    st = {"lambda": p.lambda_m, "dx": p.dx_m, "dy": p.dy_m, "z": p.z_m, "output_magnitude": True}
    out = _run("FresnelDiffraction", st, S.cuda())
    assert rel_l2(out.cpu().numpy(), O.fresnel(S, p, "f64").numpy()) < 1e-5
    n = Node("FresnelDiffraction", st, [desc_of(S.cuda())])
    assert n.update(dict(st, z=0.45))                         # z-only change keeps the task
    p.z_m = float(np.float32(0.45))
    out = n.execute(S.cuda())[0]
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), O.fresnel(S, p, "f64").numpy()) < 1e-5
    pa = O.PathParams(48, 64, 32, 32, 6, 13, "asm", z_m=0.3, pad=(64, 96), filt=(3, 20, 2, 4))
    sa = {"lambda": pa.lambda_m, "dx": pa.dx_m, "dy": pa.dy_m, "z": pa.z_m,
          "padding": {"width": 96, "height": 64},
          "filter": {"r_inner": 3, "r_outer": 20, "s_inner": 2, "s_outer": 4}}
    out = _run("AngularSpectrum", sa, S.cuda())
    # the node keeps its optics in float32 as the reference's settings do; the transfer
    # phase 2*pi*z/lambda (~2e6 rad here) makes that rounding visible in the complex field
    f = lambda v: float(np.float32(v))  # noqa: E731
    pa.lambda_m, pa.dx_m, pa.dy_m, pa.z_m = f(pa.lambda_m), f(pa.dx_m), f(pa.dy_m), f(pa.z_m)
    assert rel_l2(out.cpu().numpy(), O.angular_spectrum(S, pa, "f64").numpy()) < 1e-5


def test_reductions(gpu):
    z = torch.randn(2, 6, 5, 7, dtype=torch.complex64, device="cuda")
    out = _run("MeanAbs", {"axis": [-3], "keepdims": True}, z)
    ref = z.cpu().abs().double().mean(dim=1, keepdim=True)
    assert rel_l2(out.cpu().numpy(), ref.numpy()) < 1e-6
    r = torch.randn(3, 4, 5, device="cuda")
    out = _run("Mean", {"axis": [0], "keepdims": False}, r)
    assert rel_l2(out.cpu().numpy(), r.cpu().double().mean(dim=0).numpy()) < 1e-6
    out = _run("Abs", {}, z)
    assert rel_l2(out.cpu().numpy(), z.cpu().abs().numpy()) < 1e-6


# ------------------------------------------------------------------ the reference graph
def _graph_moments(frames_u16, p: O.PathParams, n_pre=2):
    """The reference's node chain for one time method and propagation, node by node."""
    T, H, W = p.window, p.height, p.width
    kb, ke = p.bin_begin, p.bin_end
    x = _u16(frames_u16[:T])
    conv = Node("Conversion", {"target": "F32", "strategy": "Real"}, [desc_of(x)])
    q = AsyncNode("BatchQueue", {"target_capacity": 2 * n_pre, "output_size": n_pre, "output_stride": n_pre},
                  [desc([1, T, H, W], "F32")])
    for w in range(n_pre):
        xw = _u16(frames_u16[w * p.stride:w * p.stride + T])
        f = conv.execute(xw)[0]
        f4 = _run("Reshape", {"shape": [1, T, H, W], "copy": False}, f)
        assert q.push(f4)
    batch = q.pop()                                               # [n_pre, T, H, W]
    if not p.both_sidebands:
        FH = _run("RFFT", {"axis": 1}, batch)
        FH = _run("Copy", {}, _run("Slice", {"slices": [{}, {"start": kb, "stop": ke}, {}, {}]}, FH))
    else:
        F = _run("FFT", {"axis": 1}, batch)
        pos = _run("Slice", {"slices": [{}, {"start": kb, "stop": ke}, {}, {}]}, F)
        neg = _run("Slice", {"slices": [{}, {"start": T - ke, "stop": T - kb}, {}, {}]}, F)
        FH = _run("Concatenate", {"axis": 1}, pos, neg)
    df = 37e3 / T
    f0, f1 = kb * df, ke * df
    freqs = _run("Arange", {"start": f0, "stop": f1, "step": df, "dtype": "F32"})
    if p.both_sidebands:
        neg_f = _run("Arange", {"start": f0 - 37e3, "stop": f1 - 37e3, "step": df, "dtype": "F32"})
        freqs = _run("Concatenate", {"axis": 0}, freqs, neg_f)
    nf = FH.shape[1]
    freqs = _run("Reshape", {"shape": [1, nf, 1, 1]}, freqs)
    f2 = _run("Multiply", {}, freqs, freqs)
    results = []
    for i in range(n_pre):
        cur = _run("Copy", {}, _run("Slice", {"slices": [{"start": i, "stop": i + 1}, {}, {}, {}]}, FH))
        if p.prop == "fresnel":
            Z = _run("FresnelDiffraction", {"lambda": p.lambda_m, "dx": p.dx_m, "dy": p.dy_m, "z": p.z_m,
                                            "axes": [-2, -1], "output_magnitude": True}, cur)
            absS = Z
            m0 = _run("Mean", {"axis": [-3], "keepdims": True}, Z)
        else:
            Z = _run("AngularSpectrum", {"lambda": p.lambda_m, "dx": p.dx_m, "dy": p.dy_m, "z": p.z_m}, cur)
            m0 = _run("MeanAbs", {"axis": [-3], "keepdims": True}, Z)
            absS = _run("Abs", {}, Z)
        m1 = _run("Mean", {"axis": [-3], "keepdims": True}, _run("Multiply", {}, freqs, absS))
        m2 = _run("Mean", {"axis": [-3], "keepdims": True}, _run("Multiply", {}, f2, absS))
        ms = []
        for m in (m0, m1, m2):
            if p.fftshift:
                m = _run("FFTShiftNp", {"axes": [-2, -1]}, m)
            ms.append(_run("Mean", {"axis": [0], "keepdims": False}, m)[0])
        results.append(torch.stack(ms).cpu().numpy())
    return results


@pytest.mark.parametrize("prop,both,shift", [("fresnel", False, True), ("asm", False, False),
                                             ("fresnel", True, False)])
def test_reference_graph_matches_oracle_and_fused_node(gpu, prop, both, shift):
    T, H, W, S = 32, 64, 64, 8
    kb, ke = O.band_bins(6e3, 15e3, 37e3, T)
    p = O.PathParams(H, W, T, S, kb, ke, prop, z_m=0.3, both_sidebands=both, fftshift=shift)
    frames, _ = synth_frames(SynthParams(H, W, 20.0, seed=5, stream_id=1), 0, T + S)
    graph = _graph_moments(frames, p)
    fused = Node("DopplerMoments", {"lambda": p.lambda_m, "dx": p.dx_m, "z": p.z_m, "time_z_begin": kb,
                                    "time_z_end": ke, "stride": S,
                                    "prop": "fresnel" if prop == "fresnel" else "angular_spectrum",
                                    "fftshift": shift, "both_sidebands": both},
                 [desc([T, H, W], "U16")])
    for w in range(2):
        ref = O.doppler_moments(frames[w * S:w * S + T], p, "f64")
        assert rel_l2(graph[w], ref) < 1e-5, f"graph window {w}"
        out = fused.execute(_u16(frames[w * S:w * S + T]))[0]
        torch.cuda.synchronize()
        assert rel_l2(out.cpu().numpy(), ref) < 1e-5, f"fused window {w}"
        assert rel_l2(out.cpu().numpy(), graph[w]) < 1e-5
