"""Node-kind inference and validation through the node ABI (no GPU needed: infer only).

Cases follow the reference's *InferTest fixtures (test/holonp/*_test.cu,
test/holotask/*_test.cu) for the kinds the Doppler-moment graph wires
(graph_builder.cc:232-305, 522-566, 578-647)."""
import pytest

from holoflow_amd import nodes
from holoflow_amd._lib import HfxInvalidArgument, HfxError

D = nodes.desc


def test_registry_lists_path_kinds():
    k = set(nodes.kinds())
    assert {"BatchQueue", "Conversion", "Reshape", "Slice", "Copy", "Concatenate", "RFFT", "FFT",
            "FresnelDiffraction", "AngularSpectrum", "MeanAbs", "Mean", "Abs", "Arange", "Multiply",
            "FFTShiftNp", "DopplerMoments"} <= k


def test_unknown_kind_is_invalid():
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("NoSuchKind", {}, [])


# ---------------------------------------------------------------- Conversion
def test_conversion_u16_to_f32():
    r = nodes.infer("Conversion", {"target": "F32", "strategy": "Real"}, [D([32, 8, 8], "U16")])
    assert r["output_descs"][0]["shape"] == [32, 8, 8] and r["output_descs"][0]["dtype"] == "F32"
    assert r["kind"] == "Sync" and r["in_place"] == []


def test_conversion_rejects_host_and_other_strategies():
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("Conversion", {"target": "F32", "strategy": "Real"}, [D([4], "U16", mem_loc="Host")])
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("Conversion", {"target": "U8", "strategy": "Scaled"}, [D([4], "U16")])


# ---------------------------------------------------------------- Reshape (reshape_test.cu:88-114)
def test_reshape_view_path_for_contiguous_input():
    r = nodes.infer("Reshape", {"shape": [1, 32, 8, 8], "copy": False}, [D([32, 8, 8], "F32")])
    od = r["output_descs"][0]
    assert od["shape"] == [1, 32, 8, 8] and r["in_place"] == [[0, 0]]
    assert od["strides"][1:] == [256, 32, 4]


def test_reshape_copy_path_when_forced():
    r = nodes.infer("Reshape", {"shape": [-1], "copy": True}, [D([4, 6], "F32")])
    assert r["output_descs"][0]["shape"] == [24] and r["in_place"] == []


def test_reshape_rejects_invalid_shape_spec():
    for bad in ([-1, -1], [5, 5], [-2, 12]):
        with pytest.raises(HfxInvalidArgument):
            nodes.infer("Reshape", {"shape": bad}, [D([4, 6], "F32")])


def test_reshape_view_of_strided_input_or_copy():
    # [4, 6] sliced to columns 0..3 (row stride 24 B): merging the axes needs a copy
    sl = D([4, 3], "F32", strides=[24, 4])
    r = nodes.infer("Reshape", {"shape": [12]}, [sl])
    assert r["in_place"] == []
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("Reshape", {"shape": [12], "copy": False}, [sl])
    # splitting an axis of a strided view stays a view
    r = nodes.infer("Reshape", {"shape": [2, 2, 3]}, [sl])
    assert r["in_place"] == [[0, 0]] and r["output_descs"][0]["strides"] == [48, 24, 4]


# ---------------------------------------------------------------- Slice (slice_test.cu:85-120)
def test_slice_mixed_items_shape_strides_offset():
    d = D([2, 5, 6], "CF32")
    r = nodes.infer("Slice", {"slices": [1, {"start": 1, "stop": 5, "step": 2}, {}]}, [d])
    od = r["output_descs"][0]
    assert od["shape"] == [2, 6]
    assert od["strides"] == [2 * 48, 8]
    assert od["offset"] == 1 * 240 + 1 * 48
    assert r["in_place"] == [[0, 0]]


def test_slice_band_of_time_spectrum():
    # graph_builder.cc:266: FH[:, z0:z1, :, :] of an rfft output [8, 17, H, W]
    d = D([8, 17, 4, 4], "CF32")
    r = nodes.infer("Slice", {"slices": [{}, {"start": 3, "stop": 9}, {}, {}]}, [d])
    od = r["output_descs"][0]
    assert od["shape"] == [8, 6, 4, 4] and od["offset"] == 3 * 128


def test_slice_rejects_out_of_range_index_and_count():
    with pytest.raises(HfxError):
        nodes.infer("Slice", {"slices": [7, {}]}, [D([5, 6], "F32")])
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("Slice", {"slices": [{}]}, [D([5, 6], "F32")])
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("Slice", {"slices": [{"step": -1}, {}]}, [D([5, 6], "F32")])


# ---------------------------------------------------------------- Concatenate (concatenate_test.cu:118-190)
def test_concatenate_axes():
    a, b = D([2, 3], "F32"), D([4, 3], "F32")
    assert nodes.infer("Concatenate", {}, [a, b])["output_descs"][0]["shape"] == [6, 3]
    c, e = D([2, 3], "F32"), D([2, 5], "F32")
    assert nodes.infer("Concatenate", {"axis": 1}, [c, e])["output_descs"][0]["shape"] == [2, 8]
    assert nodes.infer("Concatenate", {"axis": -1}, [c, e])["output_descs"][0]["shape"] == [2, 8]
    assert nodes.infer("Concatenate", {"axis": None}, [a, e])["output_descs"][0]["shape"] == [16]


def test_concatenate_rejections():
    a = D([2, 3], "F32")
    for ins, st in (([], {}), ([D([2, 3], "F32", mem_loc="Host")], {}), ([a, D([2, 3], "CF32")], {}),
                    ([a, a], {"axis": 2}), ([a, D([2, 3, 1], "F32")], {"axis": 0}),
                    ([a, D([2, 4], "F32")], {"axis": 0})):
        with pytest.raises(HfxInvalidArgument):
            nodes.infer("Concatenate", st, ins)


def test_concatenate_accepts_non_contiguous_input():
    sl = D([2, 3], "F32", strides=[24, 4], offset=8)
    assert nodes.infer("Concatenate", {"axis": 1}, [sl, sl])["output_descs"][0]["shape"] == [2, 6]


# ---------------------------------------------------------------- RFFT / FFT
def test_rfft_and_fft_shapes():
    x = D([8, 32, 4, 4], "F32")
    assert nodes.infer("RFFT", {"axis": 1}, [x])["output_descs"][0]["shape"] == [8, 17, 4, 4]
    r = nodes.infer("FFT", {"axis": 1, "norm": "ortho"}, [x])["output_descs"][0]
    assert r["shape"] == [8, 32, 4, 4] and r["dtype"] == "CF32"
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("RFFT", {"axis": 4}, [x])
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("FFT", {"axis": 1, "norm": "bogus"}, [x])


# ---------------------------------------------------------------- Fresnel (fresnel_diffraction_test.cu)
FR = {"lambda": 852e-9, "dx": 12e-6, "dy": 12e-6, "z": 0.3, "output_magnitude": True}


def test_fresnel_output_descriptor():
    r = nodes.infer("FresnelDiffraction", FR, [D([8, 6, 16, 16], "CF32")])
    assert r["output_descs"][0]["dtype"] == "F32" and r["output_descs"][0]["shape"] == [8, 6, 16, 16]
    r = nodes.infer("FresnelDiffraction", dict(FR, output_magnitude=False), [D([16, 16], "F32")])
    assert r["output_descs"][0]["dtype"] == "CF32"


@pytest.mark.parametrize("bad", [{"dx": 11e-6}, {"lambda": 0.0}, {"dy": -1.0}, {"axes": [-1, -1]},
                                 {"axes": [0]}])
def test_fresnel_rejections(bad):
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("FresnelDiffraction", dict(FR, **bad), [D([2, 16, 16], "CF32")])


def test_fresnel_rejects_rank1_u16_host():
    for d in (D([16], "CF32"), D([16, 16], "U16"), D([16, 16], "CF32", mem_loc="Host")):
        with pytest.raises(HfxInvalidArgument):
            nodes.infer("FresnelDiffraction", FR, [d])


# ---------------------------------------------------------------- AngularSpectrum (angular_spectrum_test.cu)
AS = {"lambda": 852e-9, "dx": 12e-6, "dy": 12e-6, "z": 0.3}


def test_asm_in_place_and_padding():
    r = nodes.infer("AngularSpectrum", AS, [D([3, 16, 16], "CF32")])
    assert r["in_place"] == [[0, 0]] and r["output_descs"][0]["shape"] == [3, 16, 16]
    r = nodes.infer("AngularSpectrum", dict(AS, padding={"width": 24, "height": 20}), [D([3, 16, 16], "CF32")])
    assert r["in_place"] == [] and r["output_descs"][0]["shape"] == [3, 20, 24]
    r = nodes.infer("AngularSpectrum", AS, [D([16, 16], "F32")])
    assert r["in_place"] == [] and r["output_descs"][0]["dtype"] == "CF32"


@pytest.mark.parametrize("pad", [{"width": 15, "height": 16}, {"width": 17, "height": 16},
                                 {"width": 0, "height": 16}])
def test_asm_rejects_bad_padding(pad):
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("AngularSpectrum", dict(AS, padding=pad), [D([16, 16], "CF32")])


# ---------------------------------------------------------------- reductions
def test_mean_and_mean_abs_keepdims():
    x = D([8, 6, 4, 4], "F32")
    r = nodes.infer("Mean", {"axis": [-3], "keepdims": True}, [x])
    assert r["output_descs"][0]["shape"] == [8, 1, 4, 4]
    r = nodes.infer("Mean", {"axis": [0], "keepdims": False}, [D([8, 1, 4, 4], "F32")])
    assert r["output_descs"][0]["shape"] == [1, 4, 4]
    r = nodes.infer("MeanAbs", {"axis": [-3], "keepdims": True}, [D([8, 6, 4, 4], "CF32")])
    assert r["output_descs"][0]["shape"] == [8, 1, 4, 4] and r["output_descs"][0]["dtype"] == "F32"
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("Mean", {"axis": [5]}, [x])


def test_abs_output():
    r = nodes.infer("Abs", {}, [D([2, 3], "CF32")])
    assert r["output_descs"][0]["dtype"] == "F32" and r["output_descs"][0]["shape"] == [2, 3]


# ---------------------------------------------------------------- Arange (arange_test.cu:127-205)
def test_arange_shapes():
    assert nodes.infer("Arange", {"start": 0, "stop": 5, "step": 1}, [])["output_descs"][0] == D([5], "F32")
    assert nodes.infer("Arange", {"start": 0, "stop": 5, "step": 1, "dtype": "U8"}, [])["output_descs"][0]["dtype"] == "U8"
    assert nodes.infer("Arange", {"start": 0, "stop": 1, "step": 0.3}, [])["output_descs"][0]["shape"] == [4]
    assert nodes.infer("Arange", {"start": 5, "stop": 0, "step": -2}, [])["output_descs"][0]["shape"] == [3]
    assert nodes.infer("Arange", {"start": 5, "stop": 5, "step": 1}, [])["output_descs"][0]["shape"] == [0]
    assert nodes.infer("Arange", {"start": 0, "stop": 1e7, "step": 1}, [])["output_descs"][0]["shape"] == [10_000_000]


def test_arange_rejections():
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("Arange", {"start": 0, "stop": 5, "step": 1}, [D([1], "F32")])
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("Arange", {"start": 0, "stop": 5, "step": 0}, [])
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("Arange", {"start": 0, "stop": 5, "step": 1, "device": "Host"}, [])


# ---------------------------------------------------------------- Multiply / FFTShift
def test_multiply_broadcast_and_promotion():
    r = nodes.infer("Multiply", {}, [D([1, 6, 1, 1], "F32"), D([8, 6, 4, 4], "F32")])
    assert r["output_descs"][0]["shape"] == [8, 6, 4, 4] and r["output_descs"][0]["dtype"] == "F32"
    r = nodes.infer("Multiply", {}, [D([4, 4], "CF32"), D([4], "F32")])
    assert r["output_descs"][0]["dtype"] == "CF32"
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("Multiply", {}, [D([4], "F32")])
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("Multiply", {}, [D([4], "F32", mem_loc="Host"), D([4], "F32")])


def test_fftshift_keeps_descriptor_and_rejects_duplicate_axes():
    d = D([8, 1, 4, 6], "F32")
    assert nodes.infer("FFTShiftNp", {"axes": [-2, -1]}, [d])["output_descs"][0] == d
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("FFTShiftNp", {"axes": [-1, 3]}, [d])


# ---------------------------------------------------------------- BatchQueue (batch_queue.cc:190-222)
def test_batch_queue_infer():
    r = nodes.infer("BatchQueue", {"target_capacity": 16, "output_size": 8, "output_stride": 8},
                    [D([1, 32, 4, 4], "F32")])
    assert r["kind"] == "Async" and r["output_descs"][0]["shape"] == [8, 32, 4, 4]
    assert r["owned_inputs"] == [True] and r["owned_outputs"] == [True]


@pytest.mark.parametrize("st", [{"target_capacity": 0, "output_size": 1, "output_stride": 1},
                                {"target_capacity": 4, "output_size": 0, "output_stride": 1},
                                {"target_capacity": 4, "output_size": 2, "output_stride": 3}])
def test_batch_queue_rejections(st):
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("BatchQueue", st, [D([1, 4], "F32")])
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("BatchQueue", {"target_capacity": 4, "output_size": 1, "output_stride": 1},
                    [D([2, 3], "F32", strides=[24, 4])])


# ---------------------------------------------------------------- DopplerMoments (fused)
def test_doppler_moments_infer():
    st = {"lambda": 852e-9, "dx": 12e-6, "z": 0.3, "time_z_begin": 6, "time_z_end": 13}
    r = nodes.infer("DopplerMoments", st, [D([32, 64, 64], "U16")])
    assert r["output_descs"][0]["shape"] == [3, 64, 64]
    r = nodes.infer("DopplerMoments", dict(st, split_bin=9), [D([1, 32, 64, 64], "U16")])
    assert r["output_descs"][0]["shape"] == [5, 64, 64]
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("DopplerMoments", dict(st, prop="bogus"), [D([32, 64, 64], "U16")])
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("DopplerMoments", st, [D([32, 64, 64], "F32")])


# ---------------------------------------------------------------- display chain (next to the path)
@pytest.mark.parametrize("dt,target,strategy,out", [
    ("U8", "F32", "Real", "F32"), ("U8", "CF32", "Real", "CF32"), ("U16", "CF32", "Real", "CF32"),
    ("U16", "F32", "Real", "F32"), ("F32", "U8", "Scaled", "U8"), ("F32", "U16", "Scaled", "U16"),
    ("F32", "CF32", "Real", "CF32"), ("CF32", "F32", "Modulus", "F32"), ("CF32", "F32", "Argument", "F32")])
def test_conversion_table(dt, target, strategy, out):
    r = nodes.infer("Conversion", {"target": target, "strategy": strategy}, [D([2, 3], dt)])
    assert r["output_descs"][0]["dtype"] == out and r["output_descs"][0]["shape"] == [2, 3]


@pytest.mark.parametrize("dt,target,strategy", [("F32", "U8", "Real"), ("CF32", "U8", "Scaled"),
                                                ("U16", "U8", "Scaled"), ("F32", "F32", "Real")])
def test_conversion_rejects_unsupported(dt, target, strategy):
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("Conversion", {"target": target, "strategy": strategy}, [D([4], dt)])
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("Conversion", {"target": "X", "strategy": "Real"}, [D([4], "U8")])


def test_sliding_average_infer():
    st = {"target_capacity": 4, "window_size": 3}
    r = nodes.infer("SlidingAverage", st, [D([1, 8, 8], "F32")])
    assert r["kind"] == "Async" and r["output_descs"][0]["shape"] == [1, 8, 8]
    r = nodes.infer("SlidingAverage", st, [D([1, 8, 8], "F32"), D([1], "U8", mem_loc="Host")])
    assert r["owned_inputs"] == [True, False]
    for bad in ([D([2, 8, 8], "F32")], [D([1, 8, 8], "U16")], [D([8, 8], "F32")],
                [D([1, 8, 8], "F32"), D([1], "U8")]):
        with pytest.raises(HfxInvalidArgument):
            nodes.infer("SlidingAverage", st, bad)
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("SlidingAverage", {"target_capacity": 0, "window_size": 3}, [D([1, 8, 8], "F32")])


def test_dual_reader_batch_queue_infer():
    r = nodes.infer("DualReaderBatchQueue", {"target_capacity": 4, "window_size": 5}, [D([8, 3, 4, 4], "CF32")])
    o = r["output_descs"]
    assert [d["shape"] for d in o] == [[1, 3, 4, 4], [1, 3, 4, 4], [1]]
    assert o[2]["dtype"] == "U8" and o[2]["mem_loc"] == "Host"
    assert r["owned_outputs"] == [True, True, False]
    with pytest.raises(HfxInvalidArgument):
        nodes.infer("DualReaderBatchQueue", {"target_capacity": 0, "window_size": 5}, [D([8, 3], "F32")])


# ---------------------------------------------------------------- Filter2D / Flatfield (filter2d_test.cu, flatfield_test.cu)
F2D = {"r_inner": 0, "r_outer": 100, "s_inner": 0, "s_outer": 0}


def test_filter2d_infer_real_and_complex():
    r = nodes.infer("Filter2D", F2D, [D([2, 2], "F32")])
    assert r["kind"] == "Sync" and r["in_place"] == []
    assert r["output_descs"][0]["shape"] == [2, 2] and r["output_descs"][0]["dtype"] == "F32"
    r = nodes.infer("Filter2D", F2D, [D([3, 2, 2], "CF32")])
    assert r["in_place"] == [[0, 0]] and r["output_descs"][0]["dtype"] == "CF32"


@pytest.mark.parametrize("inp,st,msg", [
    (D([4], "F32"), F2D, "rank 2 or higher"),
    (D([4, 4], "U16"), F2D, "complex float32 or real float32"),
    (D([4, 4], "F32", mem_loc="Host"), F2D, "device memory"),
    (D([4, 4], "F32", strides=[4, 16]), F2D, "C-contiguous"),
    (D([4, 4], "F32"), dict(F2D, r_inner=-1), "r_inner must be non-negative"),
    (D([4, 4], "F32"), dict(F2D, s_outer=-2), "s_outer must be non-negative"),
])
def test_filter2d_rejections(inp, st, msg):
    with pytest.raises(HfxInvalidArgument, match=msg):
        nodes.infer("Filter2D", st, [inp])


def test_flatfield_infer():
    r = nodes.infer("Flatfield", {"sigma_y": 1.0, "sigma_x": 1.0}, [D([2, 3, 4], "F32")])
    assert r["kind"] == "Sync" and r["in_place"] == []
    assert r["output_descs"][0]["shape"] == [2, 3, 4] and r["output_descs"][0]["dtype"] == "F32"


@pytest.mark.parametrize("inp,st,msg", [
    (D([2, 2], "F32"), {"sigma_y": 0.0, "sigma_x": 1.0}, "sigma_y must be positive"),
    (D([2, 2], "F32"), {"sigma_y": 1.0, "sigma_x": 0.0}, "sigma_x must be positive"),
    (D([2, 2], "CF32"), {"sigma_y": 1.0, "sigma_x": 1.0}, "input must be float32"),
    (D([0, 2], "F32"), {"sigma_y": 1.0, "sigma_x": 1.0}, "must not be empty"),
])
def test_flatfield_rejections(inp, st, msg):
    with pytest.raises(HfxInvalidArgument, match=msg):
        nodes.infer("Flatfield", st, [inp])
