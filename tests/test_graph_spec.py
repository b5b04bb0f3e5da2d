"""GraphSpec JSON, dot output and compile-time validation of the runtime (on the CPU: every
check here fails before the compiler allocates anything).

Cases follow the reference's graph_spec.cc:145-215 checks and compiler.cc:381-475,
771-785 (unique names, registered kinds, one edge per input, no cycles, no consecutive
asynchronous nodes, factory inference errors)."""
import ctypes as C
import json

import numpy as np
import pytest

from holoflow_amd import graph
from holoflow_amd._lib import HfxInvalidArgument, lib
from oracle import holofile as OH


def _compile_error(spec):
    h = C.c_void_p()
    rc = lib().hfx_graph_compile(json.dumps(spec).encode() if not isinstance(spec, str) else spec.encode(),
                                 None, C.byref(h))
    assert rc != 0 and not h.value
    return lib().hfx_last_error().decode()


def _node(kind, **params):
    return {"type": kind, "params": params}


@pytest.fixture
def holo(tmp_path):
    p = tmp_path / "g.holo"
    p.write_bytes(OH.pack(np.zeros((8, 8, 8), np.uint16)))
    return str(p)


@pytest.mark.parametrize("text,msg", [
    ("[]", "root must be an object"),
    ('{"edges": []}', "missing required key 'nodes'"),
    ('{"nodes": []}', "'nodes' must be an object"),
    ('{"nodes": {"a": 1}}', "node 'a' must be an object"),
    ('{"nodes": {"a": {"params": {}}}}', "node 'a': missing 'type'"),
    ('{"nodes": {"a": {"type": 3, "params": {}}}}', "node 'a': 'type' must be a string"),
    ('{"nodes": {"a": {"type": "Copy"}}}', "node 'a': missing 'params'"),
    ('{"nodes": {"a": {"type": "Copy", "params": {}, "debug": 1}}}', "node 'a': 'debug' must be boolean"),
    ('{"nodes": {}, "edges": {}}', "'edges' must be an array if present"),
    ('{"nodes": {"a": {"type": "Copy", "params": {}}}, "edges": [{"to": "a", "out": 0, "in": 0}]}',
     "edge\\[0\\]: missing 'from'"),
    ('{"nodes": {"a": {"type": "Copy", "params": {}}}, "edges": [{"from": "b", "to": "a", "out": 0, "in": 0}]}',
     "edge\\[0\\]: unknown 'from' node 'b'"),
    ('{"nodes": {"a": {"type": "Copy", "params": {}}}, "edges": [{"from": "a", "to": "a", "out": 0.5, "in": 0}]}',
     "edge\\[0\\]: 'out' must be int"),
])
def test_graph_json_errors(text, msg):
    with pytest.raises(HfxInvalidArgument, match="GraphSpec JSON: " + msg):
        graph.to_dot(json.loads(text))


def test_dot_numbers_vertices_in_sorted_name_order():
    spec = {"nodes": {"b": _node("Copy"), "a": _node("Arange", start=0, stop=4, step=1)},
            "edges": [{"from": "a", "to": "b", "out": 0, "in": 0}]}
    dot = graph.to_dot(spec)
    assert dot.startswith("digraph holoflow_graph {")
    assert 'v0 [label="a\\n(Arange)' in dot and 'v1 [label="b\\n(Copy)' in dot
    assert 'v0 -> v1 [taillabel="0" headlabel="0"];' in dot


def test_compile_rejects_unknown_kind_and_double_input():
    assert "Unknown node kind 'Nope'" in _compile_error({"nodes": {"a": _node("Nope")}})
    spec = {"nodes": {"a": _node("Arange", start=0, stop=4, step=1), "b": _node("Arange", start=0, stop=4, step=1),
                      "c": _node("Copy")},
            "edges": [{"from": "a", "to": "c", "out": 0, "in": 0}, {"from": "b", "to": "c", "out": 0, "in": 0}]}
    assert "Multiple edges targeting: c:0" in _compile_error(spec)


def test_compile_rejects_cycles_and_bad_indices():
    spec = {"nodes": {"a": _node("Copy"), "b": _node("Copy")},
            "edges": [{"from": "a", "to": "b", "out": 0, "in": 0}, {"from": "b", "to": "a", "out": 0, "in": 0}]}
    assert "cycle" in _compile_error(spec)
    spec = {"nodes": {"a": _node("Arange", start=0, stop=4, step=1), "c": _node("Copy")},
            "edges": [{"from": "a", "to": "c", "out": 1, "in": 0}]}
    assert "Output index out of bounds" in _compile_error(spec)


def test_compile_reports_inference_errors_with_the_node(holo):
    spec = {"nodes": {"src": _node("Holofile", path=holo, load_kind="Live", start_frame=0, end_frame=9, batch_size=2)}}
    assert "Node 'src'" in _compile_error(spec) and "end_frame exceeds total frames" in _compile_error(spec)


def test_compile_rejects_consecutive_async_nodes(holo):
    spec = {"nodes": {"src": _node("Holofile", path=holo, load_kind="Live", start_frame=0, end_frame=8, batch_size=2),
                      "cp": _node("Memcpy", target="Device"),
                      "q1": _node("BatchQueue", target_capacity=4, output_size=2, output_stride=2),
                      "q2": _node("BatchQueue", target_capacity=4, output_size=2, output_stride=2)},
            "edges": [{"from": "src", "to": "cp", "out": 0, "in": 0}, {"from": "cp", "to": "q1", "out": 0, "in": 0},
                      {"from": "q1", "to": "q2", "out": 0, "in": 0}]}
    assert "Consecutive asynchronous nodes 'q1' and 'q2' are not supported" in _compile_error(spec)


def test_doppler_graph_builder_matches_reference_chain(holo):
    s = graph.doppler_graph(holo, 8, 8, 4, 1, 3, 8, moment="M2", fftshift=True, flatfield_sigma=(2.0, 2.0))
    kinds = sorted(n["type"] for n in s["nodes"].values())
    assert kinds == sorted(["Holofile", "Memcpy", "BatchQueue", "Conversion", "Reshape", "BatchQueue", "RFFT",
                            "Slice", "Copy", "DualReaderBatchQueue", "FresnelDiffraction", "Arange", "Reshape",
                            "Multiply", "Multiply", "Mean", "FFTShiftNp", "Flatfield", "Mean", "SlidingAverage",
                            "Conversion"])
    # every edge names existing nodes and every input index is connected once
    seen = set()
    for e in s["edges"]:
        assert e["from"] in s["nodes"] and e["to"] in s["nodes"]
        assert (e["to"], e["in"]) not in seen
        seen.add((e["to"], e["in"]))
    assert "digraph" in graph.to_dot(s)
