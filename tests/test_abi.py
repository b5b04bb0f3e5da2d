"""The C-ABI library loads without a GPU and exports every symbol include/*.h declares.

No compute is called here (this runs on CPU-only machines)."""
import ctypes
import os
import re

from holoflow_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


HEADERS = ("hfx.h", "hfx_nodes.h", "hfx_serve.h", "hfx_holofile.h", "hfx_graph.h")


def _declared():
    names = set()
    for h in HEADERS:
        with open(os.path.join(ROOT, "include", h)) as f:
            text = f.read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)   # declarations only, not comments
        names |= set(re.findall(r"\b(hfx_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 20
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert _declared() == set(_lib.SIGNATURES)


def test_binding_loads_and_reports_version():
    L = _lib.lib()
    assert L.hfx_abi_version() == 1
    # an invalid call (no device work) reports its reason through hfx_last_error
    assert L.hfx_plan_create(None, None, None) == _lib.HFX_E_INVALID
    assert L.hfx_last_error() == b"null argument"


def test_descriptor_defaults_are_reference_compatible():
    d = _lib.PlanDesc()
    _lib.lib().hfx_plan_desc_init(ctypes.byref(d))
    assert d.weight_fs_hz == 37e3          # graph_builder.cc:580
    assert d.precision == _lib.PREC_FP32 and d.flags == 0
