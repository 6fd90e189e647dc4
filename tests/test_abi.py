"""CPU checks of the boundary: librfx_hip.so loads, exports every entry point
include/rfx_ra.h declares, and the Python binding's signature table covers
the same names.  No compute calls (there is no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rfx_ra.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:const\s+)?[\w\*]+\s+\**(\w+)\s*\(", text, re.M))


def test_header_declares_both_layers():
    names = declared()
    for n in ("rfx_sort_pairs", "rfx_join_find_bounds", "rfx_join_main", "rfx_reduce_by_key_f64",
              "rfx_set_op", "rfx_select", "rfx_q1", "set_b40csortpair", "mgpu_join_find_bounds",
              "mgpu_join_main", "set_reduce", "set_union", "set_unique", "rfx_harmony_call"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import redfox_amd

    lib = ctypes.CDLL(redfox_amd.LIB_PATH)
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_matches_header():
    from redfox_amd._lib import SIGNATURES

    assert set(SIGNATURES) == declared()


def test_status_strings_without_gpu():
    import redfox_amd

    L = redfox_amd.lib()
    assert L.rfx_strerror(0) == b"ok"
    assert L.rfx_strerror(-1) == b"invalid argument"
    assert L.rfx_version().startswith(b"rfx")


def test_argument_errors_are_reported_not_raised():
    """Bad widths are rejected before any device work."""
    import redfox_amd

    L = redfox_amd.lib()
    assert L.rfx_sort_pairs(None, None, None, 10, 4, 32, 4) == -1
    assert L.rfx_set_op(None, 0, None, None, None, 0, None, 0, 4) == -1


def test_missing_library_fails_loudly(tmp_path):
    from redfox_amd._lib import load

    with pytest.raises(ImportError):
        load(str(tmp_path / "nope.so"))


def test_context_methods_are_plain_methods():
    """The binding's op wrappers are callables (a stray decorator would turn
    one into a property); only workspace_bytes is a property."""
    import inspect

    from redfox_amd.rfx import Context

    for name, member in inspect.getmembers(Context):
        if name.startswith("_"):
            continue
        raw = inspect.getattr_static(Context, name)
        if name == "workspace_bytes":
            assert isinstance(raw, property)
        else:
            assert inspect.isfunction(raw), name
