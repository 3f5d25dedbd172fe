"""The C-ABI library loads and exports every symbol include/*.h declares (no GPU)."""
import ctypes
import os
import re

import pytest

from elas3dxtal import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(e3x_[a-z0-9_]+)\s*\(", txt)))


def test_core_library_exports_header_symbols():
    lib = ctypes.CDLL(L.LIB_PATH)
    names = declared("e3x.h")
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.SIGNATURES), "binding and header disagree"


def test_serving_library_exports_header_symbols():
    lib = ctypes.CDLL(L.LIB_PATH)
    names = declared("e3x_serve.h")
    assert len(names) >= 8
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.SERVE_SIGNATURES), "binding and header disagree"


def test_request_response_layout_matches_header():
    """The ctypes structs have the C layout (offsets of the last members)."""
    assert ctypes.sizeof(L.Request) == 4 * 4 + 2 * 8 + 3 * 8 + 4 * 8 + 4 + 4 + 6 * 8 + 4 + 4 + 4 + 4 + 2 * 8
    assert L.Response.t_done.offset == ctypes.sizeof(L.Response) - 8


def test_version_and_error_string_without_gpu():
    lib = L.load()
    assert lib.e3x_version() >= 10000
    assert isinstance(lib.e3x_last_error(), bytes)


def test_one_node_thick_grid_is_rejected_without_gpu():
    """e3x_create checks the extents before it touches a device (include/e3x.h)."""
    from elas3dxtal.solver import Elas3D
    for shape in [(1, 4, 4), (4, 1, 4), (4, 4, 1)]:
        with pytest.raises(L.E3xError) as e:
            Elas3D(shape, 2)
        assert e.value.code == L.E3X_EINVAL


def test_grid_beyond_int32_nodes_is_rejected_without_gpu():
    """Node indices inside the kernels are 32-bit: e3x_create refuses 2^31 nodes or more."""
    from elas3dxtal.solver import Elas3D
    with pytest.raises(L.E3xError) as e:
        Elas3D((1291, 1291, 1291), 2)
    assert e.value.code == L.E3X_EINVAL


def test_stop_rule_needs_a_handle_without_gpu():
    lib = L.load()
    assert lib.e3x_set_stop_rule(None, L.STOP_NIST_ABS) == L.E3X_EINVAL
    assert b"NULL" in lib.e3x_last_error()
