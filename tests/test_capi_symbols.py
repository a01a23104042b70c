"""The C-ABI library loads and exports every entry point include/gandalf_gpu.h declares, and
the Python binding declares a prototype for each.  No device calls: runs on the CPU."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "gandalf_gpu.h")
LIB = os.path.join(ROOT, "gandalf_amd", "libgandalf_gpu.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gdf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "gdf_update_all_sph_properties" in names
    assert len(names) >= 25


@pytest.mark.skipif(not os.path.exists(LIB), reason="libgandalf_gpu.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from gandalf_amd import gpu
    assert set(declared_symbols()) == set(gpu.PROTOTYPES)


def test_python_field_widths_match_library_table():
    """gdf_get_field writes w*N doubles for a field of width w: the binding must size its
    buffer from the same widths the library's field table (gdf_capi.hip) gives."""
    from gandalf_amd import gpu
    src = open(os.path.join(ROOT, "gandalf_amd", "csrc", "gdf_capi.hip")).read()
    tab = re.findall(r'\{"([a-z0-9_]+)",\s*[01],\s*[A-Z0-9_]+,\s*([13])\}', src)
    assert len(tab) >= 20
    for name, w in tab:
        assert (name in gpu.FIELDS3) == (w == "3"), name
    assert set(gpu.FIELDS3) <= {n for n, _ in tab}
