"""libdc_hip.so loads without a GPU and exports every function that the
headers in include/ declare (the reference's backend ABI plus dcx_*)."""
import ctypes
import os
import re

from conftest import REPO


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(REPO, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(REPO, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(dc[x]?_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_headers_declare_the_reference_abi():
    names = declared_functions()
    for ref in ("dc_propagate", "dc_device_data_init", "dc_device_data_free",
                "dc_device_data_get_results", "dc_device_swap_arrays",
                "dc_device_add_source", "dc_device_extract_halo_face",
                "dc_device_insert_halo_face"):
        assert ref in names


def test_library_exports_every_declared_symbol(H):
    lib = ctypes.CDLL(H.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_every_declared_symbol(H):
    assert declared_functions() <= set(H.SIGNATURES)


def test_struct_layouts_match_the_reference_headers(H):
    P = ctypes.sizeof(ctypes.c_void_p)
    assert ctypes.sizeof(H.PrecompVars) == 10 * P
    assert ctypes.sizeof(H.DeviceDataStruct) == 16 * P
    # dc_process_t: ints, floats and size_t/pointers in reference order.
    assert H.Process.sizes.offset == 4 * (1 + 3 + 27 + 3 + 1 + 1 + 4)
    assert H.Process.anisotropy_vars.offset == H.Process.sizes.offset + 3 * 8
    assert H.Process.num_workers.offset == \
        H.Process.hostnames.offset + P
