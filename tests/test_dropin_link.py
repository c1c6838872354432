"""The C drop-in: the reference's own program links against libdc_hip.so.

oracle/Makefile's `dropin` target compiles the reference's C sources with its
own headers -- every file its Makefile builds except the backend pair
openmp_propagate.c and device_data.c (Makefile:27-58) -- and links them
against libdc_hip.so with --no-undefined.  The link must succeed, and the
program's only unresolved backend symbols must be exactly the eight the
reference's worker calls (include/propagate.h:13-19,
include/device_data.h:17-37), all exported by the library.  The binary is
never run here (no GPU) nor on a GPU.  Needs the reference sources, so it is
skipped where they are absent.
"""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
LIB = os.path.join(REPO, "distributed-cube-average_amd", "libdc_hip.so")
BIN = os.path.join(REPO, "oracle", "_ref", "dc_hip_backend")
BACKEND = {"dc_propagate", "dc_device_data_init", "dc_device_data_free",
           "dc_device_data_get_results", "dc_device_swap_arrays",
           "dc_device_add_source", "dc_device_extract_halo_face",
           "dc_device_insert_halo_face"}


def _symbols(args):
    out = subprocess.run(["nm", *args], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources absent")
@pytest.mark.skipif(not os.path.exists(LIB), reason="libdc_hip.so not built")
def test_reference_program_links_against_the_library():
    r = subprocess.run(["make", "-s", "-B", "-C", os.path.join(REPO, "oracle"),
                        "dropin"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    undefined = _symbols(["-u", BIN])
    backend_calls = {s for s in undefined if s.startswith("dc_")}
    assert backend_calls == BACKEND
    exported = _symbols(["-D", "--defined-only", LIB])
    assert BACKEND <= exported
    needed = subprocess.run(["readelf", "-d", BIN], capture_output=True,
                            text=True, check=True).stdout
    assert "[libdc_hip.so]" in needed
