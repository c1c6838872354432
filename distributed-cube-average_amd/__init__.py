"""MI355X-native Fletcher TTI time step behind the reference's backend ABI.

The product is libdc_hip.so (csrc/, built by build.py): the reference's
dc_propagate / dc_device_* entry points implemented with hand-written HIP
kernels for gfx950, plus the library's own dcx_* API.  dc_hip.py holds the
ctypes bindings; velocity_model.py the synthetic model generators of the
benchmark configurations.
"""
from . import dc_hip  # noqa: F401
