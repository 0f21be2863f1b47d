"""vship_amd - MI355X-native SSIMULACRA2 engine with the C ABI of libvship.

The product is libvship_amd.so (C ABI in include/VshipAPI.h and
include/vship_amd.h, gfx950 HIP kernels in csrc/). This package holds its
Python host mirror (ssimu2.py) and the synthetic-content generators used by
tests and benchmarks (synth.py).
"""
from .ssimu2 import (Colorspace, SSIMU2Handler, VshipError, DeviceBuffer, colorspace, srgb_rgb_full,  # noqa: F401
                     load_library, device_count, error_message)
