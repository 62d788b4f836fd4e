"""yaoray_amd — MI355X (gfx950) wavefront ray-casting / path-tracing backend.

The device path is hand-written HIP behind the C ABI in ``include/yrh.h``
(``yaoray_amd/lib/libyrh.so``); scenes, the SAH builder and workloads live in
the C++20 host library (``yaoray_amd/lib/libyaoray_host.so``).
"""
from .api import Device, DeviceBuffer, Scene, YrhError, random_rays  # noqa: F401
