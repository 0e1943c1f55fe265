"""rtpg: MI355X-native path tracing and path guiding (RTLib tracer path).

The compute path is librtpg.so (host C++ + gfx950 HIP kernels) behind the C
ABI in include/rtpg.h; this package binds it with ctypes and mirrors the
reference's C++ class names.
"""
from ._lib import RtpgError, load  # noqa: F401

__all__ = ["RtpgError", "load"]
