"""z2d_amd: MI355X-native backend for z2d's painter.fill / painter.stroke /
compositor hot path.  Python side: ctypes bindings over the C ABI declared in
include/z2d_gpu.h (libz2d_gpu.so, built in-tree)."""
from . import abi  # noqa: F401
