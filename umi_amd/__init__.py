"""umi_amd — MI355X-native (gfx950) UMI operator.

Python host side of libumi_hip.so: device buffers, streams and the operator
calls.  The C++ drop-in of the reference API (UMI.h / UMI_gpu.h /
SyncedMemory.h) lives in ``umi_amd/src``; the C ABI in ``include/umi_hip.h``.
"""
from ._lib import LIB_PATH, UmiError, lib  # noqa: F401
from .device import DeviceBuffer, Graph, HostBuffer, Stream, Timer, device_count, device_name, synchronize  # noqa: F401
from . import batch, ops, workloads  # noqa: F401  (umi_amd.proposer is imported on use)

__all__ = ["LIB_PATH", "UmiError", "lib", "DeviceBuffer", "Graph", "HostBuffer", "Stream", "Timer", "device_count",
           "device_name", "synchronize", "batch", "ops", "workloads"]
