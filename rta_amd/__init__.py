"""rta_amd -- MI355X-native (gfx950) bbvh-hip: the GPU BVH path of librta.

Karras LBVH construction and closest-hit child-is traversal as hand-written
HIP kernels behind a C ABI (include/rta_hip.h), with a C++ plugin that loads
through librta's plugin contract and this thin Python host layer over the
same ABI (ctypes, no framework types).
"""
from .abi import (AABB_DTYPE, FLT_MAX, HIT_DTYPE, NODE_DTYPE, TRI_DTYPE, RtaError, call,
                  declared_functions, hip, host)
from .device import DeviceBuffer, Event, Stream, device_count, device_name, synchronize
from .bvh import (BinaryLbvh, CisShadowTracer, CisTracer, GpuBouncer, LbvhConstructor,
                  MedianConstructor, PackedBvh, RayBuffers, ShirleyRayGenerator, diffuse_bounce,
                  shadow_rays)
from . import scenes

__all__ = [
    "AABB_DTYPE", "FLT_MAX", "HIT_DTYPE", "NODE_DTYPE", "TRI_DTYPE", "RtaError", "call",
    "declared_functions", "hip", "host", "DeviceBuffer", "Event", "Stream", "device_count",
    "device_name", "synchronize", "BinaryLbvh", "CisShadowTracer", "CisTracer", "GpuBouncer",
    "LbvhConstructor", "MedianConstructor", "PackedBvh", "RayBuffers", "ShirleyRayGenerator",
    "diffuse_bounce", "shadow_rays", "scenes",
]
