"""MI355X-native HashDAG hot path: tracing, edits, hash table and GC on one GPU.

The compute runs in libhashdag_hip.so (HIP kernels for gfx950 behind the C ABI in
include/hashdag_hip.h); this package is the host-side mirror of the reference's
interface for that path.
"""
from ._lib import HashDAGError, lib  # noqa: F401  (loads the HIP library; raises if missing)
from .dag import Geometry, HashDAG, HostImage  # noqa: F401
from .scene import Scene, TerrainParams, generate_terrain, terrain_height  # noqa: F401
from .tracer import CameraView, DAGInfo, DAGTracer, ray_params  # noqa: F401
