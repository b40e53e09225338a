"""MI355X-native kazen_proto hot path: BVH8 closest-hit traversal and a
wavefront path integrator in hand-written HIP behind a C ABI
(include/kp_hip.h).  This package holds the HIP sources (csrc/), the C++
adapters that mirror the reference's Accelerator/Integrator/Renderer
interfaces (cpp/), the ctypes binding (native.py) and the procedural
workloads (scenes.py)."""
from . import native, scenes  # noqa: F401

__all__ = ["native", "scenes"]
