"""MI355X (gfx950) per-timestep hot path of LeachingSim.

The product is the C ABI library ``liblsgpu.so`` (include/lsgpu.h); this
package holds its HIP sources (csrc/), the C++ host mirror of the reference's
Simulation API (host/), a ctypes binding (lsgpu) and the seeded geometry
generator (geometry).
"""
from .lsgpu import GpuSimulation, Params, StepReport, default_params, load_library  # noqa: F401
