"""wasl_amd: MI355X-native engine for the wasl RKDG hot path.

The numerical work runs in hand-written HIP kernels for gfx950 behind the C
ABI declared in include/wasl_hip.h and include/wasl_hip_serve.h
(wasl_amd/libwasl_hip.so) and the case layer of include/wasl_hip_case.h
(wasl_amd/libwasl_hip_case.so); this package is the Python mirror of those
interfaces plus the synthetic mesh generator.
"""

from . import meshgen, serve  # noqa: F401
from .engine import (  # noqa: F401
    BF16, EULER, FP64, IC_EULER_REGIONS, IC_RAREFACTION, IC_SCALAR_CIRCLES, IC_SCALAR_PULSE,
    IC_SCALAR_REGIONS, IC_SCALAR_WAVE, IC_SMOOTH, IC_VORTEX, INFLOW, MIXED, OUTFLOW, PERIODIC,
    SCALAR_ADVECTION, Engine, Mesh, SimulationError, WaslHipError, fp64_ceilings, library,
)
