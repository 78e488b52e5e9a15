"""wasl_amd: MI355X-native engine for the wasl RKDG hot path.

The numerical work runs in hand-written HIP kernels for gfx950 behind the C
ABI declared in include/wasl_hip.h (wasl_amd/libwasl_hip.so); this package is
the Python mirror of that interface plus the synthetic mesh generator.
"""

from . import meshgen  # noqa: F401
from .engine import (  # noqa: F401
    BF16, EULER, FP64, IC_EULER_REGIONS, IC_RAREFACTION, IC_SCALAR_PULSE, IC_SCALAR_REGIONS,
    IC_SMOOTH, IC_VORTEX, INFLOW, MIXED, OUTFLOW, PERIODIC, SCALAR_ADVECTION, Engine, Mesh,
    SimulationError, WaslHipError, library,
)
