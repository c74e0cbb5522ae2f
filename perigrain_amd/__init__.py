"""perigrain_amd -- MI355X (gfx950) PeriDEM time stepper.

A drop-in for the time loop of debdeepbh/perigrain (run_timeloop,
src/particle/timeloop.h:500-1015): hand-written HIP kernels behind the
libpg C ABI (include/perigrain.h), driven from Python through ctypes or
from the C++ drop-in executables.
"""
from .setups import Params, Setup  # noqa: F401
from .stepper import Stepper  # noqa: F401

__all__ = ["Params", "Setup", "Stepper"]
