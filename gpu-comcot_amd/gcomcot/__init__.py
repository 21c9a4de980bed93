"""gcomcot — COMCOT's shallow-water step on MI355X through a C ABI.

The compute lives in libgcomcot_hip.so (HIP kernels for gfx950 plus the
reference's C entry points); this package is the host mirror used by the
tests and the benchmark: scenario generation, layer registration, the
step loop and state access.
"""
from ._lib import GcomcotError, LIB_PATH, lib
from .scenarios import LayerSpec, Scenario, linear_layer, nonlinear_layer, two_level_nest
from .engine import DeviceRun

__all__ = ["GcomcotError", "LIB_PATH", "lib", "LayerSpec", "Scenario", "linear_layer",
           "nonlinear_layer", "two_level_nest", "DeviceRun"]
