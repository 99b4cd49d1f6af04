"""grim_amd: MI355X (gfx950) implementation of grim's per-timestep implicit
hot path -- the batched per-zone Newton solve of timeStepper::solve and the
fluidElement residual it iterates on -- behind the C ABI of
``include/grim_hip.h``.

``grim_amd.capi`` binds that ABI with ctypes; ``grim_amd.zones`` builds the
synthetic isolated-zone batches used by the tests and the benchmark.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
