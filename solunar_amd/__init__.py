"""MI355X (gfx950) SPH particle-fluid step with the reference FluidWorld API.

The compute path is libsolunar_fluid.so (hand-written HIP kernels behind the C
ABI in include/solunar_fluid.h); this package is its Python binding.
"""
from .types import (DFSPH, WCSPH, FluidReport, FluidSettings,  # noqa: F401
                    PhysicsError)


def FluidWorld(*args, **kwargs):  # noqa: N802 - mirrors the reference type name
    from .fluid import FluidWorld as _World

    return _World(*args, **kwargs)


def FluidBatch(*args, **kwargs):  # noqa: N802
    from .fluid import FluidBatch as _Batch

    return _Batch(*args, **kwargs)
