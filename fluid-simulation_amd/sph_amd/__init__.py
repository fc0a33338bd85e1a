"""MI355X-native SPH particle hot path (HIP, gfx950) behind the reference's
NeighborSearch / PressureSolver / BoundaryHandling plug-in API.

``scenes`` builds seeded inputs; ``system`` mirrors the reference's API over
libsphhip.so. Importing this package does not touch the GPU; the first call
that needs the library loads it and fails loudly when it is missing.
"""
from . import scenes  # noqa: F401
from .scenes import Params, Scene  # noqa: F401


def __getattr__(name):
    if name in ("System", "SpatialHashing", "NeighborList"):
        from . import system
        return getattr(system, name)
    raise AttributeError(name)
