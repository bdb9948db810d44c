"""ufo_amd — MI355X-native particle tracking with the API of UFO.

``import ufo_amd as ufo`` gives the reference's public names (ufo/__init__.py):
lattice and element classes, the simulation drivers, flags and the
module-level slice defaults that element constructors read.
"""
from .flags import LINEAR, FIVED, EXACT, KICK, RADIATION, DOUBLE_PRECISION, ACHROMATIC
from .constants import *  # noqa: F401,F403
from .elements import (Element, Marker, Drift, Multipole, Quadrupole, Sbend, Rbend, Sextupole,
                       Octupole, Wire, Aperture, CodeElement)
from .lattice import Lattice, Line, Beam, dump
from .mad import load as load_mad
from .radiation import emittance, RadiationIntegrals
from .chroma import chromaticity
from .rdt import rdt

DEFAULT_QUADRUPOLE_SLICES = 8
DEFAULT_BEND_SLICES = 8
DEFAULT_MULTIPOLE_SLICES = 4
DEFAULT_SEXTUPOLE_SLICES = 4
DEFAULT_OCTUPOLE_SLICES = 4

# Extra hipRTC options appended to every kernel build.  The reference's
# OpenCL options (ufo/__init__.py:22) have no hipRTC meaning: the kernels
# keep IEEE semantics and write their fused multiply-adds explicitly.
DEFAULT_CL_OPTIONS = ""

__all__ = ["Lattice", "Beam", "Line", "Track"]


def __getattr__(name):
    # Drivers and the device layer load lazily: importing the package needs
    # neither torch nor a GPU (the reference opened an OpenCL context at
    # import, ufo/__init__.py:23).
    if name in ("Track",):
        from .track import Track
        return Track
    if name in ("StableAperture",):
        from .sa import StableAperture
        return StableAperture
    if name in ("ClosedOrbit", "TangentMap"):
        from . import co
        return getattr(co, name)
    if name in ("Optics",):
        from .optics import Optics
        return Optics
    if name in ("BatchedAperture",):
        from .batch import BatchedAperture
        return BatchedAperture
    if name in ("AdjointTracker",):
        from .adjoint import AdjointTracker
        return AdjointTracker
    if name in ("list_devices", "context", "DEFAULT_CONTEXT"):
        from . import runtime
        if name == "DEFAULT_CONTEXT":
            return runtime.default_context()
        return getattr(runtime, name)
    raise AttributeError(name)
