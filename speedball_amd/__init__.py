"""speedball_amd — MI355X-native DDGI probe update (HIP for gfx950) behind the
reference's probe-field API. See include/sbgi.h for the C ABI."""
from . import _native
from ._native import lib, SbgiError
from .field import ProbeField
from .scene import Scene, SceneBuilder, uber_material, point_light, directional_light

__all__ = ["lib", "SbgiError", "ProbeField", "Scene", "SceneBuilder", "uber_material", "point_light",
           "directional_light", "_native"]
