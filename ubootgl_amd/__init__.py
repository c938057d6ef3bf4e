"""ubootgl_amd: MI355X-native (gfx950 HIP) fluid step of te42kyfo/ubootgl.

The compute path is libubg.so (ubootgl_amd/csrc, C ABI in include/ubg_c.h);
this package is its Python mirror of the reference's Simulation / MG
interface plus the seeded workload generators.
"""
from ._lib import UbgError, device_count, lib  # noqa: F401
from .simulation import (MG, Params, Simulation, flag_components, level_sizes,  # noqa: F401
                         remove_component_means, terrain_flags)

__all__ = ["MG", "Params", "Simulation", "UbgError", "device_count", "flag_components",
           "level_sizes", "lib", "remove_component_means"]
