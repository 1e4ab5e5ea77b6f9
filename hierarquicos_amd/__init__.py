"""hierarquicos_amd -- MI355X (gfx950) hierarchical gravity force path.

GPU implementation of the Barnes-Hut / grouped / Dehnen octree force
evaluators and the direct sum of Potalej/hierarquicos behind the reference's
own API surface (the f2py ``api`` module and the OctreeType methods).
"""
from . import api, ics  # noqa: F401
from . import integrate  # noqa: F401
from ._lib import HQError, lib  # noqa: F401
from .device import DeviceArray, direct_forces, morton_keys  # noqa: F401
from .morton import MortonTreeType  # noqa: F401
from .octree import OctreeType, tune_table_clear, tune_table_load, tune_table_save  # noqa: F401

__all__ = ["api", "ics", "integrate", "OctreeType", "MortonTreeType", "DeviceArray", "direct_forces", "morton_keys", "lib",
           "HQError", "tune_table_save", "tune_table_load", "tune_table_clear"]
