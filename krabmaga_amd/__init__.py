"""krabmaga_amd -- MI355X (gfx950) agent-step engine for krABMaga's Field2D
and flockers model.

The compute path is libkrabgpu.so (HIP kernels behind the C ABI declared in
include/krabgpu.h); this package is its Python front end.
"""
from ._abi import KgError, lib  # noqa: F401
from .ensemble import Ensemble, RunSpec, Tuning, footprint  # noqa: F401
from .field import Field2D, FlockParams  # noqa: F401

__all__ = ["Ensemble", "Field2D", "FlockParams", "KgError", "RunSpec", "Tuning", "footprint",
           "lib"]
