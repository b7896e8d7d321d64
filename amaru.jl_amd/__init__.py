"""amaru.jl_amd — MI355X (gfx950) implementation of Amaru.jl's FE assembly
and solve path.

The compute path is ``libamaru_hip.so`` (hand-written HIP kernels behind the
C ABI of ``include/amaru_hip.h``); this package is the host side that mirrors
the reference's user API:

    Block → Mesh → FEModel(mesh, [filter => MechSolid => Material(...)])
    → MechAnalysis(model) → addstage(ana, bcs; nincs, nouts) → solve(ana; ...)
"""
from ._lib import AmaruError, device_count, lib
from .analysis import MechAnalysis, addstage, solve
from .mesh import HEX8, HEX20, TET10, Block, Mesh
from .model import (OUTPUT_FIELDS, BodyC, DeviceModel, DruckerPrager, EdgeBC, FaceBC, FEModel, LinearElastic,
                    MechSolid, NodeBC, SurfaceBC, VonMises, output_fields)

__all__ = ["AmaruError", "device_count", "lib", "MechAnalysis", "addstage", "solve", "HEX8", "HEX20", "TET10",
           "Block", "Mesh", "BodyC", "DeviceModel", "DruckerPrager", "EdgeBC", "FaceBC", "FEModel", "LinearElastic",
           "MechSolid", "NodeBC", "SurfaceBC", "VonMises", "OUTPUT_FIELDS", "output_fields"]
