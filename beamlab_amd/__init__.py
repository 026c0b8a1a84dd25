"""MI355X (gfx950) soft-body step of BeamLab's node/beam physics behind a C ABI (include/beamlab_gpu.h).

The compute path is hand-written HIP in beamlab_amd/csrc, built into beamlab_amd/libbeamlab_gpu.so; this package
is its host-side mirror of the reference's World / SoftBody / StaticWorld interface (src/phys).
"""
from . import abi
from .world import SoftBody, StaticWorld, World, make_box

__all__ = ["abi", "SoftBody", "StaticWorld", "World", "make_box"]
