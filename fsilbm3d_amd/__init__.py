"""MI355X-native FSILBM3D fluid/coupling hot path.

Device code: fsilbm3d_amd/csrc (HIP, gfx950) behind the C-ABI of
include/fsilbm3d_amd.h, built into fsilbm3d_amd/libfsil.so.
Host mirror of the reference's block interface: LBMBlock (block.py), rigid
bodies (solid.py) and the coupled step (fsi.py).
"""
__all__ = ["LBMBlock", "FSIBlock", "RigidPlate"]


def __getattr__(name):
    if name == "LBMBlock":
        from .block import LBMBlock
        return LBMBlock
    if name == "FSIBlock":
        from .fsi import FSIBlock
        return FSIBlock
    if name == "RigidPlate":
        from .solid import RigidPlate
        return RigidPlate
    raise AttributeError(name)
