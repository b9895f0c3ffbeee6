"""``from spheral_amd.Spheral3d import *``: the reference's 3-D Python names for the SPH +
TreeGravity path, bound to the GPU engine (see spheral_amd.mirror)."""
from .mirror import namespace as _namespace

globals().update(_namespace(3))
__all__ = sorted(k for k in globals() if not k.startswith("_"))
