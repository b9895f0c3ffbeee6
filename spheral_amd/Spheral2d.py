"""``from spheral_amd.Spheral2d import *``: the reference's 2-D Python names for the SPH +
TreeGravity path, bound to the GPU engine (see spheral_amd.mirror)."""
from .mirror import namespace as _namespace

globals().update(_namespace(2))
__all__ = sorted(k for k in globals() if not k.startswith("_"))
