"""osinco3d time-step path on MI355X (gfx950).

libo3d.so (HIP) holds the kernels and the C ABI (include/o3d.h); the Fortran
drop-in modules in fortran/ and the Python modules here are its host side.
"""
from .config import Config, tgv_config, mixing_layer_config  # noqa: F401

__version__ = "0.1.0"
