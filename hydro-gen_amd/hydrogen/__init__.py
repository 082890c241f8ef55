"""hydrogen — host-side Python binding of the MI355X erosion library.

The compute path is libhydrogen.so (HIP kernels for gfx950 behind the C ABI
of include/hg.h).  This package only binds it (ctypes) and adds the PyTorch
autograd harness; nothing here computes erosion on the CPU.
"""
from ._abi import *  # noqa: F401,F403
from ._abi import load, Terrain, Batch, HGError, erosion_settings, rain_settings, fbm_params, tile_variants, droplet_settings  # noqa: F401
