"""MI355X-native sparse-voxel SDF rasterizer with the `svraster_cuda` Python surface.

Submodules mirror the reference package (cuda/svraster_cuda/__init__.py:1-6): meta, utils,
renderer, sparse_adam, marching_cubes, grid_loss_bw.  The compute lives in
libsvraster_hip.so (C-ABI in include/svraster_hip.h), bound by `_C`.  `batched` adds the
multi-view persistent path and `synthetic` the seeded scene/camera generator.
"""
from . import meta
from . import utils
from . import renderer
from . import sparse_adam
from . import marching_cubes
from . import grid_loss_bw

__all__ = ["meta", "utils", "renderer", "sparse_adam", "marching_cubes", "grid_loss_bw"]
