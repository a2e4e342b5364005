"""MI355X-native sparse-voxel rasterizer with the module layout of the
reference backend package `new_svraster_cuda`
(backends/new_cuda/new_svraster_cuda/__init__.py:11-15): submodules
renderer, sparse_adam, utils, meta and grid_loss_bw over the HIP library
libsvraster_hip.so, plus `batch` (multi-view rendering, no reference
counterpart).

Importing the package loads the library and fails if it has not been built
(`python sv_raster_amd/build.py`); there is no CPU fallback.
"""
from . import _C  # noqa: F401  (loads libsvraster_hip.so)
from . import meta  # noqa: F401
from . import utils  # noqa: F401
from . import renderer  # noqa: F401
from . import sparse_adam  # noqa: F401
from . import grid_loss_bw  # noqa: F401
from . import batch  # noqa: F401
from . import checkpoint  # noqa: F401
