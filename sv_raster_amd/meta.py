"""Read-only constants of the rasterizer (backends/new_cuda/new_svraster_cuda/meta.py:13-17)."""
from . import _C

MAX_NUM_LEVELS = _C.MAX_NUM_LEVELS
STEP_SZ_SCALE = _C.STEP_SZ_SCALE
MAX_N_SAMP = _C.MAX_N_SAMP
MAX_SORT_KEY_BITS = _C.MAX_SORT_KEY_BITS
MAX_PACKED_IMAGE_DIM = _C.MAX_PACKED_IMAGE_DIM
MAX_RENDER_TILES = _C.MAX_RENDER_TILES
