// C-ABI plumbing: error strings, constants, fills, debug sync, state sizes.
#include <stdarg.h>
#include <stdio.h>

#include "svr_device.h"
#include "svr_host.h"

namespace svr {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ dst, uint32_t value, int64_t count)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) dst[i] = value;
}

// Clears device memory with a kernel instead of hipMemsetAsync so that a
// captured graph never records a memset node.
int fill_u32(void* dst, uint32_t value, int64_t count, hipStream_t stream)
{
    if (count <= 0) return SVR_OK;
    int64_t blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    fill_u32_kernel<<<(unsigned)blocks, 256, 0, stream>>>((uint32_t*)dst, value, count);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

}  // namespace svr

using namespace svr;

extern "C" {

const char* svr_last_error(void) { return g_err; }

int svr_get_constants(svr_constants* out)
{
    SVR_REQUIRE(out != nullptr, SVR_ERR_INVALID, "svr_get_constants: null output");
    out->max_num_levels = MAX_NUM_LEVELS;
    out->step_sz_scale = STEP_SZ_SCALE;
    out->max_n_samp = MAX_N_SAMP;
    out->max_sort_key_bits = MAX_SORT_KEY_BITS;
    out->max_packed_image_dim = MAX_PACKED_IMAGE_DIM;
    out->max_render_tiles = MAX_RENDER_TILES;
    out->tile_x = TILE_X;
    out->tile_y = TILE_Y;
    return SVR_OK;
}

size_t svr_geom_state_size(int64_t n)
{
    size_t bytes = 0;
    GeomState::carve(nullptr, n < 0 ? 0 : n, &bytes);
    return bytes;
}

size_t svr_order_table_size(int64_t n)
{
    size_t bytes = 0;
    OrderTable::carve(nullptr, n < 0 ? 0 : n, &bytes);
    return bytes;
}

size_t svr_binning_state_size(int64_t R)
{
    size_t bytes = 0;
    BinningState::carve(nullptr, R < 0 ? 0 : R, &bytes);
    return bytes;
}

size_t svr_image_state_size(int32_t W, int32_t H)
{
    size_t bytes = 0;
    const int64_t T = (int64_t)tiles_x(W) * tiles_y(H);
    ImageState::carve(nullptr, (int64_t)W * H, T, &bytes);
    return bytes;
}

int svr_debug_sync(svr_stream_t stream)
{
    SVR_CHECK_HIP(hipStreamSynchronize(stream));
    SVR_CHECK_HIP(hipGetLastError());
    return SVR_OK;
}

}  // extern "C"
