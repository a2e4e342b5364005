// C-ABI plumbing: error strings, constants, fills, debug sync, state sizes.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>

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

// 16-byte stores over the aligned body of the range; the elements before the
// first 16-byte boundary (head) and after the last one (tail) are stored
// singly by the first threads of the grid.
__global__ void fill_u32_kernel(uint32_t* __restrict__ dst, uint32_t value, int64_t count, int64_t head)
{
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t nvec = (count - head) >> 2;
    uint4* __restrict__ body = reinterpret_cast<uint4*>(dst + head);
    const uint4 v4 = make_uint4(value, value, value, value);
    for (int64_t i = t0; i < nvec; i += stride) body[i] = v4;
    if (t0 < head) dst[t0] = value;
    const int64_t tail0 = head + 4 * nvec;
    if (t0 < count - tail0) dst[tail0 + t0] = value;
}

// Two ranges in one launch: row blockIdx.y of the grid fills range y.
__global__ void fill_u32_pair_kernel(uint32_t* __restrict__ d0, int64_t n0, int64_t h0, uint32_t* __restrict__ d1,
                                     int64_t n1, int64_t h1, uint32_t value)
{
    uint32_t* dst = blockIdx.y ? d1 : d0;
    const int64_t count = blockIdx.y ? n1 : n0, head = blockIdx.y ? h1 : h0;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t nvec = (count - head) >> 2;
    uint4* __restrict__ body = reinterpret_cast<uint4*>(dst + head);
    const uint4 v4 = make_uint4(value, value, value, value);
    for (int64_t i = t0; i < nvec; i += stride) body[i] = v4;
    if (t0 < head) dst[t0] = value;
    const int64_t tail0 = head + 4 * nvec;
    if (t0 < count - tail0) dst[tail0 + t0] = value;
}

// Elements before the first 16-byte boundary of dst (at most count), and the
// workgroups a fill of the rest takes (at least one: head and tail need the
// first threads only; at most 4096, a grid-stride loop does the rest).
static int64_t fill_head(const void* dst, int64_t count)
{
    const int64_t head = (int64_t)(((16u - ((uintptr_t)dst & 15u)) & 15u) >> 2);
    return head > count ? count : head;
}
static int64_t fill_blocks(int64_t count, int64_t head)
{
    const int64_t b = ((count - head) / 4 + 255) / 256;
    return b < 1 ? 1 : (b > 4096 ? 4096 : b);
}

// Clears device memory with a kernel instead of hipMemsetAsync so that a
// captured graph never records a memset node.
int fill_u32(void* dst, uint32_t value, int64_t count, hipStream_t stream)
{
    if (count <= 0) return SVR_OK;
    SVR_REQUIRE(((uintptr_t)dst & 3u) == 0, SVR_ERR_INVALID, "fill: destination not 4-byte aligned");
    const int64_t head = fill_head(dst, count);
    fill_u32_kernel<<<(unsigned)fill_blocks(count, head), 256, 0, stream>>>((uint32_t*)dst, value, count, head);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

// fill_u32 of two ranges with one launch.
int fill_u32_pair(void* dst0, int64_t n0, void* dst1, int64_t n1, uint32_t value, hipStream_t stream)
{
    if (n0 <= 0) return fill_u32(dst1, value, n1, stream);
    if (n1 <= 0) return fill_u32(dst0, value, n0, stream);
    SVR_REQUIRE((((uintptr_t)dst0 | (uintptr_t)dst1) & 3u) == 0, SVR_ERR_INVALID,
                "fill: destination not 4-byte aligned");
    const int64_t h0 = fill_head(dst0, n0), h1 = fill_head(dst1, n1);
    const int64_t b0 = fill_blocks(n0, h0), b1 = fill_blocks(n1, h1);
    fill_u32_pair_kernel<<<dim3((unsigned)(b0 > b1 ? b0 : b1), 2), 256, 0, stream>>>(
        (uint32_t*)dst0, n0, h0, (uint32_t*)dst1, n1, h1, value);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

hipError_t thread_event(hipEvent_t* ev)
{
    constexpr int MAX_DEV = 64;
    thread_local hipEvent_t events[MAX_DEV] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEV) return hipErrorInvalidDevice;
    if (!events[dev])
    {
        e = hipEventCreateWithFlags(&events[dev], hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    *ev = events[dev];
    return hipSuccess;
}

hipError_t thread_pinned_words(uint32_t** words)
{
    thread_local uint32_t* buf = nullptr;
    if (!buf)
    {
        void* p = nullptr;
        const hipError_t e = hipHostMalloc(&p, 64 * sizeof(uint32_t), hipHostMallocDefault);
        if (e != hipSuccess) return e;
        buf = (uint32_t*)p;
    }
    *words = buf;
    return hipSuccess;
}

struct ProfPair
{
    hipEvent_t start, stop;
};

static bool g_prof_on = false;
static ProfPair* g_prof[PROF_NUM_STAGES] = {nullptr};
static int g_prof_cap[PROF_NUM_STAGES] = {0};
static int g_prof_n[PROF_NUM_STAGES] = {0};

void prof_begin(int stage, hipStream_t stream)
{
    if (!g_prof_on) return;
    if (g_prof_n[stage] == g_prof_cap[stage])
    {
        const int cap = g_prof_cap[stage] ? 2 * g_prof_cap[stage] : 64;
        ProfPair* np = (ProfPair*)realloc(g_prof[stage], sizeof(ProfPair) * cap);
        if (!np)
        {
            g_prof_on = false;
            return;
        }
        g_prof[stage] = np;
        for (int i = g_prof_cap[stage]; i < cap; ++i)
        {
            if (hipEventCreate(&np[i].start) != hipSuccess || hipEventCreate(&np[i].stop) != hipSuccess)
            {
                g_prof_on = false;
                return;
            }
            g_prof_cap[stage] = i + 1;
        }
    }
    (void)hipEventRecord(g_prof[stage][g_prof_n[stage]].start, stream);
}

void prof_end(int stage, hipStream_t stream)
{
    if (!g_prof_on || g_prof_n[stage] >= g_prof_cap[stage]) return;
    (void)hipEventRecord(g_prof[stage][g_prof_n[stage]].stop, stream);
    ++g_prof_n[stage];
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

size_t svr_image_state_size(int32_t W, int32_t H, int64_t R)
{
    size_t bytes = 0;
    const int64_t T = (int64_t)tiles_x(W) * tiles_y(H);
    ImageState::carve(nullptr, (int64_t)W * H, T, &bytes, R > 0 ? R : 0);
    return bytes;
}

int svr_profile_enable(int32_t on)
{
    g_prof_on = on != 0;
    for (int i = 0; i < PROF_NUM_STAGES; ++i) g_prof_n[i] = 0;
    return SVR_OK;
}

int svr_profile_collect(int32_t stage, double* total_ms_host, int64_t* count_host)
{
    SVR_REQUIRE(stage >= 0 && stage < PROF_NUM_STAGES && total_ms_host && count_host, SVR_ERR_INVALID,
                "svr_profile_collect: bad stage or null output");
    double total = 0.0;
    for (int i = 0; i < g_prof_n[stage]; ++i)
    {
        SVR_CHECK_HIP(hipEventSynchronize(g_prof[stage][i].stop));
        float ms = 0.f;
        SVR_CHECK_HIP(hipEventElapsedTime(&ms, g_prof[stage][i].start, g_prof[stage][i].stop));
        total += ms;
    }
    *total_ms_host = total;
    *count_host = g_prof_n[stage];
    g_prof_n[stage] = 0;
    return SVR_OK;
}

int svr_debug_sync(svr_stream_t stream)
{
    SVR_CHECK_HIP(hipStreamSynchronize(stream));
    SVR_CHECK_HIP(hipGetLastError());
    return SVR_OK;
}

}  // extern "C"
