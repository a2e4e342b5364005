// Voxel preprocess (screen bbox, camera quadrants, entry counts) and the
// prefix scan of the entry counts.
//
// Reference: backends/new_cuda/src/preprocess.cu:24-199 (kernel + host),
// forward.cu:649-663 (scan + reading the number of rendered entries).
#include "voxel_ops.h"
#include "svr_host.h"

namespace svr {

// One thread per voxel.
__global__ __launch_bounds__(256) void preprocess_kernel(
    int64_t N, int W, int H, int gx, int gy, float focal_x, float focal_y, float cx, float cy,
    const float* __restrict__ w2c, const float* __restrict__ c2w, float near_plane,
    const float* __restrict__ centers, const float* __restrict__ lengths,
    int32_t* __restrict__ out_n_dup, uint32_t* __restrict__ n_dup, uint2* __restrict__ bboxes,
    uint32_t* __restrict__ qbits_out, OctBox* __restrict__ octs)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    float m[12], cw[12];
#pragma unroll
    for (int i = 0; i < 12; ++i)
    {
        m[i] = w2c[i];
        cw[i] = c2w[i];
    }
    uint32_t nd, bits;
    uint2 bb;
    OctBox oc;
    preprocess_voxel(centers[3 * v + 0], centers[3 * v + 1], centers[3 * v + 2], 0.5f * lengths[v], W, H, gx, gy,
                     focal_x, focal_y, cx, cy, m, cw, near_plane, nd, bb, bits, oc);
    octs[v] = oc;
    out_n_dup[v] = (int32_t)nd;
    n_dup[v] = nd;
    bboxes[v] = bb;
    qbits_out[v] = bits;
}

// ---- reduce-then-scan over n_dup -------------------------------------

__device__ inline uint32_t wave_incl_scan(uint32_t x, int lane)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

// Inclusive block scan for SCAN_BLOCK threads; returns the block total in *total.
__device__ inline uint32_t block_incl_scan(uint32_t x, uint32_t* lds, uint32_t* total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    constexpr int NW = SCAN_BLOCK / 64;
    x = wave_incl_scan(x, lane);
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    if (wid == 0)
    {
        uint32_t s = lane < NW ? lds[lane] : 0u;
        s = wave_incl_scan(s, lane);
        if (lane < NW) lds[lane] = s;
    }
    __syncthreads();
    const uint32_t prefix = wid > 0 ? lds[wid - 1] : 0u;
    *total = lds[NW - 1];
    __syncthreads();
    return x + prefix;
}

__global__ __launch_bounds__(SCAN_BLOCK) void scan_reduce_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                                 uint32_t* __restrict__ block_sums)
{
    __shared__ uint32_t lds[SCAN_BLOCK / 64];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
        if (base + i < n) s += in[base + i];
    uint32_t total;
    block_incl_scan(s, lds, &total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Exclusive scan of the block sums in place (one workgroup); total -> counters[0..1] (u64).
__device__ inline void scan_sums_block(uint32_t* __restrict__ block_sums, int64_t nb, uint32_t* __restrict__ counters)
{
    __shared__ uint32_t lds[SCAN_BLOCK / 64];
    __shared__ uint64_t carry_s;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (int64_t off = 0; off < nb; off += SCAN_BLOCK)
    {
        const int64_t i = off + threadIdx.x;
        const uint32_t x = i < nb ? block_sums[i] : 0u;
        uint32_t total;
        const uint32_t incl = block_incl_scan(x, lds, &total);
        const uint64_t carry = carry_s;
        if (i < nb) block_sums[i] = (uint32_t)(carry + incl - x);
        __syncthreads();
        if (threadIdx.x == 0) carry_s = carry + total;
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        const uint64_t R = carry_s;
        counters[0] = (uint32_t)R;
        counters[1] = (uint32_t)(R >> 32);
    }
}

__global__ __launch_bounds__(SCAN_BLOCK) void scan_sums_kernel(uint32_t* __restrict__ block_sums, int64_t nb,
                                                               uint32_t* __restrict__ counters)
{
    scan_sums_block(block_sums, nb, counters);
}

// Workgroup 0 scans the entry-count block sums (total -> counters[0..1]),
// workgroup 1 the in-frustum block counts (total -> counters[6..7]).  With
// host_words (pinned host memory, device-accessible), both totals also go
// there at the same offsets, so the host reads them without a copy.
__global__ __launch_bounds__(SCAN_BLOCK) void scan_sums2_kernel(uint32_t* __restrict__ block_sums,
                                                                uint32_t* __restrict__ vis_sums, int64_t nb,
                                                                uint32_t* __restrict__ counters,
                                                                uint32_t* __restrict__ host_words)
{
    const int o = blockIdx.x == 0 ? 0 : 6;
    scan_sums_block(blockIdx.x == 0 ? block_sums : vis_sums, nb, counters + o);
    if (host_words && threadIdx.x == 0)
    {
        host_words[o] = counters[o];
        host_words[o + 1] = counters[o + 1];
    }
}

// Per scan block: the entry total (block_sums) and the number of in-frustum
// voxels, n_dup > 0 (vis_sums).
__global__ __launch_bounds__(SCAN_BLOCK) void count_reduce_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                                  uint32_t* __restrict__ block_sums,
                                                                  uint32_t* __restrict__ vis_sums)
{
    __shared__ uint32_t lds[SCAN_BLOCK / 64];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0, c = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
    {
        const uint32_t x = (base + i < n) ? in[base + i] : 0u;
        s += x;
        c += (x != 0u);
    }
    uint32_t total;
    block_incl_scan(s, lds, &total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
    block_incl_scan(c, lds, &total);
    if (threadIdx.x == 0) vis_sums[blockIdx.x] = total;
}

// Ascending ids of the voxels with n_dup > 0; scan block b writes from
// vis_sums[b] (exclusive-scanned) on (renderer.py:81, torch.where(n_dup > 0)).
__global__ __launch_bounds__(SCAN_BLOCK) void visible_write_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                                   const uint32_t* __restrict__ vis_sums,
                                                                   int64_t* __restrict__ idx)
{
    __shared__ uint32_t lds[SCAN_BLOCK / 64];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    bool f[SCAN_ITEMS];
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
    {
        f[i] = (base + i < n) && in[base + i] != 0u;
        c += f[i];
    }
    uint32_t total;
    const uint32_t incl = block_incl_scan(c, lds, &total);
    int64_t o = (int64_t)vis_sums[blockIdx.x] + (int64_t)(incl - c);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
        if (f[i]) idx[o++] = base + i;
}

// ---- exclusive scan of the rank-ordered pair counts, in place ---------

// PSCAN_ITEMS consecutive elements of thread t (16-byte aligned runs).
__device__ inline void pscan_load(const uint32_t* __restrict__ c, int64_t n, int64_t base, uint32_t (&v)[PSCAN_ITEMS])
{
    if (base + PSCAN_ITEMS <= n)
    {
#pragma unroll
        for (int i = 0; i < PSCAN_ITEMS / 4; ++i)
        {
            const uint4 q = *reinterpret_cast<const uint4*>(c + base + 4 * i);
            v[4 * i + 0] = q.x; v[4 * i + 1] = q.y; v[4 * i + 2] = q.z; v[4 * i + 3] = q.w;
        }
    }
    else
    {
#pragma unroll
        for (int i = 0; i < PSCAN_ITEMS; ++i) v[i] = (base + i < n) ? c[base + i] : 0u;
    }
}

__global__ __launch_bounds__(SCAN_BLOCK) void pscan_reduce_kernel(const uint32_t* __restrict__ c, int64_t n,
                                                                  uint32_t* __restrict__ sums)
{
    __shared__ uint32_t lds[SCAN_BLOCK / 64];
    uint32_t v[PSCAN_ITEMS];
    pscan_load(c, n, (int64_t)blockIdx.x * PSCAN_TILE + (int64_t)threadIdx.x * PSCAN_ITEMS, v);
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < PSCAN_ITEMS; ++i) s += v[i];
    uint32_t total;
    block_incl_scan(s, lds, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Every element becomes its exclusive prefix (only those of nonzero counts
// are read later; writing whole 16-byte runs keeps the stores dense).
__global__ __launch_bounds__(SCAN_BLOCK) void pscan_apply_kernel(uint32_t* __restrict__ c, int64_t n,
                                                                 const uint32_t* __restrict__ sums)
{
    __shared__ uint32_t lds[SCAN_BLOCK / 64];
    const int64_t base = (int64_t)blockIdx.x * PSCAN_TILE + (int64_t)threadIdx.x * PSCAN_ITEMS;
    uint32_t v[PSCAN_ITEMS];
    pscan_load(c, n, base, v);
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < PSCAN_ITEMS; ++i) s += v[i];
    uint32_t total;
    uint32_t run = block_incl_scan(s, lds, &total) - s + sums[blockIdx.x];
#pragma unroll
    for (int i = 0; i < PSCAN_ITEMS; ++i)
    {
        const uint32_t x = v[i];
        v[i] = run;
        run += x;
    }
    if (base + PSCAN_ITEMS <= n)
    {
#pragma unroll
        for (int i = 0; i < PSCAN_ITEMS / 4; ++i)
            *reinterpret_cast<uint4*>(c + base + 4 * i) = make_uint4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
    }
    else
    {
#pragma unroll
        for (int i = 0; i < PSCAN_ITEMS; ++i)
            if (base + i < n) c[base + i] = v[i];
    }
}

int pair_scan_launch(uint32_t* c, int64_t n, uint32_t* sums, hipStream_t stream)
{
    if (n <= 0) return SVR_OK;
    const int64_t nb = (n + PSCAN_TILE - 1) / PSCAN_TILE;
    SVR_REQUIRE(nb < (1ll << 31), SVR_ERR_INVALID, "pair scan of %lld elements too large", (long long)n);
    pscan_reduce_kernel<<<(unsigned)nb, SCAN_BLOCK, 0, stream>>>(c, n, sums);
    SVR_CHECK_LAUNCH();
    scan_sums_kernel<<<1, SCAN_BLOCK, 0, stream>>>(sums, nb, sums + nb);
    SVR_CHECK_LAUNCH();
    pscan_apply_kernel<<<(unsigned)nb, SCAN_BLOCK, 0, stream>>>(c, n, sums);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

}  // namespace svr

using namespace svr;

extern "C" {

int svr_rasterize_preprocess(int32_t W, int32_t H, float tanfovx, float tanfovy, float cx, float cy,
                             const float* w2c, const float* c2w, float near_plane, int64_t N,
                             const float* centers, const float* lengths, int32_t* out_n_dup, void* geom_state,
                             svr_stream_t stream)
{
    SVR_REQUIRE(W > 0 && H > 0 && W <= MAX_PACKED_IMAGE_DIM && H <= MAX_PACKED_IMAGE_DIM, SVR_ERR_INVALID,
                "image size %dx%d outside [1, %d]", W, H, MAX_PACKED_IMAGE_DIM);
    SVR_REQUIRE(N >= 0 && N <= (int64_t)VOX_ID_MASK + 1, SVR_ERR_INVALID, "voxel count %lld out of range",
                (long long)N);
    if (N == 0) return SVR_OK;
    SVR_REQUIRE(geom_state && centers && lengths && out_n_dup && w2c && c2w, SVR_ERR_INVALID,
                "svr_rasterize_preprocess: null pointer");
    GeomState g = GeomState::carve(geom_state, N, nullptr);
    const float focal_x = 0.5f * (float)W / tanfovx;
    const float focal_y = 0.5f * (float)H / tanfovy;
    const int64_t blocks = (N + 255) / 256;
    prof_begin(PROF_PREPROCESS, stream);
    preprocess_kernel<<<(unsigned)blocks, 256, 0, stream>>>(N, W, H, tiles_x(W), tiles_y(H), focal_x, focal_y, cx,
                                                            cy, w2c, c2w, near_plane, centers, lengths, out_n_dup,
                                                            g.n_dup, g.bbox, g.qbits, g.oct);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_PREPROCESS, stream);
    return SVR_OK;
}

int svr_scan_entries(int64_t N, void* geom_state, svr_stream_t stream)
{
    SVR_REQUIRE(N >= 0 && N < (1ll << 32), SVR_ERR_INVALID, "entry-count array length out of range");
    if (N == 0) return SVR_OK;
    SVR_REQUIRE(geom_state != nullptr, SVR_ERR_INVALID, "svr_scan_entries: null geometry state");
    GeomState g = GeomState::carve(geom_state, N, nullptr);
    const int64_t nb = (N + SCAN_TILE - 1) / SCAN_TILE;
    prof_begin(PROF_SCAN, stream);
    scan_reduce_kernel<<<(unsigned)nb, SCAN_BLOCK, 0, stream>>>(g.n_dup, N, g.block_sums);
    SVR_CHECK_LAUNCH();
    scan_sums_kernel<<<1, SCAN_BLOCK, 0, stream>>>(g.block_sums, nb, g.counters);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_SCAN, stream);
    return SVR_OK;
}

int svr_num_rendered(int64_t N, void* geom_state, svr_stream_t stream, int64_t* R_host)
{
    SVR_REQUIRE(R_host != nullptr, SVR_ERR_INVALID, "svr_num_rendered: null output");
    *R_host = 0;
    if (N == 0) return SVR_OK;
    const int rc = svr_scan_entries(N, geom_state, stream);
    if (rc != SVR_OK) return rc;
    GeomState g = GeomState::carve(geom_state, N, nullptr);
    uint32_t cnt[2] = {0, 0};
    SVR_CHECK_HIP(hipMemcpyAsync(cnt, g.counters, sizeof(cnt), hipMemcpyDeviceToHost, stream));
    SVR_CHECK_HIP(hipStreamSynchronize(stream));
    const uint64_t R = ((uint64_t)cnt[1] << 32) | cnt[0];
    SVR_REQUIRE(R < (1ull << 32), SVR_ERR_CAPACITY, "number of rendered entries %llu exceeds 2^32 - 1",
                (unsigned long long)R);
    *R_host = (int64_t)R;
    return SVR_OK;
}

int svr_visible_voxels(int64_t N, void* geom_state, const void* order_table, int64_t* idx_out, svr_stream_t stream,
                       int64_t* R_host, int64_t* n_visible_host)
{
    SVR_REQUIRE(R_host != nullptr && n_visible_host != nullptr, SVR_ERR_INVALID, "svr_visible_voxels: null output");
    *R_host = 0;
    *n_visible_host = 0;
    SVR_REQUIRE(N >= 0 && N < (1ll << 32), SVR_ERR_INVALID, "voxel count out of range");
    if (N == 0) return SVR_OK;
    SVR_REQUIRE(geom_state != nullptr && idx_out != nullptr, SVR_ERR_INVALID, "svr_visible_voxels: null pointer");
    GeomState g = GeomState::carve(geom_state, N, nullptr);
    const int64_t nb = (N + SCAN_TILE - 1) / SCAN_TILE;
    prof_begin(PROF_SCAN, stream);
    count_reduce_kernel<<<(unsigned)nb, SCAN_BLOCK, 0, stream>>>(g.n_dup, N, g.block_sums, g.vis_sums);
    SVR_CHECK_LAUNCH();
    // the two totals go straight to pinned host memory (no copy); the host
    // waits for them only, while the list, the pair offsets and whatever the
    // caller queues next keep the device busy
    uint32_t* cnt = nullptr;
    SVR_CHECK_HIP(thread_pinned_words(&cnt));
    uint32_t* cnt_dev = nullptr;
    SVR_CHECK_HIP(hipHostGetDevicePointer((void**)&cnt_dev, cnt, 0));
    scan_sums2_kernel<<<2, SCAN_BLOCK, 0, stream>>>(g.block_sums, g.vis_sums, nb, g.counters, cnt_dev);
    SVR_CHECK_LAUNCH();
    hipEvent_t ev = nullptr;
    SVR_CHECK_HIP(thread_event(&ev));
    SVR_CHECK_HIP(hipEventRecord(ev, stream));
    visible_write_kernel<<<(unsigned)nb, SCAN_BLOCK, 0, stream>>>(g.n_dup, N, g.vis_sums, idx_out);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_SCAN, stream);
    if (order_table)
    {
        OrderTable ot = OrderTable::carve((void*)order_table, N, nullptr);
        const int rc = pair_prepare(N, g, ot, stream);
        if (rc != SVR_OK) return rc;
    }
    SVR_CHECK_HIP(hipEventSynchronize(ev));
    const uint64_t R = ((uint64_t)cnt[1] << 32) | cnt[0];
    SVR_REQUIRE(R < (1ull << 32), SVR_ERR_CAPACITY, "number of rendered entries %llu exceeds 2^32 - 1",
                (unsigned long long)R);
    *R_host = (int64_t)R;
    *n_visible_host = (int64_t)(((uint64_t)cnt[7] << 32) | cnt[6]);
    return SVR_OK;
}

}  // extern "C"
