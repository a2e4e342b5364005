// Per-tile front-to-back compositor and its reverse-order adjoint.
//
// Reference: backends/new_cuda/src/forward.cu:43-394 (renderCUDA),
// backward.cu:37-530 (renderCUDA backward), host sides forward.cu:607-879 and
// backward.cu:654-772.
//
// Mapping: one 256-thread workgroup per 16x16 tile; each of its four waves owns
// one 8x8 sub-tile (lane = 8 * row + column) and walks the tile's entry list on
// its own, 64 entries per step, with no workgroup barrier inside the loop:
//   1. lane l loads entry base+l, tests its packed bbox against the sub-tile
//      and its quadrant against the quadrants of the still-active pixels,
//      producing a 64-bit pixel mask (row of a 64x64 bit matrix), and stages
//      the entry's payload in LDS at slot l;
//   2. the bit matrix is transposed across the wave (6 shuffle steps), giving
//      each pixel lane the mask of entries whose bbox covers it;
//   3. every lane walks its own mask in list order (lock-step over hits, not
//      over entries), doing the ray-box test and the compositing.
// Per-pixel results do not depend on this mapping: each pixel visits exactly
// the entries the reference visits, in the same order, with the same
// arithmetic.  The backward accumulates the 12 per-entry gradients of a chunk
// in LDS and flushes them into one partial row per list entry; a per-voxel
// segmented sum over those rows gives the voxel gradients.
#include <mutex>

#include "compositor.h"
#include "svr_host.h"

namespace svr {

// One launch per direction and frame.  render_fwd_kernel takes the tiles in
// queue order (workgroup b: tile order[b]) and composites with the hardware
// min/max in the ray-box test when all 256 rays of the tile have finite 1/rd;
// a tile with another ray takes the reference's ternary min/max (whose NaN
// handling differs) through a workgroup-uniform branch, at no register cost,
// and is also appended to exact_list for the extended backward.
// render_bwd_kernel makes the same choice per tile.

template <int NS, bool EXACT>
__device__ inline void fwd_tile(Stage& st, uint32_t (&s_last)[NWAVE], int tile, const PixRay& ray, int W, int H,
                                int gx, float bg, const uint2* __restrict__ ranges,
                                const uint32_t* __restrict__ vox_list, const uint2* __restrict__ bboxes, const OctBox* __restrict__ octs,
                                const float* __restrict__ centers, const float* __restrict__ lengths,
                                const float* __restrict__ geos, const float* __restrict__ rgbs,
                                uint32_t* __restrict__ tile_last, uint32_t* __restrict__ n_contrib,
                                float* __restrict__ out_color, float* __restrict__ out_T,
                                uint64_t* __restrict__ hit_masks)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tx = tile % gx, ty = tile / gx;
    const int sx0 = tx * TILE_X + (w & 1) * 8, sy0 = ty * TILE_Y + (w >> 1) * 8;
    const int sx1 = min(sx0 + 7, W - 1), sy1 = min(sy0 + 7, H - 1);
    const int px = sx0 + (lane & 7), py = sy0 + (lane >> 3);
    const bool inside = px < W && py < H;
    const uint2 range = ranges[tile];

    float T = 1.f, C0 = 0.f, C1 = 0.f, C2 = 0.f;
    uint64_t* mask_out = hit_masks + (((size_t)(range.x >> 6) + tile) * NWAVE + w) * 64 + lane;
    const uint32_t last = composite_subtile<NS, EXACT, NWAVE, false, true>(
        st, w, lane, ray, inside, range, sx0, sy0, sx1, sy1, vox_list, bboxes, octs, centers, lengths, geos, rgbs, T, C0, C1,
        C2, nullptr, nullptr, mask_out);

    if (inside)
    {
        const size_t pid = (size_t)py * W + px;
        const size_t HW = (size_t)H * W;
        n_contrib[pid] = last;
        out_color[pid] = C0 + T * bg;
        out_color[HW + pid] = C1 + T * bg;
        out_color[2 * HW + pid] = C2 + T * bg;
        out_T[pid] = T;
    }
    const uint32_t wm = wave_max_u32(inside ? last : 0u);
    if (lane == 0) s_last[w] = wm;
    __syncthreads();
    if (threadIdx.x == 0)
        tile_last[tile] = range.x + max(max(s_last[0], s_last[1]), max(s_last[2], s_last[3]));
}

__device__ inline PixRay tile_ray(int tile, int W, int H, int gx, float tanfovx, float tanfovy, float cx, float cy,
                                  const float* __restrict__ c2w)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tx = tile % gx, ty = tile / gx;
    const int px = tx * TILE_X + (w & 1) * 8 + (lane & 7), py = ty * TILE_Y + (w >> 1) * 8 + (lane >> 3);
    return make_ray(px, py, W, H, tanfovx, tanfovy, cx, cy, c2w);
}

template <int NS>
__global__ __launch_bounds__(256) void render_fwd_kernel(
    int W, int H, int gx, const uint32_t* __restrict__ tile_order, float tanfovx, float tanfovy, float cx, float cy,
    const float* __restrict__ c2w, float bg, const uint2* __restrict__ ranges, const uint32_t* __restrict__ vox_list,
    const uint2* __restrict__ bboxes, const OctBox* __restrict__ octs, const float* __restrict__ centers, const float* __restrict__ lengths,
    const float* __restrict__ geos, const float* __restrict__ rgbs, uint32_t* __restrict__ tile_last,
    uint32_t* __restrict__ n_contrib, float* __restrict__ out_color, float* __restrict__ out_T,
    uint64_t* __restrict__ hit_masks, uint32_t* __restrict__ exact_ctr, uint32_t* __restrict__ exact_list)
{
    __shared__ Stage st;
    __shared__ uint32_t s_last[NWAVE];
    const int tile = (int)tile_order[blockIdx.x];
    const PixRay ray = tile_ray(tile, W, H, gx, tanfovx, tanfovy, cx, cy, c2w);
    if (__syncthreads_or(!ray_inv_finite(ray)) != 0)
    {
        // still listed: the extended backward (after a plain forward when only
        // its regularizers are on) walks the list in its exact launch
        if (threadIdx.x == 0) exact_list[atomicAdd(exact_ctr, 1u)] = (uint32_t)tile;
        fwd_tile<NS, true>(st, s_last, tile, ray, W, H, gx, bg, ranges, vox_list, bboxes, octs, centers, lengths, geos,
                           rgbs, tile_last, n_contrib, out_color, out_T, hit_masks);
    }
    else
        fwd_tile<NS, false>(st, s_last, tile, ray, W, H, gx, bg, ranges, vox_list, bboxes, octs, centers, lengths, geos,
                            rgbs, tile_last, n_contrib, out_color, out_T, hit_masks);
}

template <int NS, int SYNC_MIN, bool EXACT>
__device__ inline void bwd_tile(WindowLDS<BWD_WINDOW>& sh, int tile, const PixRay& ray, int W, int H, int gx,
                                float bg, const uint2* __restrict__ ranges, const uint32_t* __restrict__ vox_list,
                                const float* __restrict__ centers, const float* __restrict__ lengths,
                                const float* __restrict__ geos, const float* __restrict__ rgbs,
                                const float* __restrict__ out_T, const uint32_t* __restrict__ n_contrib,
                                const float* __restrict__ dL_dcolor, const float* __restrict__ dL_dT,
                                const uint32_t* __restrict__ orig, const uint64_t* __restrict__ hit_masks,
                                const uint32_t* __restrict__ tile_last, float* __restrict__ partial)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tx = tile % gx, ty = tile / gx;
    const int px = tx * TILE_X + (w & 1) * 8 + (lane & 7), py = ty * TILE_Y + (w >> 1) * 8 + (lane >> 3);
    const bool inside = px < W && py < H;
    PlainHit<NS, EXACT> hit;
    hit.ray = ray;
    const uint2 range = ranges[tile];
    const uint32_t C = tile_chunks(range, tile_last[tile]);
    const size_t HW = (size_t)H * W;
    const size_t pid = inside ? (size_t)py * W + px : 0;
    const uint64_t* masks = hit_masks + (((size_t)(range.x >> 6) + tile) * NWAVE + w) * 64 + lane;
    adjoint_window<SYNC_MIN, Tile16, BWD_WINDOW>(
        sh, w, lane, C, [&]() { return hit.init(inside, pid, HW, bg, out_T, n_contrib, dL_dcolor, dL_dT); }, hit, range,
        masks, vox_list, centers, lengths, geos, rgbs, orig, partial);
}

template <int NS, int SYNC_MIN>
__global__ __launch_bounds__(256) void render_bwd_kernel(
    int W, int H, int gx, const uint32_t* __restrict__ tile_order, float tanfovx, float tanfovy, float cx, float cy,
    const float* __restrict__ c2w, float bg, const uint2* __restrict__ ranges, const uint32_t* __restrict__ vox_list,
    const float* __restrict__ centers, const float* __restrict__ lengths, const float* __restrict__ geos,
    const float* __restrict__ rgbs, const float* __restrict__ out_T, const uint32_t* __restrict__ n_contrib,
    const float* __restrict__ dL_dcolor, const float* __restrict__ dL_dT, const uint32_t* __restrict__ orig,
    const uint64_t* __restrict__ hit_masks, const uint32_t* __restrict__ tile_last, float* __restrict__ partial)
{
    __shared__ WindowLDS<BWD_WINDOW> sh;
    const int tile = (int)tile_order[blockIdx.x];
    const PixRay ray = tile_ray(tile, W, H, gx, tanfovx, tanfovy, cx, cy, c2w);
    // a tile with a non-finite 1/rd takes the reference's min/max here (the
    // same workgroup-uniform choice as the forward's exact launch); its larger
    // register count costs no occupancy, which the LDS window bounds
    if (__syncthreads_or(!ray_inv_finite(ray)) != 0)
        bwd_tile<NS, SYNC_MIN, true>(sh, tile, ray, W, H, gx, bg, ranges, vox_list, centers, lengths, geos, rgbs,
                                     out_T, n_contrib, dL_dcolor, dL_dT, orig, hit_masks, tile_last, partial);
    else
        bwd_tile<NS, SYNC_MIN, false>(sh, tile, ray, W, H, gx, bg, ranges, vox_list, centers, lengths, geos, rgbs,
                                      out_T, n_contrib, dL_dcolor, dL_dT, orig, hit_masks, tile_last, partial);
}

// Per-voxel gradient = sum of the partial rows of its entries: one run of
// consecutive rows per quadrant (run_offset), quadrants ascending.  Voxels with
// more than BIG_DUP entries (listed in big_list by the duplication pass) are
// summed by a whole workgroup; those with more than HUGE_DUP are split over
// several workgroups, whose partial sums land in scratch rows that
// grad_reduce_huge_kernel adds in a fixed order (no atomics: the result does
// not depend on scheduling).

__device__ inline void write_grad(size_t v, const float* a, float* __restrict__ dgeos, float* __restrict__ drgbs,
                                  float* __restrict__ dsub)
{
    *reinterpret_cast<float4*>(dgeos + 8 * v) = make_float4(a[0], a[1], a[2], a[3]);
    *reinterpret_cast<float4*>(dgeos + 8 * v + 4) = make_float4(a[4], a[5], a[6], a[7]);
    drgbs[3 * v + 0] = a[8];
    drgbs[3 * v + 1] = a[9];
    drgbs[3 * v + 2] = a[10];
    dsub[v] = a[11];
}

__device__ inline void add_row(float (&a)[12], const float* __restrict__ partial, uint32_t row)
{
    const float4* r = reinterpret_cast<const float4*>(partial + 12 * (size_t)row);
    const float4 p0 = r[0], p1 = r[1], p2 = r[2];
    a[0] += p0.x; a[1] += p0.y; a[2] += p0.z; a[3] += p0.w;
    a[4] += p1.x; a[5] += p1.y; a[6] += p1.z; a[7] += p1.w;
    a[8] += p2.x; a[9] += p2.y; a[10] += p2.z; a[11] += p2.w;
}

// One thread per item with at most BIG_DUP entries (item v).  A run ending
// beyond `cap` was never emitted (a frame over a fixed capacity) and counts as
// zero.
__device__ inline void grad_reduce_small(const GeomState& g, int64_t N, uint32_t cap, int64_t v,
                                         const float* __restrict__ partial, float* __restrict__ dgeos,
                                         float* __restrict__ drgbs, float* __restrict__ dsub)
{
    if (v >= N) return;
    const uint32_t nd = g.n_dup[v];
    if (nd > BIG_DUP) return;
    float a[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) a[i] = 0.f;
    if (nd > 0)
    {
        const uint32_t bits = g.qbits[v];
        const uint32_t tiles = nd / (uint32_t)__popc(bits);
        for (uint32_t b = bits; b; b &= b - 1u)
        {
            const uint32_t off = run_offset(g, (size_t)v, bits, (uint32_t)__builtin_ctz(b));
            if ((uint64_t)off + tiles > cap) continue;
            // rows in pairs: both rows' loads are in flight before the adds
            // (which keep the row order)
            uint32_t e = 0;
            for (; e + 2 <= tiles; e += 2)
            {
                const float4* r0 = reinterpret_cast<const float4*>(partial + 12 * (size_t)(off + e));
                const float4 p0 = r0[0], p1 = r0[1], p2 = r0[2], q0 = r0[3], q1 = r0[4], q2 = r0[5];
                a[0] += p0.x; a[1] += p0.y; a[2] += p0.z; a[3] += p0.w;
                a[4] += p1.x; a[5] += p1.y; a[6] += p1.z; a[7] += p1.w;
                a[8] += p2.x; a[9] += p2.y; a[10] += p2.z; a[11] += p2.w;
                a[0] += q0.x; a[1] += q0.y; a[2] += q0.z; a[3] += q0.w;
                a[4] += q1.x; a[5] += q1.y; a[6] += q1.z; a[7] += q1.w;
                a[8] += q2.x; a[9] += q2.y; a[10] += q2.z; a[11] += q2.w;
            }
            if (e < tiles) add_row(a, partial, off + e);
        }
    }
    write_grad((size_t)v, a, dgeos, drgbs, dsub);
}

// Row ranges per huge item (HUGE_SPLIT_ROWS / items, at least 1).
__device__ inline uint32_t huge_split(uint32_t n_huge) { return n_huge ? max(1u, HUGE_SPLIT_ROWS / n_huge) : 1u; }

// Launched with HUGE_SPLIT_ROWS workgroups.
// Workgroup bid of nblk (the large and huge items; see grad_reduce_big_kernel).
__device__ inline void grad_reduce_big(const GeomState& g, int64_t N, uint32_t cap, uint32_t bid, uint32_t nblk,
                                       const float* __restrict__ partial, float* __restrict__ dgeos,
                                       float* __restrict__ drgbs, float* __restrict__ dsub,
                                       float* __restrict__ huge_rows)
{
    __shared__ float red[12][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // pass 0: one workgroup per large voxel, blocks striding over the queue;
    // pass 1: the huge voxels (queued from the end), each split into K row
    // ranges taken by K different workgroups (item = chunk * huge + voxel);
    // with K > 1 item `it` writes its sums to scratch row `it`
    const uint32_t counts[2] = {g.counters[2], g.counters[3]};
    const uint32_t K = huge_split(counts[1]);
    for (uint32_t pass = 0; pass < 2; ++pass)
    for (uint32_t it = bid; it < (pass == 0 ? counts[0] : counts[1] * K); it += nblk)
    {
        const uint32_t i = pass == 0 ? it : it % counts[1];
        const uint32_t v = pass == 0 ? g.big_list[i] : g.big_list[N - 1 - i];
        const uint32_t nd = g.n_dup[v];
        const uint32_t bits = g.qbits[v];
        const uint32_t tiles = nd / (uint32_t)__popc(bits);
        const uint32_t c = pass == 0 ? 0u : it / counts[1];
        const uint32_t e_lo = pass == 0 ? 0u : (uint32_t)(((uint64_t)nd * c) / K);
        const uint32_t e_hi = pass == 0 ? nd : (uint32_t)(((uint64_t)nd * (c + 1)) / K);
        float a[12];
#pragma unroll
        for (int i = 0; i < 12; ++i) a[i] = 0.f;
        uint32_t qi = 0;
        for (uint32_t b = bits; b; b &= b - 1u, ++qi)
        {
            // the part of [e_lo, e_hi) in this quadrant's run
            const uint32_t r_lo = max(e_lo, qi * tiles), r_hi = min(e_hi, (qi + 1) * tiles);
            if (r_lo >= r_hi) continue;
            const uint32_t off = run_offset(g, v, bits, (uint32_t)__builtin_ctz(b));
            if ((uint64_t)off + tiles > cap) continue;
            const uint32_t row0 = off - qi * tiles;
            for (uint32_t e = r_lo + threadIdx.x; e < r_hi; e += blockDim.x) add_row(a, partial, row0 + e);
        }
#pragma unroll
        for (int i = 0; i < 12; ++i)
        {
            const float sum = wave_sum_f32(a[i]);
            if (lane == 0) red[i][w] = sum;
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            float t[12];
#pragma unroll
            for (int i = 0; i < 12; ++i) t[i] = (red[i][0] + red[i][1]) + (red[i][2] + red[i][3]);
            if (pass == 0 || K == 1)
            {
                write_grad((size_t)v, t, dgeos, drgbs, dsub);
            }
            else
            {
                float4* row = reinterpret_cast<float4*>(huge_rows + 12 * (size_t)it);
                row[0] = make_float4(t[0], t[1], t[2], t[3]);
                row[1] = make_float4(t[4], t[5], t[6], t[7]);
                row[2] = make_float4(t[8], t[9], t[10], t[11]);
            }
        }
        __syncthreads();
    }
}

// Huge items split over K > 1 workgroups: item i adds its K scratch rows
// (c * huge + i, c ascending), one wave per item.
__global__ __launch_bounds__(64) void grad_reduce_huge_kernel(GeomState g, int64_t N,
                                                              const float* __restrict__ huge_rows,
                                                              float* __restrict__ dgeos, float* __restrict__ drgbs,
                                                              float* __restrict__ dsub)
{
    const uint32_t n_huge = g.counters[3];
    const uint32_t K = huge_split(n_huge);
    if (K == 1) return;
    const int lane = threadIdx.x;
    for (uint32_t i = blockIdx.x; i < n_huge; i += gridDim.x)
    {
        float a[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) a[k] = 0.f;
        for (uint32_t c = (uint32_t)lane; c < K; c += 64) add_row(a, huge_rows, c * n_huge + i);
        float t[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) t[k] = wave_sum_f32(a[k]);
        if (lane == 0) write_grad((size_t)g.big_list[N - 1 - i], t, dgeos, drgbs, dsub);
    }
}

// Launched with HUGE_SPLIT_ROWS workgroups.
__global__ __launch_bounds__(256) void grad_reduce_big_kernel(GeomState g, int64_t N, uint32_t cap,
                                                              const float* __restrict__ partial,
                                                              float* __restrict__ dgeos, float* __restrict__ drgbs,
                                                              float* __restrict__ dsub, float* __restrict__ huge_rows)
{
    grad_reduce_big(g, N, cap, blockIdx.x, gridDim.x, partial, dgeos, drgbs, dsub, huge_rows);
}

// Both in one launch: the first HUGE_SPLIT_ROWS workgroups take the large and
// huge items (they run longest, so they start first), the rest one small item
// per thread; the two sets share nothing, so the large items' sums overlap the
// small ones instead of following them.
__global__ __launch_bounds__(256) void grad_reduce_all_kernel(GeomState g, int64_t N, uint32_t cap,
                                                              const float* __restrict__ partial,
                                                              float* __restrict__ dgeos, float* __restrict__ drgbs,
                                                              float* __restrict__ dsub, float* __restrict__ huge_rows)
{
    if (blockIdx.x < HUGE_SPLIT_ROWS)
        grad_reduce_big(g, N, cap, blockIdx.x, HUGE_SPLIT_ROWS, partial, dgeos, drgbs, dsub, huge_rows);
    else
        grad_reduce_small(g, N, cap, (int64_t)(blockIdx.x - HUGE_SPLIT_ROWS) * blockDim.x + threadIdx.x, partial,
                          dgeos, drgbs, dsub);
}

int grad_reduce_big_launch(int64_t N, const GeomState& g, const float* partial, float* dgeos, float* drgbs,
                           float* dsub, float* huge_rows, hipStream_t stream, uint32_t cap)
{
    grad_reduce_big_kernel<<<HUGE_SPLIT_ROWS, 256, 0, stream>>>(g, N, cap, partial, dgeos, drgbs, dsub, huge_rows);
    SVR_CHECK_LAUNCH();
    grad_reduce_huge_kernel<<<64, 64, 0, stream>>>(g, N, huge_rows, dgeos, drgbs, dsub);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

int grad_reduce_launch(int64_t N, const GeomState& g, const float* partial, float* dgeos, float* drgbs, float* dsub,
                       float* huge_rows, hipStream_t stream, uint32_t cap)
{
    grad_reduce_all_kernel<<<(unsigned)(HUGE_SPLIT_ROWS + (N + 255) / 256), 256, 0, stream>>>(
        g, N, cap, partial, dgeos, drgbs, dsub, huge_rows);
    SVR_CHECK_LAUNCH();
    grad_reduce_huge_kernel<<<64, 64, 0, stream>>>(g, N, huge_rows, dgeos, drgbs, dsub);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

template <int NS>
int launch_fwd(int W, int H, float tanfovx, float tanfovy, float cx, float cy, const float* c2w, float bg,
               const GeomState& g, const BinningState& b, const ImageState& img, const float* centers,
               const float* lengths, const float* geos, const float* rgbs, float* out_color, float* out_T,
               hipStream_t stream)
{
    const int gx = tiles_x(W), gy = tiles_y(H);
    const int T = gx * gy;
    prof_begin(PROF_RENDER_FWD, stream);
    render_fwd_kernel<NS><<<(unsigned)T, 256, 0, stream>>>(
        W, H, gx, img.order, tanfovx, tanfovy, cx, cy, c2w, bg, img.ranges, b.vox_list, g.bbox, g.oct, centers, lengths, geos,
        rgbs, img.tile_last, img.n_contrib, out_color, out_T, img.hit_masks, img.ctr, img.exact_list);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_RENDER_FWD, stream);
    return SVR_OK;
}

template <int NS>
int launch_bwd(int W, int H, float tanfovx, float tanfovy, float cx, float cy, const float* c2w, float bg,
               int64_t N, int64_t R, const GeomState& g, const BinningState& b, const ImageState& img,
               const float* centers, const float* lengths, const float* geos, const float* rgbs, const float* out_T,
               const float* dL_dcolor, const float* dL_dT, float* partial, float* dgeos, float* drgbs, float* dsub,
               hipStream_t stream)
{
    const int gx = tiles_x(W), gy = tiles_y(H);
    const int T = gx * gy;
    prof_begin(PROF_RENDER_BWD, stream);
    render_bwd_kernel<NS, BWD_SYNC_MIN_MASKED><<<(unsigned)T, 256, 0, stream>>>(
        W, H, gx, img.order, tanfovx, tanfovy, cx, cy, c2w, bg, img.ranges, b.vox_list, centers, lengths, geos, rgbs,
        out_T, img.n_contrib, dL_dcolor, dL_dT, b.orig, img.hit_masks, img.tile_last, partial);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_RENDER_BWD, stream);
    prof_begin(PROF_GRAD_REDUCE, stream);
    const int rc = grad_reduce_launch(N, g, partial, dgeos, drgbs, dsub, huge_rows(partial, R), stream);
    prof_end(PROF_GRAD_REDUCE, stream);
    return rc;
}

}  // namespace svr

using namespace svr;

namespace svr {

int forward_setup(int32_t n_samp, int32_t W, int32_t H, const float* c2w, int64_t N, const float* centers,
                  const float* lengths, const float* geos, const float* rgbs, const void* geom_state,
                  const void* order_table, int64_t R, void* binning_state, void* image_state, float* out_color,
                  float* out_T, GeomState& g, BinningState& b, ImageState& img, hipStream_t stream,
                  int32_t flags)
{
    SVR_REQUIRE(n_samp >= 1 && n_samp <= MAX_N_SAMP, SVR_ERR_INVALID, "n_samp_per_vox should be in range [1, %d].",
                MAX_N_SAMP);
    SVR_REQUIRE(W > 0 && H > 0 && W <= MAX_PACKED_IMAGE_DIM && H <= MAX_PACKED_IMAGE_DIM, SVR_ERR_INVALID,
                "image size %dx%d outside [1, %d]", W, H, MAX_PACKED_IMAGE_DIM);
    SVR_REQUIRE(R >= 0 && R < (1ll << 32), SVR_ERR_INVALID, "num_rendered out of range");
    SVR_REQUIRE(image_state && out_color && out_T && c2w, SVR_ERR_INVALID, "svr_rasterize_forward: null pointer");
    SVR_REQUIRE(N == 0 || (geom_state && order_table && centers && lengths && geos && rgbs), SVR_ERR_INVALID,
                "svr_rasterize_forward: null pointer");
    SVR_REQUIRE(R == 0 || binning_state, SVR_ERR_INVALID, "svr_rasterize_forward: null binning state");
    SVR_REQUIRE(((uintptr_t)geos & 15u) == 0, SVR_ERR_INVALID, "geos must be 16-byte aligned");
    const int64_t T = (int64_t)tiles_x(W) * tiles_y(H);
    g = GeomState::carve((void*)geom_state, N, nullptr);
    OrderTable ot = OrderTable::carve((void*)order_table, N, nullptr);
    b = BinningState::carve(binning_state, R, nullptr);
    img = ImageState::carve(image_state, (int64_t)W * H, T, nullptr, R);
    if (flags & SVR_FWD_BINNED) return SVR_OK;
    return binning_forward(W, H, N, g, ot, R, b, img, stream, (flags & SVR_FWD_PAIRS_READY) != 0);
}

// Returns 1 (with the outputs zeroed) when there is nothing to composite.
int backward_setup(int32_t n_samp, int32_t W, int32_t H, const float* c2w, int64_t N, const float* geos,
                   const void* geom_state, int64_t R, const void* binning_state, const void* image_state,
                   const float* out_T, const float* dL_dcolor, const float* dL_dT, void* workspace, float* dgeos,
                   float* drgbs, float* dsub, GeomState& g, BinningState& b, ImageState& img, hipStream_t stream)
{
    SVR_REQUIRE(n_samp >= 1 && n_samp <= MAX_N_SAMP, SVR_ERR_INVALID, "n_samp_per_vox should be in range [1, %d].",
                MAX_N_SAMP);
    if (N == 0) return 1;
    SVR_REQUIRE(geom_state && image_state && out_T && dL_dcolor && dgeos && drgbs && dsub && c2w,
                SVR_ERR_INVALID, "svr_rasterize_backward: null pointer");
    SVR_REQUIRE(((uintptr_t)geos & 15u) == 0 && ((uintptr_t)dgeos & 15u) == 0, SVR_ERR_INVALID,
                "geos and dL_dgeos must be 16-byte aligned");
    if (R == 0)
    {
        if (fill_u32(dgeos, 0u, 8 * N, stream) != SVR_OK) return SVR_ERR_HIP;
        if (fill_u32(drgbs, 0u, 3 * N, stream) != SVR_OK) return SVR_ERR_HIP;
        if (fill_u32(dsub, 0u, N, stream) != SVR_OK) return SVR_ERR_HIP;
        return 1;
    }
    SVR_REQUIRE(binning_state && workspace, SVR_ERR_INVALID, "svr_rasterize_backward: null binning state or workspace");
    SVR_REQUIRE(((uintptr_t)workspace & 15u) == 0, SVR_ERR_INVALID, "backward workspace must be 16-byte aligned");
    const int64_t T = (int64_t)tiles_x(W) * tiles_y(H);
    g = GeomState::carve((void*)geom_state, N, nullptr);
    b = BinningState::carve((void*)binning_state, R, nullptr);
    img = ImageState::carve((void*)image_state, (int64_t)W * H, T, nullptr, R);
    return SVR_OK;
}

// Scratch of svr_rasterize_backward_noalloc: the workspace of
// svr_rasterize_backward (one 12-float partial row per entry, then the rows of
// split huge voxels) for up to SVR_NOALLOC_MAX_ENTRIES entries, as a device
// array of this library.  Every GPU the library runs on holds its own copy
// from the code object's load until the process ends.  Nothing in it needs
// clearing: the backward writes every partial row it later reads (zeros for
// entries no pixel used) and the huge-voxel rows before their sum.
constexpr size_t NOALLOC_SCRATCH_BYTES =
    ((sizeof(float) * 12 * (size_t)SVR_NOALLOC_MAX_ENTRIES + 255) / 256) * 256 + sizeof(float) * 12 * HUGE_SPLIT_ROWS;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// Device address of `scratch` on the current device, looked up once per
// device (hipGetSymbolAddress with that device current) and kept: the array's
// name in host code is a host-side stand-in, never a device address.
int noalloc_scratch(void** out)
{
    constexpr int MAX_DEV = 64;
    static void* ptrs[MAX_DEV] = {};
    static std::mutex mu;
    int dev = 0;
    SVR_CHECK_HIP(hipGetDevice(&dev));
    SVR_REQUIRE(dev >= 0 && dev < MAX_DEV, SVR_ERR_INVALID, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (!ptrs[dev])
    {
        void* p = nullptr;
        SVR_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SVR_REQUIRE(p != nullptr, SVR_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
        ptrs[dev] = p;
    }
    *out = ptrs[dev];
    return SVR_OK;
}

}  // namespace svr

extern "C" {

int svr_rasterize_forward(int32_t n_samp, int32_t W, int32_t H, float tanfovx, float tanfovy, float cx, float cy,
                          const float* c2w, float bg, int64_t N, const float* centers, const float* lengths,
                          const float* geos, const float* rgbs, const void* geom_state, const void* order_table,
                          int64_t R, void* binning_state, void* image_state, float* out_color, float* out_T,
                          int32_t flags, svr_stream_t stream)
{
    GeomState g;
    BinningState b;
    ImageState img;
    const int rc = forward_setup(n_samp, W, H, c2w, N, centers, lengths, geos, rgbs, geom_state, order_table, R,
                                 binning_state, image_state, out_color, out_T, g, b, img, stream, flags);
    if (rc != SVR_OK) return rc;
    switch (n_samp)
    {
    case 1: return launch_fwd<1>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, g, b, img, centers, lengths, geos, rgbs,
                                 out_color, out_T, stream);
    case 2: return launch_fwd<2>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, g, b, img, centers, lengths, geos, rgbs,
                                 out_color, out_T, stream);
    default: return launch_fwd<3>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, g, b, img, centers, lengths, geos, rgbs,
                                  out_color, out_T, stream);
    }
}

int svr_rasterize_binning(int32_t W, int32_t H, int64_t N, const void* geom_state, const void* order_table,
                          int64_t R, void* binning_state, void* image_state, int32_t flags, svr_stream_t stream)
{
    SVR_REQUIRE(W > 0 && H > 0 && W <= MAX_PACKED_IMAGE_DIM && H <= MAX_PACKED_IMAGE_DIM, SVR_ERR_INVALID,
                "image size %dx%d outside [1, %d]", W, H, MAX_PACKED_IMAGE_DIM);
    SVR_REQUIRE(R >= 0 && R < (1ll << 32), SVR_ERR_INVALID, "num_rendered out of range");
    SVR_REQUIRE(image_state, SVR_ERR_INVALID, "svr_rasterize_binning: null image state");
    SVR_REQUIRE(N == 0 || (geom_state && order_table), SVR_ERR_INVALID, "svr_rasterize_binning: null pointer");
    SVR_REQUIRE(R == 0 || binning_state, SVR_ERR_INVALID, "svr_rasterize_binning: null binning state");
    const int64_t T = (int64_t)tiles_x(W) * tiles_y(H);
    GeomState g = GeomState::carve((void*)geom_state, N, nullptr);
    OrderTable ot = OrderTable::carve((void*)order_table, N, nullptr);
    BinningState b = BinningState::carve(binning_state, R, nullptr);
    ImageState img = ImageState::carve(image_state, (int64_t)W * H, T, nullptr, R);
    return binning_forward(W, H, N, g, ot, R, b, img, stream, (flags & SVR_FWD_PAIRS_READY) != 0);
}

size_t svr_backward_workspace_size(int64_t R)
{
    return align_up(sizeof(float) * 12 * (size_t)(R > 0 ? R : 0), 256) + sizeof(float) * 12 * HUGE_SPLIT_ROWS;
}

int svr_rasterize_backward(int32_t n_samp, int32_t W, int32_t H, float tanfovx, float tanfovy, float cx, float cy,
                           const float* c2w, float bg, int64_t N, const float* centers, const float* lengths,
                           const float* geos, const float* rgbs, const void* geom_state, int64_t R,
                           const void* binning_state, const void* image_state, const float* out_T,
                           const float* dL_dcolor, const float* dL_dT, void* workspace, float* dgeos, float* drgbs,
                           float* dsub, svr_stream_t stream)
{
    GeomState g;
    BinningState b;
    ImageState img;
    const int rc = backward_setup(n_samp, W, H, c2w, N, geos, geom_state, R, binning_state, image_state, out_T,
                                  dL_dcolor, dL_dT, workspace, dgeos, drgbs, dsub, g, b, img, stream);
    if (rc != SVR_OK) return rc == 1 ? SVR_OK : rc;
    float* partial = (float*)workspace;
    switch (n_samp)
    {
    case 1: return launch_bwd<1>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, N, R, g, b, img, centers, lengths, geos,
                                 rgbs, out_T, dL_dcolor, dL_dT, partial, dgeos, drgbs, dsub, stream);
    case 2: return launch_bwd<2>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, N, R, g, b, img, centers, lengths, geos,
                                 rgbs, out_T, dL_dcolor, dL_dT, partial, dgeos, drgbs, dsub, stream);
    default: return launch_bwd<3>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, N, R, g, b, img, centers, lengths, geos,
                                  rgbs, out_T, dL_dcolor, dL_dT, partial, dgeos, drgbs, dsub, stream);
    }
}

int svr_rasterize_backward_noalloc(int32_t n_samp, int32_t W, int32_t H, float tanfovx, float tanfovy, float cx,
                                   float cy, const float* c2w, float bg, int64_t N, const float* centers,
                                   const float* lengths, const float* geos, const float* rgbs, const void* geom_state,
                                   int64_t R, const void* binning_state, const void* image_state, const float* out_T,
                                   const float* dL_dcolor, const float* dL_dT, float* dgeos, float* drgbs,
                                   float* dsub, svr_stream_t stream)
{
    SVR_REQUIRE(R >= 0 && R <= SVR_NOALLOC_MAX_ENTRIES, SVR_ERR_CAPACITY,
                "svr_rasterize_backward_noalloc: %lld entries exceed its fixed scratch of %lld entries",
                (long long)R, (long long)SVR_NOALLOC_MAX_ENTRIES);
    void* ws = nullptr;
    const int rc = noalloc_scratch(&ws);
    if (rc != SVR_OK) return rc;
    return svr_rasterize_backward(n_samp, W, H, tanfovx, tanfovy, cx, cy, c2w, bg, N, centers, lengths, geos, rgbs,
                                  geom_state, R, binning_state, image_state, out_T, dL_dcolor, dL_dT, ws, dgeos,
                                  drgbs, dsub, stream);
}

int svr_unpack_image_state(int32_t W, int32_t H, const void* image_state, int32_t* ranges, int32_t* tile_last,
                           int32_t* n_contrib, svr_stream_t stream)
{
    const int64_t T = (int64_t)tiles_x(W) * tiles_y(H);
    ImageState img = ImageState::carve((void*)image_state, (int64_t)W * H, T, nullptr);
    SVR_CHECK_HIP(hipMemcpyAsync(ranges, img.ranges, sizeof(uint32_t) * 2 * T, hipMemcpyDeviceToDevice, stream));
    SVR_CHECK_HIP(hipMemcpyAsync(tile_last, img.tile_last, sizeof(uint32_t) * T, hipMemcpyDeviceToDevice, stream));
    SVR_CHECK_HIP(hipMemcpyAsync(n_contrib, img.n_contrib, sizeof(uint32_t) * W * H, hipMemcpyDeviceToDevice,
                                 stream));
    return SVR_OK;
}

}  // extern "C"
