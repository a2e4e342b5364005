// Multi-view batch: V cameras of one octree rendered, differentiated and
// colour-evaluated together.
//
// No reference counterpart as a whole: the reference renders one camera per
// call (backends/new_cuda/new_svraster_cuda/renderer.py:39-127) and a training
// loop over views repeats every stage per view.  Per (view, voxel) pair the
// arithmetic here is the single-view path's (preprocess.cu:24-132,
// forward.cu:507-604 and 43-394, backward.cu:37-530, sh_compute.cu:55-240), so
// each view's images, n_contrib and tile lists equal the single-view results.
//
// What the batch changes:
//   * one launch per stage over all (view, voxel) pairs instead of one per view;
//   * one sort of all views' entries: keys (global tile << pos_bits) | pos, the
//     global tile id being the view's tile_base plus its tile index;
//   * the compositor forward and backward each run as one launch over the
//     tiles of all views, queued by list length (longest first; workgroup b
//     takes queue slot b, and the hardware dispatches workgroups in index
//     order), so views of different resolution share the chip with no tail
//     per view; the few tiles with a non-finite 1/rd ray are deferred to a
//     small persistent launch that loops over them;
//   * SH colours of all views are computed from one read of each voxel's 48
//     coefficients, and the SH backward sums the views' gradients in registers
//     before one write.
#include "batch_common.h"
#include "voxel_ops.h"

namespace svr {

// ---- preprocess and key duplication over (view, voxel) pairs -------------

template <typename TS>
__global__ __launch_bounds__(256) void batch_preprocess_kernel(const svr_view* __restrict__ views, int64_t N,
                                                               const float* __restrict__ centers,
                                                               const float* __restrict__ lengths,
                                                               int32_t* __restrict__ out_n_dup,
                                                               uint32_t* __restrict__ n_dup,
                                                               uint2* __restrict__ bboxes,
                                                               uint32_t* __restrict__ qbits_out,
                                                               OctBox* __restrict__ octs)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const svr_view& vw = views[blockIdx.y];
    const int W = vw.width, H = vw.height;
    const float focal_x = 0.5f * (float)W / vw.tanfovx;
    const float focal_y = 0.5f * (float)H / vw.tanfovy;
    float m[12], cw[12];
#pragma unroll
    for (int i = 0; i < 12; ++i)
    {
        m[i] = vw.w2c[i];
        cw[i] = vw.c2w[i];
    }
    uint32_t nd, bits;
    uint2 bb;
    OctBox oc;
    preprocess_voxel<TS::TX, TS::TY>(centers[3 * v + 0], centers[3 * v + 1], centers[3 * v + 2], 0.5f * lengths[v], W,
                                     H, tiles_along(W, TS::TX), tiles_along(H, TS::TY), focal_x, focal_y, vw.cx,
                                     vw.cy, m, cw, vw.near_plane, nd, bb, bits, oc);
    const size_t p = (size_t)blockIdx.y * (size_t)N + (size_t)v;
    out_n_dup[p] = (int32_t)nd;
    n_dup[p] = nd;
    // every reader of a pair's box, quadrant bits and octagon first finds
    // entries in n_dup, so a pair without entries leaves them unwritten
    if (nd == 0) return;
    bboxes[p] = bb;
    qbits_out[p] = bits;
    octs[p] = oc;
}

// forward.cu:523-577 per (view, voxel) pair: each quadrant's run at its
// rank-ordered offset within the view (recorded in qoff); pairs with many
// entries are queued (as u * N + v) for batch_duplicate_big_kernel.  Runs that
// would end beyond `cap` (a frame over a fixed capacity) emit nothing.
template <typename TS>
__global__ __launch_bounds__(256) void batch_duplicate_kernel(GeomState g, const svr_view* __restrict__ views,
                                                              int64_t N, int pos_bits,
                                                              const uint32_t* __restrict__ order_pos,
                                                              uint64_t* __restrict__ keys, uint32_t cap)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const size_t p = (size_t)blockIdx.y * (size_t)N + (size_t)v;
    const uint32_t nd = g.n_dup[p];
    if (nd == 0) return;
    const bool large = nd > BIG_DUP;
    const svr_view& vw = views[blockIdx.y];
    const int gx = tiles_along(vw.width, TS::TX), gy = tiles_along(vw.height, TS::TY);
    int txmin, tymin, txmax, tymax;
    rect_of<TS::TX, TS::TY>(g.bbox[p], gx, gy, txmin, tymin, txmax, tymax);
    const uint32_t bits = g.qbits[p];
    const uint32_t tiles = nd / (uint32_t)__popc(bits);
    const uint32_t* blk = g.pos_off + 8 * (size_t)blockIdx.y * (size_t)N;
    for (uint32_t b = bits; b; b &= b - 1u)
    {
        const uint32_t q = (uint32_t)__builtin_ctz(b);
        const uint64_t pos = order_pos[8 * v + q];
        uint32_t off = blk[pos];
        set_run_offset(g, p, bits, q, off);
        if (large || (uint64_t)off + tiles > cap) continue;
        for (int y = tymin; y <= tymax; ++y)
            for (int x = txmin; x <= txmax; ++x)
            {
                const uint32_t g = (uint32_t)vw.tile_base + (uint32_t)y * (uint32_t)gx + (uint32_t)x;
                keys[off] = ((uint64_t)g << pos_bits) | pos;
                ++off;
            }
    }
    if (large) queue_large(g.big_list, g.counters + 2, (int64_t)gridDim.y * N, nd, (uint32_t)p);
}

template <typename TS>
__global__ __launch_bounds__(256) void batch_duplicate_big_kernel(GeomState g, int V, const svr_view* __restrict__ views,
                                                                  int64_t N, int pos_bits,
                                                                  const uint32_t* __restrict__ order_pos,
                                                                  uint64_t* __restrict__ keys, uint32_t cap)
{
    // large pairs one workgroup each, then the huge ones (queued from the end
    // of big_list) spread over every workgroup
    const uint32_t count = g.counters[2], huge = g.counters[3];
    const int64_t n_items = (int64_t)V * N;
    const uint32_t K = huge ? max(1u, gridDim.x / huge) : 1u;
    for (uint32_t pass = 0; pass < 2; ++pass)
    {
    for (uint32_t it = blockIdx.x; it < (pass == 0 ? count : huge * K); it += gridDim.x)
    {
        const uint32_t i = pass == 0 ? it : it % huge;
        const uint32_t p = pass == 0 ? g.big_list[i] : g.big_list[n_items - 1 - i];
        const uint32_t u = (uint32_t)(p / (uint64_t)N), v = (uint32_t)(p - (uint64_t)u * (uint64_t)N);
        const svr_view& vw = views[u];
        const int gx = tiles_along(vw.width, TS::TX), gy = tiles_along(vw.height, TS::TY);
        const uint32_t nd = g.n_dup[p];
        int txmin, tymin, txmax, tymax;
        rect_of<TS::TX, TS::TY>(g.bbox[p], gx, gy, txmin, tymin, txmax, tymax);
        const uint32_t w = (uint32_t)(txmax - txmin + 1);
        const uint32_t tiles = w * (uint32_t)(tymax - tymin + 1);
        const uint32_t bits = g.qbits[p];
        const uint32_t c = pass == 0 ? 0u : it / huge;
        const uint32_t e_lo = pass == 0 ? 0u : (uint32_t)(((uint64_t)nd * c) / K);
        const uint32_t e_hi = pass == 0 ? nd : (uint32_t)(((uint64_t)nd * (c + 1)) / K);
        uint32_t qi = 0;
        for (uint32_t b = bits; b; b &= b - 1u, ++qi)
        {
            // the part of [e_lo, e_hi) in this quadrant's run
            const uint32_t q = (uint32_t)__builtin_ctz(b);
            const uint32_t r_lo = max(e_lo, qi * tiles), r_hi = min(e_hi, (qi + 1) * tiles);
            if (r_lo >= r_hi) continue;
            const uint32_t off = run_offset(g, p, bits, q);
            if ((uint64_t)off + tiles > cap) continue;
            const uint64_t pos = order_pos[8 * (size_t)v + q];
            for (uint32_t e = r_lo + threadIdx.x; e < r_hi; e += blockDim.x)
            {
                const uint32_t rem = e - qi * tiles;
                const uint32_t y = (uint32_t)tymin + rem / w, x = (uint32_t)txmin + rem % w;
                const uint32_t gt = (uint32_t)vw.tile_base + y * (uint32_t)gx + x;
                keys[off + rem] = ((uint64_t)gt << pos_bits) | pos;
            }
        }
    }
    }
}

// One tile of the batch forward.  The fast form (EXACT = false) defers a tile
// holding a ray whose 1/rd is not finite to exact_list (count ctr[1]).
// MAX_W: weights-only pass (no colour read or written) that raises
// max_w[view * N + voxel] to each hit's blending weight.
template <int NS, bool EXACT, typename TS, bool MAX_W>
__device__ inline void batch_fwd_tile(StageT<TS::NW>& st, uint32_t (*sv)[64], uint32_t* s_last, uint32_t g, int V,
                                      const svr_view* __restrict__ views, int64_t N, uint32_t* __restrict__ ctr,
                                      uint32_t* __restrict__ exact_list, const uint2* __restrict__ ranges,
                                      const uint32_t* __restrict__ vox_list, const uint2* __restrict__ bboxes, const OctBox* __restrict__ octs,
                                      const float* __restrict__ centers, const float* __restrict__ lengths,
                                      const float* __restrict__ geos, const float* __restrict__ rgbs,
                                      uint32_t* __restrict__ tile_last, uint32_t* __restrict__ n_contrib,
                                      float* __restrict__ out_color, float* __restrict__ out_T,
                                      float* __restrict__ max_w, uint64_t* __restrict__ hit_masks)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const TileGeom t = tile_geom<TS>(views, V, g, w, lane);
    const svr_view& vw = views[t.u];
    const PixRay ray = make_ray(t.px, t.py, t.W, t.H, vw.tanfovx, vw.tanfovy, vw.cx, vw.cy, vw.c2w);
    const bool need_exact = __syncthreads_or(!ray_inv_finite(ray)) != 0;
    if (!EXACT && need_exact)
    {
        if (threadIdx.x == 0) exact_list[atomicAdd(&ctr[1], 1u)] = g;
        return;
    }
    const uint2 range = ranges[g];
    const size_t vo = (size_t)t.u * (size_t)N;
    float T = 1.f, C0 = 0.f, C1 = 0.f, C2 = 0.f;
    uint64_t* mask_out = MAX_W ? nullptr : hit_masks + (((size_t)(range.x >> 6) + g) * TS::NW + w) * 64 + lane;
    const uint32_t last = composite_subtile<NS, EXACT, TS::NW, MAX_W, !MAX_W>(
        st, w, lane, ray, t.inside, range, t.sx0, t.sy0, t.sx1, t.sy1, vox_list, bboxes + vo, octs + vo, centers, lengths, geos,
        MAX_W ? rgbs : rgbs + 3 * vo, T, C0, C1, C2, sv, MAX_W ? reinterpret_cast<uint32_t*>(max_w + vo) : nullptr,
        mask_out);
    if (t.inside)
    {
        const size_t HW = (size_t)t.H * t.W;
        const size_t pid = (size_t)t.py * t.W + t.px;
        const size_t pb = (size_t)vw.pix_base;
        const float bg = vw.bg_color;
        n_contrib[pb + pid] = last;
        if constexpr (!MAX_W)
        {
            out_color[3 * pb + pid] = C0 + T * bg;
            out_color[3 * pb + HW + pid] = C1 + T * bg;
            out_color[3 * pb + 2 * HW + pid] = C2 + T * bg;
        }
        out_T[pb + pid] = T;
    }
    const uint32_t wm = wave_max_u32(t.inside ? last : 0u);
    if (lane == 0) s_last[w] = wm;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t m = 0;
#pragma unroll
        for (int i = 0; i < TS::NW; ++i) m = max(m, s_last[i]);
        tile_last[g] = range.x + m;
    }
}

// Fast form: workgroup b renders tile order[b] (the hardware dispatches
// workgroups in index order, so the longest lists start first).
template <int NS, typename TS, bool MAX_W>
__global__ __launch_bounds__(TS::THREADS) void batch_fwd_kernel(
    int V, const svr_view* __restrict__ views, int64_t N, const uint32_t* __restrict__ order, uint32_t* __restrict__ ctr,
    uint32_t* __restrict__ exact_list, const uint2* __restrict__ ranges, const uint32_t* __restrict__ vox_list,
    const uint2* __restrict__ bboxes, const OctBox* __restrict__ octs, const float* __restrict__ centers, const float* __restrict__ lengths,
    const float* __restrict__ geos, const float* __restrict__ rgbs, uint32_t* __restrict__ tile_last,
    uint32_t* __restrict__ n_contrib, float* __restrict__ out_color, float* __restrict__ out_T,
    float* __restrict__ max_w, uint64_t* __restrict__ hit_masks)
{
    __shared__ StageT<TS::NW> st;
    __shared__ uint32_t sv[MAX_W ? TS::NW : 1][64];
    __shared__ uint32_t s_last[TS::NW];
    const uint32_t g = order[blockIdx.x];
    batch_fwd_tile<NS, false, TS, MAX_W>(st, sv, s_last, g, V, views, N, ctr, exact_list, ranges, vox_list, bboxes, octs,
                                         centers, lengths, geos, rgbs, tile_last, n_contrib, out_color, out_T, max_w,
                                         hit_masks);
}

// Exact form: a few workgroups loop over the deferred tiles (head ctr[2]).
template <int NS, typename TS, bool MAX_W>
__global__ __launch_bounds__(TS::THREADS) void batch_fwd_exact_kernel(
    int V, const svr_view* __restrict__ views, int64_t N, uint32_t* __restrict__ ctr, uint32_t* __restrict__ exact_list,
    const uint2* __restrict__ ranges, const uint32_t* __restrict__ vox_list, const uint2* __restrict__ bboxes, const OctBox* __restrict__ octs,
    const float* __restrict__ centers, const float* __restrict__ lengths, const float* __restrict__ geos,
    const float* __restrict__ rgbs, uint32_t* __restrict__ tile_last, uint32_t* __restrict__ n_contrib,
    float* __restrict__ out_color, float* __restrict__ out_T, float* __restrict__ max_w,
    uint64_t* __restrict__ hit_masks)
{
    __shared__ StageT<TS::NW> st;
    __shared__ uint32_t sv[MAX_W ? TS::NW : 1][64];
    __shared__ uint32_t s_last[TS::NW];
    __shared__ uint32_t s_i;
    const uint32_t count = ctr[1];
    while (true)
    {
        if (threadIdx.x == 0) s_i = atomicAdd(&ctr[2], 1u);
        __syncthreads();
        const uint32_t i = __builtin_amdgcn_readfirstlane(s_i);
        if (i >= count) break;
        const uint32_t g = __builtin_amdgcn_readfirstlane(exact_list[i]);
        batch_fwd_tile<NS, true, TS, MAX_W>(st, sv, s_last, g, V, views, N, ctr, exact_list, ranges, vox_list, bboxes, octs,
                                            centers, lengths, geos, rgbs, tile_last, n_contrib, out_color, out_T,
                                            max_w, hit_masks);
        __syncthreads();
    }
}

__device__ inline void add_row_batch(float (&a)[12], const float* __restrict__ partial, uint32_t row)
{
    const float4* r = reinterpret_cast<const float4*>(partial + 12 * (size_t)row);
    const float4 p0 = r[0], p1 = r[1], p2 = r[2];
    a[0] += p0.x; a[1] += p0.y; a[2] += p0.z; a[3] += p0.w;
    a[4] += p1.x; a[5] += p1.y; a[6] += p1.z; a[7] += p1.w;
    a[8] += p2.x; a[9] += p2.y; a[10] += p2.z; a[11] += p2.w;
}

// Per voxel, over the views in view order (the order in which a per-view
// training loop accumulates them): each (view, voxel) pair's gradient is the
// sum of its entries' partial rows (quadrant runs ascending, each run in
// emission order; pairs with more than BIG_DUP entries were summed by
// grad_reduce_big_launch into pair_geo / drgbs / pair_sub), its colour
// gradient goes to drgbs (zeros for a pair without entries), and the density
// gradient and subdivision priority are summed over the views.  A workgroup
// takes RED_VOX consecutive voxels and RED_VIEWS views at a time, one pair
// per thread (so every pair's chain of dependent loads runs in parallel); the
// pairs' density terms meet in LDS, where the first RED_VOX threads add them
// to their voxel's running sums in view order.  No per-pair density rows are
// written and read back.  A run ending beyond `cap` was never emitted and
// counts as zero.
constexpr int RED_VOX = 32, RED_VIEWS = 8;
static_assert(RED_VOX * RED_VIEWS == 256, "one pair per thread");

__global__ __launch_bounds__(256) void batch_reduce_kernel(int V, GeomState g, int64_t N, uint32_t cap,
                                                           const float* __restrict__ partial,
                                                           const float* __restrict__ pair_geo,
                                                           const float* __restrict__ pair_sub,
                                                           float* __restrict__ drgbs, float* __restrict__ dgeos,
                                                           float* __restrict__ dsub)
{
    __shared__ float part[RED_VIEWS][9][RED_VOX];
    const int t = threadIdx.x, xi = t % RED_VOX, vj = t / RED_VOX;
    const int64_t v = (int64_t)blockIdx.x * RED_VOX + xi;
    float run[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) run[i] = 0.f;
    for (int u0 = 0; u0 < V; u0 += RED_VIEWS)
    {
        const int u = u0 + vj;
        float a[12];
#pragma unroll
        for (int i = 0; i < 12; ++i) a[i] = 0.f;
        if (v < N && u < V)
        {
            const size_t p = (size_t)u * (size_t)N + (size_t)v;
            const uint32_t nd = g.n_dup[p];
            if (nd > BIG_DUP)
            {
                const float4 b0 = reinterpret_cast<const float4*>(pair_geo + 8 * p)[0];
                const float4 b1 = reinterpret_cast<const float4*>(pair_geo + 8 * p)[1];
                a[0] = b0.x; a[1] = b0.y; a[2] = b0.z; a[3] = b0.w;
                a[4] = b1.x; a[5] = b1.y; a[6] = b1.z; a[7] = b1.w;
                a[11] = pair_sub[p];
            }
            else
            {
                if (nd > 0)
                {
                    const uint32_t bits = g.qbits[p];
                    const uint32_t tiles = nd / (uint32_t)__popc(bits);
                    for (uint32_t b = bits; b; b &= b - 1u)
                    {
                        const uint32_t off = run_offset(g, p, bits, (uint32_t)__builtin_ctz(b));
                        if ((uint64_t)off + tiles > cap) continue;
                        for (uint32_t e = 0; e < tiles; ++e) add_row_batch(a, partial, off + e);
                    }
                }
                drgbs[3 * p + 0] = a[8];
                drgbs[3 * p + 1] = a[9];
                drgbs[3 * p + 2] = a[10];
            }
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) part[vj][c][xi] = a[c];
        part[vj][8][xi] = a[11];
        __syncthreads();
        if (vj == 0)
        {
            const int nk = min(RED_VIEWS, V - u0);
            for (int k = 0; k < nk; ++k)
#pragma unroll
                for (int c = 0; c < 9; ++c) run[c] += part[k][c][xi];
        }
        __syncthreads();
    }
    if (vj == 0 && v < N)
    {
        reinterpret_cast<float4*>(dgeos + 8 * v)[0] = make_float4(run[0], run[1], run[2], run[3]);
        reinterpret_cast<float4*>(dgeos + 8 * v)[1] = make_float4(run[4], run[5], run[6], run[7]);
        dsub[v] = run[8];
    }
}

// ---- SH colour of all views ----------------------------------------------

constexpr int BSH_PAD = 49;

// A workgroup takes SHF_VOX consecutive voxels: their coefficients and
// centres are staged in LDS with contiguous loads, then each thread evaluates
// one (view, voxel) pair at a time, SHF_VIEWS views per round, where the pair
// has entries (sh_compute.cu:55-106 arithmetic); zeros elsewhere.
constexpr int SHF_VOX = 32, SHF_VIEWS = 8;
static_assert(SHF_VOX * SHF_VIEWS == 256, "one pair per thread");

__global__ __launch_bounds__(256) void batch_sh_fwd_kernel(int deg, int V, const svr_view* __restrict__ views,
                                                           int64_t N, const float* __restrict__ centers,
                                                           const uint32_t* __restrict__ n_dup,
                                                           const float* __restrict__ sh0,
                                                           const float* __restrict__ shs, float* __restrict__ rgbs)
{
    __shared__ float s0[SHF_VOX * 3], s1[SHF_VOX * 45], cen[SHF_VOX * 3];
    const int t = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.x * SHF_VOX;
    const int nv = (int)min((int64_t)SHF_VOX, N - v0);
    for (int e = t; e < nv * 45; e += 256) s1[e] = shs[45 * v0 + e];
    if (t < nv * 3)
    {
        s0[t] = sh0[3 * v0 + t];
        cen[t] = centers[3 * v0 + t];
    }
    __syncthreads();
    const int xi = t % SHF_VOX, vj = t / SHF_VOX;
    if (xi >= nv) return;
    const int64_t v = v0 + xi;
    const float c[3] = {cen[3 * xi + 0], cen[3 * xi + 1], cen[3 * xi + 2]};
    for (int u = vj; u < V; u += SHF_VIEWS)
    {
        const size_t p = (size_t)u * (size_t)N + (size_t)v;
        float* out = rgbs + 3 * p;
        if (n_dup[p] == 0)
        {
            out[0] = 0.f; out[1] = 0.f; out[2] = 0.f;
            continue;
        }
        const svr_view& vw = views[u];
        const float cam[3] = {vw.c2w[3], vw.c2w[7], vw.c2w[11]};
        float x, y, z, b[16], res[3];
        sh_dir_c(c, cam, x, y, z);
        const int nb = sh_basis(deg, x, y, z, b);
        sh_eval_rows(nb, b, x, y, z, s0 + 3 * xi, s1 + 45 * xi, res);
        out[0] = res[0]; out[1] = res[1]; out[2] = res[2];
    }
}

// sh_compute.cu:110-176 per view, summed over the views in registers; the
// 48 sums of the wave's 64 voxels leave through LDS with contiguous stores.
__global__ __launch_bounds__(64) void batch_sh_bwd_kernel(int deg, int V, const svr_view* __restrict__ views,
                                                          int64_t N, const float* __restrict__ centers,
                                                          const uint32_t* __restrict__ n_dup,
                                                          const float* __restrict__ dL_drgbs,
                                                          float* __restrict__ dsh0, float* __restrict__ dshs)
{
    __shared__ float buf[64 * BSH_PAD];
    const int lane = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.x * 64;
    const int nv = (int)min((int64_t)64, N - v0);
    if (lane < nv)
    {
        const int64_t v = v0 + lane;
        float a[48];
#pragma unroll
        for (int k = 0; k < 48; ++k) a[k] = 0.f;
        const float c[3] = {centers[3 * v + 0], centers[3 * v + 1], centers[3 * v + 2]};
        // the counts and colour gradients of a batch of views are loaded
        // together (one round trip per batch)
        constexpr int VB = 8;
        for (int u0 = 0; u0 < V; u0 += VB)
        {
            uint32_t nd[VB];
            float gr[VB][3];
#pragma unroll
            for (int j = 0; j < VB; ++j)
            {
                const size_t p = (size_t)min(u0 + j, V - 1) * (size_t)N + (size_t)v;
                nd[j] = u0 + j < V ? n_dup[p] : 0u;
                gr[j][0] = dL_drgbs[3 * p + 0];
                gr[j][1] = dL_drgbs[3 * p + 1];
                gr[j][2] = dL_drgbs[3 * p + 2];
            }
#pragma unroll
            for (int j = 0; j < VB; ++j)
            {
                if (nd[j] == 0) continue;
                const svr_view& vw = views[u0 + j];
                const float cam[3] = {vw.c2w[3], vw.c2w[7], vw.c2w[11]};
                float x, y, z, b[16];
                sh_dir_c(c, cam, x, y, z);
                const int nb = sh_basis(deg, x, y, z, b);
#pragma unroll
                for (int ch = 0; ch < 3; ++ch)
                {
                    const float g = gr[j][ch];
                    a[ch] += SH_C0 * g;
#pragma unroll
                    for (int k = 1; k < 16; ++k)
                        if (k < nb) a[3 * k + ch] += b[k] * g;
                }
            }
        }
        float* row = buf + lane * BSH_PAD;
#pragma unroll
        for (int k = 0; k < 48; ++k) row[k] = a[k];
    }
    wave_sync();
    for (int e = lane; e < nv * 3; e += 64)
    {
        const int k = e / 3, f = e - 3 * k;
        dsh0[3 * v0 + e] = buf[k * BSH_PAD + f];
    }
    for (int e = lane; e < nv * 45; e += 64)
    {
        const int k = e / 45, f = e - 45 * k;
        dshs[45 * v0 + e] = buf[k * BSH_PAD + 3 + f];
    }
}

// ---- host side --------------------------------------------------------------

// Binning tile of a tile_shape code (SVR_TILE_*).
inline void tile_dims(int shape, int* tx, int* ty)
{
    *tx = shape == SVR_TILE_8X8 ? 8 : (shape == SVR_TILE_32X8 ? 32 : 16);
    *ty = shape == SVR_TILE_16X16 ? 16 : 8;
}

// Validates the view table (sizes, running sums, one tile shape) and
// returns the batch's tiles, pixels and tile shape.
int check_views(int V, const svr_view* views, int64_t* T_out, int64_t* P_out, int* shape_out = nullptr)
{
    SVR_REQUIRE(V >= 1 && V <= 65536, SVR_ERR_INVALID, "number of views %d out of range", V);
    SVR_REQUIRE(views != nullptr, SVR_ERR_INVALID, "null view table");
    const int shape = views[0].tile_shape;
    SVR_REQUIRE(shape == SVR_TILE_16X16 || shape == SVR_TILE_8X8 || shape == SVR_TILE_32X8, SVR_ERR_INVALID,
                "unknown tile shape %d", shape);
    int tw, th;
    tile_dims(shape, &tw, &th);
    int64_t T = 0, P = 0;
    for (int u = 0; u < V; ++u)
    {
        const svr_view& vw = views[u];
        SVR_REQUIRE(vw.width > 0 && vw.height > 0 && vw.width <= MAX_PACKED_IMAGE_DIM &&
                        vw.height <= MAX_PACKED_IMAGE_DIM,
                    SVR_ERR_INVALID, "view %d: image size %dx%d outside [1, %d]", u, vw.width, vw.height,
                    MAX_PACKED_IMAGE_DIM);
        SVR_REQUIRE(vw.tile_shape == shape, SVR_ERR_INVALID, "view %d: all views need the same tile shape", u);
        SVR_REQUIRE(vw.bwd_window == views[0].bwd_window && vw.bwd_sync == views[0].bwd_sync, SVR_ERR_INVALID,
                    "view %d: all views need the same backward window and sync threshold", u);
        SVR_REQUIRE(vw.bwd_window == 0 || vw.bwd_window == 2 || vw.bwd_window == 4 || vw.bwd_window == 6 ||
                        vw.bwd_window == 8,
                    SVR_ERR_INVALID, "view %d: backward window %d not in {2, 4, 6, 8}", u, vw.bwd_window);
        SVR_REQUIRE(vw.bwd_sync == 0 || vw.bwd_sync == 4 || vw.bwd_sync == 8 || vw.bwd_sync == 16, SVR_ERR_INVALID,
                    "view %d: backward sync threshold %d not in {4, 8, 16}", u, vw.bwd_sync);
        SVR_REQUIRE(vw.tile_base == T && vw.pix_base == P, SVR_ERR_INVALID,
                    "view %d: tile_base / pix_base must be the running sums of the previous views", u);
        T += (int64_t)tiles_along(vw.width, tw) * tiles_along(vw.height, th);
        P += (int64_t)vw.width * vw.height;
    }
    SVR_REQUIRE(T < (1ll << 31), SVR_ERR_INVALID, "too many tiles in the batch");
    *T_out = T;
    *P_out = P;
    if (shape_out) *shape_out = shape;
    return SVR_OK;
}

// Runs BODY with TS bound to the tile shape of a SVR_TILE_* code.
#define SVR_WITH_TILE(shape, BODY)                                   \
    switch (shape)                                                   \
    {                                                                \
    case SVR_TILE_8X8: { using TS = TileShape<8, 8>; BODY; } break;   \
    case SVR_TILE_32X8: { using TS = TileShape<32, 8>; BODY; } break; \
    default: { using TS = Tile16; BODY; } break;                     \
    }

template <int NS, typename TS, bool MAX_W>
int batch_fwd_launch(int V, const svr_view* views, int64_t N, int64_t T, const BinningState& b,
                     const BatchImageState& img, const GeomState& g, const float* centers, const float* lengths,
                     const float* geos, const float* rgbs, float* out_color, float* out_T, float* max_w,
                     hipStream_t stream)
{
    batch_fwd_kernel<NS, TS, MAX_W><<<(unsigned)T, TS::THREADS, 0, stream>>>(
        V, views, N, img.order, img.ctr, img.exact_list, img.ranges, b.vox_list, g.bbox, g.oct, centers, lengths, geos, rgbs,
        img.tile_last, img.n_contrib, out_color, out_T, max_w, img.hit_masks);
    SVR_CHECK_LAUNCH();
    batch_fwd_exact_kernel<NS, TS, MAX_W><<<(unsigned)std::min<int64_t>(EXACT_GRID, T), TS::THREADS, 0, stream>>>(
        V, views, N, img.ctr, img.exact_list, img.ranges, b.vox_list, g.bbox, g.oct, centers, lengths, geos, rgbs,
        img.tile_last, img.n_contrib, out_color, out_T, max_w, img.hit_masks);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

template <int NS, typename TS>
int batch_fwd_dispatch(int V, const svr_view* views, int64_t N, int64_t T, const BinningState& b,
                       const BatchImageState& img, const GeomState& g, const float* centers, const float* lengths,
                       const float* geos, const float* rgbs, float* out_color, float* out_T, float* max_w,
                       hipStream_t stream)
{
    if (max_w)
        return batch_fwd_launch<NS, TS, true>(V, views, N, T, b, img, g, centers, lengths, geos, rgbs, out_color,
                                              out_T, max_w, stream);
    return batch_fwd_launch<NS, TS, false>(V, views, N, T, b, img, g, centers, lengths, geos, rgbs, out_color, out_T,
                                           nullptr, stream);
}

}  // namespace svr

using namespace svr;

extern "C" {

size_t svr_batch_geom_state_size(int32_t V, int64_t N)
{
    size_t bytes = 0;
    GeomState::carve(nullptr, (V > 0 ? V : 0) * (N > 0 ? N : 0), &bytes);
    return bytes;
}

size_t svr_batch_image_state_size(int64_t T, int64_t P, int64_t R)
{
    size_t bytes = 0;
    BatchImageState::carve(nullptr, T > 0 ? T : 0, P > 0 ? P : 0, &bytes, R > 0 ? R : 0);
    return bytes;
}

size_t svr_batch_backward_workspace_size(int32_t V, int64_t N, int64_t R)
{
    size_t bytes = 0;
    BatchBwdWork::carve(nullptr, (V > 0 ? V : 0) * (N > 0 ? N : 0), R > 0 ? R : 0, &bytes);
    return bytes;
}

int svr_batch_preprocess(int32_t V, const svr_view* views_host, const svr_view* views_dev, int64_t N,
                         const float* centers, const float* lengths, int32_t* out_n_dup, void* geom_state,
                         svr_stream_t stream)
{
    int64_t T = 0, P = 0;
    int shape = 0;
    int rc = check_views(V, views_host, &T, &P, &shape);
    if (rc != SVR_OK) return rc;
    SVR_REQUIRE(N >= 0 && N <= (int64_t)VOX_ID_MASK + 1, SVR_ERR_INVALID, "voxel count %lld out of range",
                (long long)N);
    SVR_REQUIRE((int64_t)V * N < (1ll << 32), SVR_ERR_INVALID, "views x voxels must stay below 2^32");
    if (N == 0) return SVR_OK;
    SVR_REQUIRE(views_dev && centers && lengths && out_n_dup && geom_state, SVR_ERR_INVALID,
                "svr_batch_preprocess: null pointer");
    GeomState g = GeomState::carve(geom_state, (int64_t)V * N, nullptr);
    prof_begin(PROF_PREPROCESS, stream);
    SVR_WITH_TILE(shape, (batch_preprocess_kernel<TS><<<dim3((unsigned)((N + 255) / 256), (unsigned)V), 256, 0,
                                                        stream>>>(views_dev, N, centers, lengths, out_n_dup, g.n_dup,
                                                                  g.bbox, g.qbits, g.oct)));
    SVR_CHECK_LAUNCH();
    prof_end(PROF_PREPROCESS, stream);
    return SVR_OK;
}

int svr_batch_sh_compute(int32_t deg, int32_t V, const svr_view* views_dev, int64_t N, const float* centers,
                         const void* geom_state, const float* sh0, const float* shs, float* rgbs, svr_stream_t stream)
{
    SVR_REQUIRE(deg >= 0 && deg <= 3, SVR_ERR_INVALID, "SH degree %d outside [0, 3]", deg);
    SVR_REQUIRE(V >= 1, SVR_ERR_INVALID, "number of views must be positive");
    if (N == 0) return SVR_OK;
    SVR_REQUIRE(views_dev && centers && geom_state && sh0 && shs && rgbs, SVR_ERR_INVALID,
                "svr_batch_sh_compute: null pointer");
    GeomState g = GeomState::carve((void*)geom_state, (int64_t)V * N, nullptr);
    prof_begin(PROF_SH_FWD, stream);
    batch_sh_fwd_kernel<<<(unsigned)((N + SHF_VOX - 1) / SHF_VOX), 256, 0, stream>>>(deg, V, views_dev, N, centers,
                                                                                    g.n_dup, sh0,
                                                                     shs, rgbs);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_SH_FWD, stream);
    return SVR_OK;
}

int svr_batch_sh_compute_bw(int32_t deg, int32_t V, const svr_view* views_dev, int64_t N, const float* centers,
                            const void* geom_state, const float* dL_drgbs, float* dsh0, float* dshs,
                            svr_stream_t stream)
{
    SVR_REQUIRE(deg >= 0 && deg <= 3, SVR_ERR_INVALID, "SH degree %d outside [0, 3]", deg);
    SVR_REQUIRE(V >= 1, SVR_ERR_INVALID, "number of views must be positive");
    if (N == 0) return SVR_OK;
    SVR_REQUIRE(views_dev && centers && geom_state && dL_drgbs && dsh0 && dshs, SVR_ERR_INVALID,
                "svr_batch_sh_compute_bw: null pointer");
    GeomState g = GeomState::carve((void*)geom_state, (int64_t)V * N, nullptr);
    prof_begin(PROF_SH_BWD, stream);
    batch_sh_bwd_kernel<<<(unsigned)((N + 63) / 64), 64, 0, stream>>>(deg, V, views_dev, N, centers, g.n_dup,
                                                                     dL_drgbs, dsh0, dshs);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_SH_BWD, stream);
    return SVR_OK;
}

}  // extern "C"

namespace svr {

// R: the exact entry count (eager form) or the capacity the buffers were
// sized for (capacity form, overflow != nullptr: the count is read on the
// device and a frame over the capacity renders no entries).
// max_w != nullptr: the weights-only pass (rgbs and out_color unused).
static int batch_forward_impl(int32_t n_samp, int32_t V, const svr_view* views_host, const svr_view* views_dev,
                              int64_t N, const float* centers, const float* lengths, const float* geos,
                              const float* rgbs, const void* geom_state, const void* order_table, int64_t R,
                              void* binning_state, void* image_state, float* out_color, float* out_T,
                              uint32_t* overflow, hipStream_t stream, float* max_w = nullptr)
{
    SVR_REQUIRE(n_samp >= 1 && n_samp <= MAX_N_SAMP, SVR_ERR_INVALID, "n_samp_per_vox should be in range [1, %d].",
                MAX_N_SAMP);
    int64_t T = 0, P = 0;
    int shape = 0;
    int rc = check_views(V, views_host, &T, &P, &shape);
    if (rc != SVR_OK) return rc;
    SVR_REQUIRE(R >= 0 && R < (1ll << 32), SVR_ERR_INVALID, "num_rendered out of range");
    SVR_REQUIRE(views_dev && image_state && (out_color || max_w) && out_T, SVR_ERR_INVALID,
                "svr_batch_rasterize_forward: null pointer");
    SVR_REQUIRE(N == 0 || (geom_state && order_table && centers && lengths && geos && (rgbs || max_w)),
                SVR_ERR_INVALID, "svr_batch_rasterize_forward: null pointer");
    if (max_w && N > 0)
        SVR_REQUIRE(fill_u32(max_w, 0u, (int64_t)V * N, stream) == SVR_OK, SVR_ERR_HIP, "fill max_w failed");
    SVR_REQUIRE(R == 0 || binning_state, SVR_ERR_INVALID, "svr_batch_rasterize_forward: null binning state");
    SVR_REQUIRE(((uintptr_t)geos & 15u) == 0, SVR_ERR_INVALID, "geos must be 16-byte aligned");
    BatchImageState img = BatchImageState::carve(image_state, T, P, nullptr, R);
    GeomState g = GeomState::carve((void*)geom_state, (int64_t)V * N, nullptr);
    OrderTable ot = OrderTable::carve((void*)order_table, N, nullptr);
    BinningState b = BinningState::carve(binning_state, R, nullptr);

    prof_begin(PROF_BINNING, stream);
    SVR_REQUIRE(fill_u32(img.ranges, 0u, 2 * T, stream) == SVR_OK, SVR_ERR_HIP, "fill ranges failed");
    SVR_REQUIRE(fill_u32(img.ctr, 0u, 8, stream) == SVR_OK, SVR_ERR_HIP, "fill counters failed");
    const int pos_bits = (int)required_bits((uint64_t)8 * (N > 0 ? N : 1));
    const int tile_bits = (int)required_bits((uint64_t)T);
    SVR_REQUIRE(radix_passes(tile_bits) <= SORT_FRAME_PASSES, SVR_ERR_INVALID, "%lld tiles are too many to sort",
                (long long)T);
    const bool capacity_mode = overflow != nullptr;
    const uint32_t cap = capacity_mode ? (uint32_t)R : 0xFFFFFFFFu;
    const uint32_t* R_dev = capacity_mode ? g.counters + 4 : nullptr;
    if (capacity_mode && N > 0)
    {
        rc = clamp_count_launch(g.counters, (uint64_t)R, overflow, stream);
        if (rc != SVR_OK) return rc;
    }
    if (R > 0)
    {
        rc = pair_offsets_launch(V, N, g, ot.pos, stream, false);
        if (rc != SVR_OK) return rc;
        // entries are emitted view by view in rank order, so the sort only
        // orders the (global) tile bits; emit into the buffer that makes the
        // sorted result land in b.keys / b.orig
        const bool odd = radix_passes(tile_bits) & 1;
        uint64_t* k0 = odd ? b.keys_alt : b.keys;
        uint32_t* o0 = odd ? b.orig_alt : b.orig;
        SVR_WITH_TILE(shape, (batch_duplicate_kernel<TS><<<dim3((unsigned)((N + 255) / 256), (unsigned)V), 256, 0,
                                                           stream>>>(g, views_dev, N, pos_bits, ot.pos, k0, cap)));
        SVR_CHECK_LAUNCH();
        SVR_WITH_TILE(shape, (batch_duplicate_big_kernel<TS><<<1024, 256, 0, stream>>>(g, V, views_dev, N, pos_bits,
                                                                                       ot.pos, k0, cap)));
        SVR_CHECK_LAUNCH();
        uint64_t* sorted = nullptr;
        uint32_t* sorted_orig = nullptr;
        rc = radix_sort_ex(k0, odd ? b.keys : b.keys_alt, o0, odd ? b.orig : b.orig_alt, R, tile_bits, b.counts,
                           stream, &sorted, &sorted_orig, R_dev, pos_bits, true);
        if (rc != SVR_OK) return rc;
        SVR_REQUIRE(sorted == b.keys && sorted_orig == b.orig, SVR_ERR_HIP, "sort ended in the wrong buffer");
        rc = ranges_launch(b.keys, R, pos_bits, ot.pair, img.ranges, b.vox_list, stream, R_dev);
        if (rc != SVR_OK) return rc;
    }
    // queue order: longest tile lists first
    rc = tile_queue_launch(T, img.ranges, img.qkeys, img.qkeys_alt, img.order, img.order_alt, img.qcounts, stream);
    if (rc != SVR_OK) return rc;
    prof_end(PROF_BINNING, stream);

    prof_begin(PROF_RENDER_FWD, stream);
    SVR_WITH_TILE(shape, (rc = n_samp == 1   ? batch_fwd_dispatch<1, TS>(V, views_dev, N, T, b, img, g, centers,
                                                                            lengths, geos, rgbs, out_color, out_T,
                                                                            max_w, stream)
                               : n_samp == 2 ? batch_fwd_dispatch<2, TS>(V, views_dev, N, T, b, img, g, centers,
                                                                            lengths, geos, rgbs, out_color, out_T,
                                                                            max_w, stream)
                                             : batch_fwd_dispatch<3, TS>(V, views_dev, N, T, b, img, g, centers,
                                                                            lengths, geos, rgbs, out_color, out_T,
                                                                            max_w, stream)));
    prof_end(PROF_RENDER_FWD, stream);
    return rc;
}

static int batch_backward_impl(int32_t n_samp, int32_t V, const svr_view* views_host, const svr_view* views_dev,
                               int64_t N, const float* centers, const float* lengths, const float* geos,
                               const float* rgbs, const void* geom_state, int64_t R, const void* binning_state,
                               const void* image_state, const float* out_T, const float* dL_dcolor, const float* dL_dT,
                               void* workspace, float* dgeos, float* drgbs, float* dsub, bool capacity_mode,
                               hipStream_t stream)
{
    SVR_REQUIRE(n_samp >= 1 && n_samp <= MAX_N_SAMP, SVR_ERR_INVALID, "n_samp_per_vox should be in range [1, %d].",
                MAX_N_SAMP);
    int64_t T = 0, P = 0;
    int shape = 0;
    int rc = check_views(V, views_host, &T, &P, &shape);
    if (rc != SVR_OK) return rc;
    if (N == 0) return SVR_OK;
    SVR_REQUIRE(views_dev && geom_state && image_state && out_T && dL_dcolor && dgeos && drgbs && dsub &&
                    workspace,
                SVR_ERR_INVALID, "svr_batch_rasterize_backward: null pointer");
    SVR_REQUIRE(((uintptr_t)geos & 15u) == 0 && ((uintptr_t)dgeos & 15u) == 0 && ((uintptr_t)workspace & 15u) == 0,
                SVR_ERR_INVALID, "geos, dL_dgeos and workspace must be 16-byte aligned");
    SVR_REQUIRE(R == 0 || binning_state, SVR_ERR_INVALID, "svr_batch_rasterize_backward: null binning state");
    const int64_t VN = (int64_t)V * N;
    GeomState g = GeomState::carve((void*)geom_state, VN, nullptr);
    BinningState b = BinningState::carve((void*)binning_state, R, nullptr);
    BatchImageState img = BatchImageState::carve((void*)image_state, T, P, nullptr, R);
    BatchBwdWork wk = BatchBwdWork::carve(workspace, VN, R, nullptr);

    if (R > 0)
    {
        if (fill_u32(img.ctr + 3, 0u, 2, stream) != SVR_OK) return SVR_ERR_HIP;
        prof_begin(PROF_RENDER_BWD, stream);
        const int kw = views_host[0].bwd_window ? views_host[0].bwd_window : BATCH_BWD_WINDOW_DEFAULT;
        const int sync = views_host[0].bwd_sync ? views_host[0].bwd_sync : BATCH_BWD_SYNC_DEFAULT;
        SVR_WITH_TILE(shape, (rc = batch_bwd_run<TS>(n_samp, kw, sync, V, views_dev, N, T, b, img, centers, lengths,
                                                     geos, rgbs, out_T, dL_dcolor, dL_dT, wk.partial, stream)));
        prof_end(PROF_RENDER_BWD, stream);
        if (rc != SVR_OK) return rc;
    }
    prof_begin(PROF_GRAD_REDUCE, stream);
    if (R > 0)
    {
        const uint32_t cap = capacity_mode ? (uint32_t)R : 0xFFFFFFFFu;
        rc = grad_reduce_big_launch(VN, g, wk.partial, wk.pair_geo, drgbs, wk.pair_sub, wk.huge, stream, cap);
        if (rc != SVR_OK) return rc;
        batch_reduce_kernel<<<(unsigned)((N + RED_VOX - 1) / RED_VOX), 256, 0, stream>>>(
            V, g, N, cap, wk.partial, wk.pair_geo, wk.pair_sub, drgbs, dgeos, dsub);
        SVR_CHECK_LAUNCH();
    }
    else
    {
        if (fill_u32(dgeos, 0u, 8 * N, stream) != SVR_OK) return SVR_ERR_HIP;
        if (fill_u32(drgbs, 0u, 3 * VN, stream) != SVR_OK) return SVR_ERR_HIP;
        if (fill_u32(dsub, 0u, N, stream) != SVR_OK) return SVR_ERR_HIP;
    }
    prof_end(PROF_GRAD_REDUCE, stream);
    return SVR_OK;
}

}  // namespace svr

extern "C" {

int svr_batch_rasterize_forward(int32_t n_samp, int32_t V, const svr_view* views_host, const svr_view* views_dev,
                                int64_t N, const float* centers, const float* lengths, const float* geos,
                                const float* rgbs, const void* geom_state, const void* order_table, int64_t R,
                                void* binning_state, void* image_state, float* out_color, float* out_T,
                                svr_stream_t stream)
{
    return batch_forward_impl(n_samp, V, views_host, views_dev, N, centers, lengths, geos, rgbs, geom_state,
                              order_table, R, binning_state, image_state, out_color, out_T, nullptr, stream);
}

int svr_batch_rasterize_forward_capacity(int32_t n_samp, int32_t V, const svr_view* views_host,
                                         const svr_view* views_dev, int64_t N, const float* centers,
                                         const float* lengths, const float* geos, const float* rgbs,
                                         const void* geom_state, const void* order_table, int64_t capacity,
                                         void* binning_state, void* image_state, float* out_color, float* out_T,
                                         uint32_t* overflow_dev, svr_stream_t stream)
{
    SVR_REQUIRE(overflow_dev != nullptr, SVR_ERR_INVALID, "svr_batch_rasterize_forward_capacity: null overflow flag");
    return batch_forward_impl(n_samp, V, views_host, views_dev, N, centers, lengths, geos, rgbs, geom_state,
                              order_table, capacity, binning_state, image_state, out_color, out_T, overflow_dev,
                              stream);
}

int svr_batch_rasterize_backward(int32_t n_samp, int32_t V, const svr_view* views_host, const svr_view* views_dev,
                                 int64_t N, const float* centers, const float* lengths, const float* geos,
                                 const float* rgbs, const void* geom_state, int64_t R, const void* binning_state,
                                 const void* image_state, const float* out_T, const float* dL_dcolor,
                                 const float* dL_dT, void* workspace, float* dgeos, float* drgbs, float* dsub,
                                 svr_stream_t stream)
{
    return batch_backward_impl(n_samp, V, views_host, views_dev, N, centers, lengths, geos, rgbs, geom_state, R,
                               binning_state, image_state, out_T, dL_dcolor, dL_dT, workspace, dgeos, drgbs, dsub,
                               false, stream);
}

int svr_batch_rasterize_backward_capacity(int32_t n_samp, int32_t V, const svr_view* views_host,
                                          const svr_view* views_dev, int64_t N, const float* centers,
                                          const float* lengths, const float* geos, const float* rgbs,
                                          const void* geom_state, int64_t capacity, const void* binning_state,
                                          const void* image_state, const float* out_T, const float* dL_dcolor,
                                          const float* dL_dT, void* workspace, float* dgeos, float* drgbs,
                                          float* dsub, svr_stream_t stream)
{
    return batch_backward_impl(n_samp, V, views_host, views_dev, N, centers, lengths, geos, rgbs, geom_state,
                               capacity, binning_state, image_state, out_T, dL_dcolor, dL_dT, workspace, dgeos, drgbs,
                               dsub, true, stream);
}

int svr_batch_unpack_image_state(int64_t T, int64_t P, const void* image_state, int32_t* ranges, int32_t* tile_last,
                                 int32_t* n_contrib, svr_stream_t stream)
{
    SVR_REQUIRE(image_state && ranges && tile_last && n_contrib, SVR_ERR_INVALID,
                "svr_batch_unpack_image_state: null pointer");
    BatchImageState img = BatchImageState::carve((void*)image_state, T, P, nullptr);
    SVR_CHECK_HIP(hipMemcpyAsync(ranges, img.ranges, sizeof(uint32_t) * 2 * T, hipMemcpyDeviceToDevice, stream));
    SVR_CHECK_HIP(hipMemcpyAsync(tile_last, img.tile_last, sizeof(uint32_t) * T, hipMemcpyDeviceToDevice, stream));
    SVR_CHECK_HIP(hipMemcpyAsync(n_contrib, img.n_contrib, sizeof(uint32_t) * P, hipMemcpyDeviceToDevice, stream));
    return SVR_OK;
}

}  // extern "C"

// ---- per-voxel statistics over a batch of cameras ---------------------------
//
// SRC/sparse_voxel_gears/adaptive.py:190-221 (compute_training_stat) and
// NC/new_svraster_cuda/renderer.py:408-438 (mark_max_samp_rate, mark_near)
// loop over cameras on the host, one render or preprocess per camera.  Here
// one batch pass gives every (view, voxel) pair's max blending weight or
// tile count, and one thread per voxel folds the views in view order.

namespace svr {

// Camera-axis distance of a voxel centre: ((c - position) * lookat).sum(-1)
// with position = c2w[:3, 3] and lookat = c2w[:3, 2] (cameras.py:30-36).
__device__ inline float axis_dist(const svr_view& vw, float c0, float c1, float c2)
{
    return (c0 - vw.c2w[3]) * vw.c2w[2] + (c1 - vw.c2w[7]) * vw.c2w[6] + (c2 - vw.c2w[11]) * vw.c2w[10];
}

__global__ __launch_bounds__(256) void training_stat_kernel(int V, const svr_view* __restrict__ views, int64_t N,
                                                            const float* __restrict__ centers,
                                                            const float* __restrict__ pix_size,
                                                            const float* __restrict__ pair_max_w,
                                                            float* __restrict__ max_w,
                                                            float* __restrict__ min_samp_interval,
                                                            float* __restrict__ view_cnt)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const float c0 = centers[3 * v], c1 = centers[3 * v + 1], c2 = centers[3 * v + 2];
    float mw = 0.f, msi = 1e30f, cnt = 0.f;
    for (int u = 0; u < V; ++u)
    {
        const float w = pair_max_w[(size_t)u * N + v];
        mw = fmaxf(mw, w);
        if (w > 0.f)
        {
            const float si = axis_dist(views[u], c0, c1, c2) * pix_size[u];
            msi = fminf(msi, si);
            cnt = cnt + 1.f;
        }
    }
    max_w[v] = mw;
    min_samp_interval[v] = msi;
    view_cnt[v] = cnt;
}

__global__ __launch_bounds__(256) void mark_samp_kernel(int V, const svr_view* __restrict__ views, int64_t N,
                                                        const float* __restrict__ centers,
                                                        const float* __restrict__ lengths,
                                                        const float* __restrict__ pix_size,
                                                        const int32_t* __restrict__ pair_n_dup,
                                                        float* __restrict__ max_samp_rate, uint8_t* __restrict__ is_near)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const float c0 = centers[3 * v], c1 = centers[3 * v + 1], c2 = centers[3 * v + 2], len = lengths[v];
    float rate = 0.f;
    bool near_hit = false;
    for (int u = 0; u < V; ++u)
    {
        if (pair_n_dup[(size_t)u * N + v] <= 0) continue;
        const svr_view& vw = views[u];
        const float z = axis_dist(vw, c0, c1, c2);
        if (z > vw.near_plane) rate = fmaxf(rate, len / (z * pix_size[u]));
        near_hit = near_hit || (z <= vw.near_plane + len);
    }
    if (max_samp_rate) max_samp_rate[v] = rate;
    if (is_near) is_near[v] = near_hit ? 1 : 0;
}

}  // namespace svr

extern "C" {

int svr_batch_rasterize_max_w(int32_t n_samp, int32_t V, const svr_view* views_host, const svr_view* views_dev,
                              int64_t N, const float* centers, const float* lengths, const float* geos,
                              const void* geom_state, const void* order_table, int64_t R, void* binning_state,
                              void* image_state, float* out_T, float* pair_max_w, svr_stream_t stream)
{
    SVR_REQUIRE(pair_max_w != nullptr || N == 0, SVR_ERR_INVALID, "svr_batch_rasterize_max_w: null max_w");
    if (N == 0) return SVR_OK;
    return batch_forward_impl(n_samp, V, views_host, views_dev, N, centers, lengths, geos, nullptr, geom_state,
                              order_table, R, binning_state, image_state, nullptr, out_T, nullptr, stream,
                              pair_max_w);
}

int svr_batch_training_stat(int32_t V, const svr_view* views_dev, int64_t N, const float* centers,
                            const float* pix_size, const float* pair_max_w, float* max_w, float* min_samp_interval,
                            float* view_cnt, svr_stream_t stream)
{
    SVR_REQUIRE(V >= 1, SVR_ERR_INVALID, "number of views must be positive");
    if (N == 0) return SVR_OK;
    SVR_REQUIRE(views_dev && centers && pix_size && pair_max_w && max_w && min_samp_interval && view_cnt,
                SVR_ERR_INVALID, "svr_batch_training_stat: null pointer");
    training_stat_kernel<<<(unsigned)((N + 255) / 256), 256, 0, stream>>>(V, views_dev, N, centers, pix_size,
                                                                          pair_max_w, max_w, min_samp_interval,
                                                                          view_cnt);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

int svr_batch_mark_samp(int32_t V, const svr_view* views_dev, int64_t N, const float* centers, const float* lengths,
                        const float* pix_size, const int32_t* pair_n_dup, float* max_samp_rate, uint8_t* is_near,
                        svr_stream_t stream)
{
    SVR_REQUIRE(V >= 1, SVR_ERR_INVALID, "number of views must be positive");
    if (N == 0) return SVR_OK;
    SVR_REQUIRE(views_dev && centers && lengths && pix_size && pair_n_dup && (max_samp_rate || is_near),
                SVR_ERR_INVALID, "svr_batch_mark_samp: null pointer");
    mark_samp_kernel<<<(unsigned)((N + 255) / 256), 256, 0, stream>>>(V, views_dev, N, centers, lengths, pix_size,
                                                                      pair_n_dup, max_samp_rate, is_near);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

}  // extern "C"
