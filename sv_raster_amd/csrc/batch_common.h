// Multi-view batch: device state layouts, tile geometry and the backward
// compositor kernels shared by batch.hip and the per-tile-shape backward
// instantiation files (batch_bwd_*.hip), which build in parallel.
//
// No reference counterpart as a whole: per (view, voxel) pair the arithmetic
// is the single-view path's (backward.cu:37-530 for the adjoint).
#pragma once

#include <algorithm>

#include "compositor.h"
#include "svr_host.h"
#include "voxel_ops.h"

namespace svr {

// Workgroups of the launches that loop over the deferred exact-ray tiles.
constexpr int64_t EXACT_GRID = 256;

// Backward depth of a batch (svr_view.bwd_window / bwd_sync; 0 selects these).
// Measured on the 16-view 2 M-voxel workload: 4 chunks and 8 pixels beat
// (6, 8), (5, 12) and (4, 12) with the then-default 8x8 binning tile.
constexpr int BATCH_BWD_WINDOW_DEFAULT = 4;
constexpr int BATCH_BWD_SYNC_DEFAULT = 8;

// Batch image state: per global tile ranges / tile_last, per pixel n_contrib,
// and the tile queue (order, deferred exact-ray tiles, counters).
struct BatchImageState
{
    uint2* ranges;        // [T]
    uint32_t* tile_last;  // [T]
    uint32_t* n_contrib;  // [P]
    uint64_t* qkeys;      // [T] queue sort keys
    uint64_t* qkeys_alt;  // [T]
    uint32_t* order;      // [T] tile order (longest list first)
    uint32_t* order_alt;  // [T]
    uint32_t* exact_list; // [T] tiles deferred to the exact ray-box launch
    uint32_t* qcounts;    // radix counts for the queue sort
    uint32_t* ctr;        // [8]: 0 fwd head, 1 exact count, 2 fwd exact head, 3 bwd head, 4 bwd exact head
    uint64_t* hit_masks;  // [hit_mask_slots(R, T)][up to 4 waves][64]: as in ImageState, global tile ids
    static BatchImageState carve(void* p, int64_t T, int64_t P, size_t* bytes, int64_t R = 0)
    {
        Carver c(p);
        BatchImageState s;
        s.ranges = c.take<uint2>(T);
        s.tile_last = c.take<uint32_t>(T);
        s.n_contrib = c.take<uint32_t>(P);
        s.qkeys = c.take<uint64_t>(T);
        s.qkeys_alt = c.take<uint64_t>(T);
        s.order = c.take<uint32_t>(T);
        s.order_alt = c.take<uint32_t>(T);
        s.exact_list = c.take<uint32_t>(T);
        s.qcounts = c.take<uint32_t>(sort_scratch_words(T, 1));
        s.ctr = c.take<uint32_t>(8);
        s.hit_masks = c.take<uint64_t>(hit_mask_slots(R, T) * 256);
        if (bytes) *bytes = c.size();
        return s;
    }
};

// Backward workspace: per-entry partial rows, then per-pair density and
// subdivision-priority gradients before the sum over views.
struct BatchBwdWork
{
    float* partial;   // [R][12]
    float* pair_geo;  // [V*N][8]
    float* pair_sub;  // [V*N]
    float* huge;      // [HUGE_SPLIT_ROWS][12] partial sums of split huge pairs
    static BatchBwdWork carve(void* p, int64_t VN, int64_t R, size_t* bytes)
    {
        Carver c(p);
        BatchBwdWork w;
        w.partial = c.take<float>(12 * R);
        w.pair_geo = c.take<float>(8 * VN);
        w.pair_sub = c.take<float>(VN);
        w.huge = c.take<float>(12 * HUGE_SPLIT_ROWS);
        if (bytes) *bytes = c.size();
        return w;
    }
};

__device__ inline int view_of_tile(const svr_view* __restrict__ views, int V, uint32_t g)
{
    int lo = 0, hi = V - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if ((uint32_t)views[mid].tile_base <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---- compositor over the tile queue ---------------------------------------

struct TileGeom
{
    int u, W, H, tx, ty, sx0, sy0, sx1, sy1, px, py;
    bool inside;
};

template <typename TS>
__device__ inline TileGeom tile_geom(const svr_view* __restrict__ views, int V, uint32_t g, int w, int lane)
{
    TileGeom t;
    t.u = view_of_tile(views, V, g);
    const svr_view& vw = views[t.u];
    t.W = vw.width;
    t.H = vw.height;
    const int gx = tiles_along(t.W, TS::TX);
    const uint32_t tile = g - (uint32_t)vw.tile_base;
    t.tx = (int)(tile % (uint32_t)gx);
    t.ty = (int)(tile / (uint32_t)gx);
    t.sx0 = t.tx * TS::TX + (w % TS::WX) * 8;
    t.sy0 = t.ty * TS::TY + (w / TS::WX) * 8;
    t.sx1 = min(t.sx0 + 7, t.W - 1);
    t.sy1 = min(t.sy0 + 7, t.H - 1);
    t.px = t.sx0 + (lane & 7);
    t.py = t.sy0 + (lane >> 3);
    t.inside = t.px < t.W && t.py < t.H;
    return t;
}

template <int NS, bool EXACT, typename TS, int KW, int SYNC>
__device__ inline void batch_bwd_tile(WindowLDS<KW, TS::NW>& sh, uint32_t g, int V, const svr_view* __restrict__ views,
                                      int64_t N, const uint2* __restrict__ ranges,
                                      const uint32_t* __restrict__ vox_list, const float* __restrict__ centers,
                                      const float* __restrict__ lengths, const float* __restrict__ geos,
                                      const float* __restrict__ rgbs, const float* __restrict__ out_T,
                                      const uint32_t* __restrict__ n_contrib, const float* __restrict__ dL_dcolor,
                                      const float* __restrict__ dL_dT, const uint32_t* __restrict__ orig,
                                      const uint64_t* __restrict__ hit_masks, const uint32_t* __restrict__ tile_last,
                                      float* __restrict__ partial)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const TileGeom t = tile_geom<TS>(views, V, g, w, lane);
    const svr_view& vw = views[t.u];
    PlainHit<NS, EXACT> hit;
    hit.ray = make_ray(t.px, t.py, t.W, t.H, vw.tanfovx, vw.tanfovy, vw.cx, vw.cy, vw.c2w);
    const bool need_exact = __syncthreads_or(!ray_inv_finite(hit.ray)) != 0;
    if (!EXACT && need_exact) return;  // rendered by the exact launch
    const size_t vo = (size_t)t.u * (size_t)N;
    const size_t pb = (size_t)vw.pix_base;
    const size_t HW = (size_t)t.H * t.W;
    const size_t pid = t.inside ? (size_t)t.py * t.W + t.px : 0;
    const uint2 range = ranges[g];
    const uint32_t C = tile_chunks(range, tile_last[g]);
    const uint64_t* masks = hit_masks + (((size_t)(range.x >> 6) + g) * TS::NW + w) * 64 + lane;
    adjoint_window<SYNC, TS, KW>(
        sh, w, lane, C,
        [&]() {
            return hit.init(t.inside, pid, HW, vw.bg_color, out_T + pb, n_contrib + pb, dL_dcolor + 3 * pb,
                            dL_dT ? dL_dT + pb : nullptr);
        },
        hit, range, masks, vox_list, centers, lengths, geos, rgbs + 3 * vo, orig, partial);
}

template <int NS, typename TS, int KW, int SYNC>
__global__ __launch_bounds__(TS::THREADS) void batch_bwd_kernel(
    int V, const svr_view* __restrict__ views, int64_t N, const uint32_t* __restrict__ order,
    const uint2* __restrict__ ranges, const uint32_t* __restrict__ vox_list, const float* __restrict__ centers,
    const float* __restrict__ lengths, const float* __restrict__ geos, const float* __restrict__ rgbs,
    const float* __restrict__ out_T, const uint32_t* __restrict__ n_contrib, const float* __restrict__ dL_dcolor,
    const float* __restrict__ dL_dT, const uint32_t* __restrict__ orig, const uint64_t* __restrict__ hit_masks,
    const uint32_t* __restrict__ tile_last, float* __restrict__ partial)
{
    __shared__ WindowLDS<KW, TS::NW> sh;
    const uint32_t g = order[blockIdx.x];
    batch_bwd_tile<NS, false, TS, KW, SYNC>(sh, g, V, views, N, ranges, vox_list, centers, lengths, geos, rgbs, out_T,
                                            n_contrib, dL_dcolor, dL_dT, orig, hit_masks, tile_last, partial);
}

// The forward's deferred tiles (exact_list, count ctr[1]; head ctr[4]).
template <int NS, typename TS, int KW, int SYNC>
__global__ __launch_bounds__(TS::THREADS) void batch_bwd_exact_kernel(
    int V, const svr_view* __restrict__ views, int64_t N, uint32_t* __restrict__ ctr,
    const uint32_t* __restrict__ exact_list, const uint2* __restrict__ ranges, const uint32_t* __restrict__ vox_list,
    const float* __restrict__ centers, const float* __restrict__ lengths, const float* __restrict__ geos,
    const float* __restrict__ rgbs, const float* __restrict__ out_T, const uint32_t* __restrict__ n_contrib,
    const float* __restrict__ dL_dcolor, const float* __restrict__ dL_dT, const uint32_t* __restrict__ orig,
    const uint64_t* __restrict__ hit_masks, const uint32_t* __restrict__ tile_last, float* __restrict__ partial)
{
    __shared__ WindowLDS<KW, TS::NW> sh;
    __shared__ uint32_t s_i;
    const uint32_t count = ctr[1];
    while (true)
    {
        if (threadIdx.x == 0) s_i = atomicAdd(&ctr[4], 1u);
        __syncthreads();
        const uint32_t i = __builtin_amdgcn_readfirstlane(s_i);
        if (i >= count) break;
        const uint32_t g = __builtin_amdgcn_readfirstlane(exact_list[i]);
        batch_bwd_tile<NS, true, TS, KW, SYNC>(sh, g, V, views, N, ranges, vox_list, centers, lengths, geos, rgbs,
                                               out_T, n_contrib, dL_dcolor, dL_dT, orig, hit_masks, tile_last, partial);
        __syncthreads();
    }
}

// Launches the batch backward compositor (fast tiles, then the deferred exact
// ones) with window depth kw (2, 4, 6, 8) and sync threshold sync (4, 8,
// 16); n_samp 2 and 3 take the defaults.  Defined for each tile shape in
// batch_bwd_<shape>.hip.
template <typename TS>
int batch_bwd_run(int n_samp, int kw, int sync, int V, const svr_view* views, int64_t N, int64_t T,
                  const BinningState& b, const BatchImageState& img, const float* centers, const float* lengths,
                  const float* geos, const float* rgbs, const float* out_T, const float* dL_dcolor, const float* dL_dT,
                  float* partial, hipStream_t stream);

// The body of batch_bwd_run for one tile shape (included by its file).
template <int NS, typename TS, int KW, int SYNC>
int batch_bwd_launch(int V, const svr_view* views, int64_t N, int64_t T, const BinningState& b,
                     const BatchImageState& img, const float* centers, const float* lengths, const float* geos,
                     const float* rgbs, const float* out_T, const float* dL_dcolor, const float* dL_dT,
                     float* partial, hipStream_t stream)
{
    batch_bwd_kernel<NS, TS, KW, SYNC><<<(unsigned)T, TS::THREADS, 0, stream>>>(
        V, views, N, img.order, img.ranges, b.vox_list, centers, lengths, geos, rgbs, out_T, img.n_contrib, dL_dcolor,
        dL_dT, b.orig, img.hit_masks, img.tile_last, partial);
    SVR_CHECK_LAUNCH();
    batch_bwd_exact_kernel<NS, TS, KW, SYNC><<<(unsigned)std::min<int64_t>(EXACT_GRID, T), TS::THREADS, 0, stream>>>(
        V, views, N, img.ctr, img.exact_list, img.ranges, b.vox_list, centers, lengths, geos, rgbs, out_T,
        img.n_contrib, dL_dcolor, dL_dT, b.orig, img.hit_masks, img.tile_last, partial);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

template <typename TS, int KW>
int batch_bwd_sync_dispatch(int sync, int V, const svr_view* views, int64_t N, int64_t T, const BinningState& b,
                            const BatchImageState& img, const float* centers, const float* lengths, const float* geos,
                            const float* rgbs, const float* out_T, const float* dL_dcolor, const float* dL_dT,
                            float* partial, hipStream_t stream)
{
    switch (sync)
    {
    case 4: return batch_bwd_launch<1, TS, KW, 4>(V, views, N, T, b, img, centers, lengths, geos, rgbs, out_T,
                                                  dL_dcolor, dL_dT, partial, stream);
    case 16: return batch_bwd_launch<1, TS, KW, 16>(V, views, N, T, b, img, centers, lengths, geos, rgbs, out_T,
                                                    dL_dcolor, dL_dT, partial, stream);
    default: return batch_bwd_launch<1, TS, KW, 8>(V, views, N, T, b, img, centers, lengths, geos, rgbs, out_T,
                                                   dL_dcolor, dL_dT, partial, stream);
    }
}

#define SVR_DEFINE_BATCH_BWD_RUN(TS_)                                                                                 \
    template <>                                                                                                       \
    int batch_bwd_run<TS_>(int n_samp, int kw, int sync, int V, const svr_view* views, int64_t N, int64_t T,         \
                           const BinningState& b, const BatchImageState& img, const float* centers,                  \
                           const float* lengths, const float* geos, const float* rgbs, const float* out_T,           \
                           const float* dL_dcolor, const float* dL_dT, float* partial, hipStream_t stream)           \
    {                                                                                                                 \
        constexpr int KD = BATCH_BWD_WINDOW_DEFAULT, SD = BATCH_BWD_SYNC_DEFAULT;                                     \
        if (n_samp == 2)                                                                                              \
            return batch_bwd_launch<2, TS_, KD, SD>(V, views, N, T, b, img, centers, lengths, geos, rgbs, out_T,       \
                                                    dL_dcolor, dL_dT, partial, stream);                               \
        if (n_samp == 3)                                                                                              \
            return batch_bwd_launch<3, TS_, KD, SD>(V, views, N, T, b, img, centers, lengths, geos, rgbs, out_T,       \
                                                    dL_dcolor, dL_dT, partial, stream);                               \
        switch (kw)                                                                                                   \
        {                                                                                                             \
        case 2: return batch_bwd_sync_dispatch<TS_, 2>(sync, V, views, N, T, b, img, centers, lengths, geos, rgbs,    \
                                                       out_T, dL_dcolor, dL_dT, partial, stream);                     \
        case 6: return batch_bwd_sync_dispatch<TS_, 6>(sync, V, views, N, T, b, img, centers, lengths, geos, rgbs,    \
                                                       out_T, dL_dcolor, dL_dT, partial, stream);                     \
        case 8: return batch_bwd_sync_dispatch<TS_, 8>(sync, V, views, N, T, b, img, centers, lengths, geos, rgbs,    \
                                                       out_T, dL_dcolor, dL_dT, partial, stream);                     \
        default: return batch_bwd_sync_dispatch<TS_, 4>(sync, V, views, N, T, b, img, centers, lengths, geos, rgbs,   \
                                                        out_T, dL_dcolor, dL_dT, partial, stream);                    \
        }                                                                                                             \
    }

}  // namespace svr
