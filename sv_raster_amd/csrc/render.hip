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
// arithmetic.  The backward accumulates the 12 per-entry gradients of a step
// in LDS (ds_add) and flushes one global atomic per (entry, component, wave).
#include "svr_device.h"
#include "svr_host.h"

namespace svr {

constexpr int NWAVE = 4;
#ifndef BWD_SYNC_MIN
#define BWD_SYNC_MIN 24
#endif

struct Stage
{
    float c[3][NWAVE][64];
    float l[NWAVE][64];
    float linv[NWAVE][64];
    float g[8][NWAVE][64];
    float rgb[3][NWAVE][64];
};

// Pixel mask (bit = 8 * row + column of the wave's 8x8 sub-tile) covered by a
// packed bbox (inclusive), intersected with qsel.
__device__ inline uint64_t bbox_pixel_mask(uint2 bb, int sx0, int sy0, int sx1, int sy1, uint64_t qsel)
{
    const int bxmin = (int)(bb.x >> 16), bymin = (int)(bb.x & 0xFFFFu);
    const int bxmax = (int)(bb.y >> 16), bymax = (int)(bb.y & 0xFFFFu);
    const int x0 = max(bxmin, sx0) - sx0, x1 = min(bxmax, sx1) - sx0;
    const int y0 = max(bymin, sy0) - sy0, y1 = min(bymax, sy1) - sy0;
    if (x0 > x1 || y0 > y1) return 0ull;
    const uint64_t col = (uint64_t)(((2u << x1) - 1u) & ~((1u << x0) - 1u));
    const uint64_t rows_hi = (y1 >= 7) ? ~0ull : ((1ull << (8 * (y1 + 1))) - 1ull);
    const uint64_t rows = rows_hi & ~((1ull << (8 * y0)) - 1ull);
    return (col * 0x0101010101010101ull) & rows & qsel;
}

__device__ inline uint64_t select_q(const uint64_t* qpix, uint32_t q)
{
    uint64_t s = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) s = (q == i) ? qpix[i] : s;
    return s;
}

// True when the packed bbox (inclusive) overlaps exactly one of the four 8x8
// sub-tiles of the 16x16 tile at (tx0, ty0): then only that sub-tile's wave can
// produce gradient terms for the entry.
__device__ inline bool single_subtile(uint2 bb, int tx0, int ty0, int W, int H)
{
    const int bxmin = (int)(bb.x >> 16), bymin = (int)(bb.x & 0xFFFFu);
    const int bxmax = (int)(bb.y >> 16), bymax = (int)(bb.y & 0xFFFFu);
    const int cl = (bxmin <= tx0 + 7 && bxmax >= tx0) ? 1 : 0;
    const int cr = (tx0 + 8 <= W - 1 && bxmax >= tx0 + 8 && bxmin <= tx0 + 15) ? 1 : 0;
    const int rt = (bymin <= ty0 + 7 && bymax >= ty0) ? 1 : 0;
    const int rb = (ty0 + 8 <= H - 1 && bymax >= ty0 + 8 && bymin <= ty0 + 15) ? 1 : 0;
    return (cl + cr) * (rt + rb) == 1;
}

__device__ inline uint32_t wave_max_u32(uint32_t x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x = max(x, (uint32_t)__shfl_xor(x, o, 64));
    return x;
}

// One list entry as loaded by a lane: voxel id, pixel mask on the wave's
// sub-tile and the render payload, kept in registers until staged in LDS.
struct EntryRegs
{
    uint32_t v;
    uint64_t pm;
    uint2 bb;
    float c0, c1, c2, l, linv;
    float4 g0, g1;
    float r0, r1, r2;
};

__device__ inline void load_entry(EntryRegs& e, bool valid, uint32_t val, const uint2* __restrict__ bboxes,
                                  const float* __restrict__ centers, const float* __restrict__ lengths,
                                  const float* __restrict__ geos, const float* __restrict__ rgbs, int sx0, int sy0,
                                  int sx1, int sy1, const uint64_t* qpix)
{
    e.pm = 0ull;
    e.v = val & VOX_ID_MASK;
    if (!valid) return;
    e.bb = bboxes[e.v];
    e.pm = bbox_pixel_mask(e.bb, sx0, sy0, sx1, sy1, select_q(qpix, val >> 29));
    if (!e.pm) return;
    const size_t v = e.v;
    e.c0 = centers[3 * v + 0];
    e.c1 = centers[3 * v + 1];
    e.c2 = centers[3 * v + 2];
    e.l = lengths[v];
    e.linv = 1.f / e.l;
    e.g0 = *reinterpret_cast<const float4*>(geos + 8 * v);
    e.g1 = *reinterpret_cast<const float4*>(geos + 8 * v + 4);
    e.r0 = rgbs[3 * v + 0];
    e.r1 = rgbs[3 * v + 1];
    e.r2 = rgbs[3 * v + 2];
}

__device__ inline void stage_entry(Stage& st, int w, int lane, const EntryRegs& e)
{
    if (!e.pm) return;
    st.c[0][w][lane] = e.c0; st.c[1][w][lane] = e.c1; st.c[2][w][lane] = e.c2; st.l[w][lane] = e.l;
    st.linv[w][lane] = e.linv;
    st.g[0][w][lane] = e.g0.x; st.g[1][w][lane] = e.g0.y; st.g[2][w][lane] = e.g0.z; st.g[3][w][lane] = e.g0.w;
    st.g[4][w][lane] = e.g1.x; st.g[5][w][lane] = e.g1.y; st.g[6][w][lane] = e.g1.z; st.g[7][w][lane] = e.g1.w;
    st.rgb[0][w][lane] = e.r0; st.rgb[1][w][lane] = e.r1; st.rgb[2][w][lane] = e.r2;
}

// Front-to-back compositing of one wave's 8x8 sub-tile over the tile's list
// (forward.cu:150-328 for colour and T).  Returns the pixel's n_contrib and
// leaves its transmittance and colour in T, C0..C2.  EXACT selects the
// ray-box min/max form (see ray_aabb); it is uniform over the wave.
template <int NS, bool EXACT>
__device__ inline uint32_t composite_subtile(Stage& st, int w, int lane, const PixRay& ray, bool inside, uint2 range,
                                             int sx0, int sy0, int sx1, int sy1,
                                             const uint32_t* __restrict__ vox_list, const uint2* __restrict__ bboxes,
                                             const float* __restrict__ centers, const float* __restrict__ lengths,
                                             const float* __restrict__ geos, const float* __restrict__ rgbs,
                                             float& T, float& C0, float& C1, float& C2)
{
    uint32_t last = 0;
    bool done = !inside;

    // The next chunk is loaded (into registers) while the current one is
    // composited; its pixel masks use the quadrant masks of the lanes active
    // at load time, and lanes that finish meanwhile drop their candidates.
    uint64_t qpix[8];
#pragma unroll
    for (uint32_t q = 0; q < 8; ++q) qpix[q] = __ballot(!done && ray.pq == q);
    EntryRegs cur;
    uint32_t base = range.x;
    {
        const uint32_t j = base + (uint32_t)lane;
        const bool valid = j < range.y;
        load_entry(cur, valid, valid ? vox_list[j] : 0u, bboxes, centers, lengths, geos, rgbs, sx0, sy0, sx1, sy1,
                   qpix);
    }
    while (base < range.y)
    {
        wave_sync();
        stage_entry(st, w, lane, cur);
        wave_sync();
        uint64_t hm = wave_transpose64(cur.pm, lane);
        if (done) hm = 0ull;

        // prefetch the next chunk
        const uint32_t nbase = base + 64;
        if (nbase < range.y)
        {
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) qpix[q] = __ballot(!done && ray.pq == q);
            const uint32_t j = nbase + (uint32_t)lane;
            const bool valid = j < range.y;
            load_entry(cur, valid, valid ? vox_list[j] : 0u, bboxes, centers, lengths, geos, rgbs, sx0, sy0, sx1,
                       sy1, qpix);
        }

        while (hm)
        {
            const int l = __builtin_ctzll(hm);
            hm &= hm - 1ull;
            const float c0 = st.c[0][w][l], c1 = st.c[1][w][l], c2 = st.c[2][w][l], len = st.l[w][l];
            float a, b;
            ray_aabb<EXACT>(c0, c1, c2, len, ray, a, b);
            if (a > b) continue;
            float g[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) g[i] = st.g[i][w][l];
            const float vol = vox_vol<NS, false>(c0, c1, c2, len, st.linv[w][l], g, a, b, ray, nullptr);
            const float alpha = fminf(MAX_ALPHA, 1.f - svr_expf(-vol));
            if (alpha < MIN_ALPHA) continue;
            const float pt_w = T * alpha;
            C0 = __builtin_fmaf(pt_w, st.rgb[0][w][l], C0);
            C1 = __builtin_fmaf(pt_w, st.rgb[1][w][l], C1);
            C2 = __builtin_fmaf(pt_w, st.rgb[2][w][l], C2);
            T = T * (1.f - alpha);
            last = base + (uint32_t)l - range.x + 1u;
            if (T < EARLY_STOP_T)
            {
                done = true;
                break;
            }
        }
        base = nbase;
        if (__ballot(!done) == 0ull) break;
    }
    return last;
}

// Launched twice per frame: EXACT = false handles the tiles whose 256 rays
// all have finite 1/rd (hardware min/max in the ray-box test), EXACT = true
// the remaining tiles (the reference's ternary min/max, whose NaN handling
// differs); each launch returns at once from the other's tiles.
template <int NS, bool EXACT>
__global__ __launch_bounds__(256) void render_fwd_kernel(
    int W, int H, int gx, int T_tiles, int per_xcd, float tanfovx, float tanfovy, float cx, float cy,
    const float* __restrict__ c2w, float bg, const uint2* __restrict__ ranges, const uint32_t* __restrict__ vox_list,
    const uint2* __restrict__ bboxes, const float* __restrict__ centers, const float* __restrict__ lengths,
    const float* __restrict__ geos, const float* __restrict__ rgbs, uint32_t* __restrict__ tile_last,
    uint32_t* __restrict__ n_contrib, float* __restrict__ out_color, float* __restrict__ out_T)
{
    __shared__ Stage st;
    __shared__ uint32_t s_last[NWAVE];
    const int tile = xcd_tile(blockIdx.x, per_xcd);
    if (tile >= T_tiles) return;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tx = tile % gx, ty = tile / gx;
    const int sx0 = tx * TILE_X + (w & 1) * 8, sy0 = ty * TILE_Y + (w >> 1) * 8;
    const int sx1 = min(sx0 + 7, W - 1), sy1 = min(sy0 + 7, H - 1);
    const int px = sx0 + (lane & 7), py = sy0 + (lane >> 3);
    const bool inside = px < W && py < H;
    const PixRay ray = make_ray(px, py, W, H, tanfovx, tanfovy, cx, cy, c2w);
    if ((__syncthreads_or(!ray_inv_finite(ray)) != 0) != EXACT) return;
    const uint2 range = ranges[tile];

    float T = 1.f, C0 = 0.f, C1 = 0.f, C2 = 0.f;
    const uint32_t last = composite_subtile<NS, EXACT>(st, w, lane, ray, inside, range, sx0, sy0, sx1, sy1, vox_list,
                                                       bboxes, centers, lengths, geos, rgbs, T, C0, C1, C2);

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

// Adjoint of one (pixel, entry) hit (backward.cu:259-527 without the
// regularizers): recovers T before the entry, updates the running dL/dT and
// writes the 12 gradient terms (8 densities, 3 colours, subdivision priority).
template <int NS, bool EXACT>
__device__ inline bool bwd_hit(const Stage& st, int w, int l, const PixRay& ray, float& T, float& last_dT, float dp0,
                               float dp1, float dp2, float* gq)
{
    const float c0 = st.c[0][w][l], c1 = st.c[1][w][l], c2 = st.c[2][w][l], len = st.l[w][l];
    float a, b;
    ray_aabb<EXACT>(c0, c1, c2, len, ray, a, b);
    if (a > b) return false;
    float g[8], dI[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) g[i] = st.g[i][w][l];
    const float vol = vox_vol<NS, true>(c0, c1, c2, len, st.linv[w][l], g, a, b, ray, dI);
    const float e = svr_expf(-vol);
    const float alpha = fminf(MAX_ALPHA, 1.f - e);
    if (alpha < MIN_ALPHA) return false;
    T = T / (1.f - alpha);
    const float pt_w = alpha * T;
    const float r0 = st.rgb[0][w][l], r1 = st.rgb[1][w][l], r2 = st.rgb[2][w][l];
    const float dpt = __builtin_fmaf(dp2, r2, __builtin_fmaf(dp1, r1, dp0 * r0));
    const float diff = dpt - last_dT;
    const float dalpha = T * diff;
    last_dT = __builtin_fmaf(alpha, diff, last_dT);
    const float dIs = dalpha * e;
#pragma unroll
    for (int i = 0; i < 8; ++i) gq[i] = dIs * dI[i];
    gq[8] = pt_w * dp0;
    gq[9] = pt_w * dp1;
    gq[10] = pt_w * dp2;
    gq[11] = fabsf(dalpha * alpha);
    return true;
}

// Reverse traversal of one wave's 8x8 sub-tile (backward.cu:175-527 for the
// colour and T losses).  EXACT selects the ray-box min/max form (see
// ray_aabb); it is uniform over the wave.
//
// Entries whose bbox covers at least SYNC_MIN pixels of the wave (coarse
// voxels, the outside shell) are "sync slots": every pixel processes them at
// the same step, at their place in its own order, and their gradient is summed
// across the wave with DPP instead of up to 64-way conflicting LDS atomics.
// Between two sync slots each lane walks its own remaining hits in lock-step
// and accumulates with ds_add.
//
// The chunk loop is software-pipelined: chunk c+1 is loaded into registers
// before chunk c's gradients are flushed with global atomics, because on this
// architecture a later load's wait also waits for all earlier stores/atomics.
template <int NS, int SYNC_MIN, bool EXACT>
__device__ inline void adjoint_subtile(Stage& st, float (&acc)[12][NWAVE][64], int w, int lane, int W, int H, int tx,
                                       int ty, int sx0, int sy0, int sx1, int sy1, bool inside, size_t pid, size_t HW,
                                       const PixRay& ray, uint2 range, float bg,
                                       const uint32_t* __restrict__ vox_list, const uint2* __restrict__ bboxes,
                                       const float* __restrict__ centers, const float* __restrict__ lengths,
                                       const float* __restrict__ geos, const float* __restrict__ rgbs,
                                       const float* __restrict__ out_T, const uint32_t* __restrict__ n_contrib,
                                       const float* __restrict__ dL_dcolor, const float* __restrict__ dL_dT,
                                       const uint32_t* __restrict__ orig, float* __restrict__ partial)
{
    float T = 0.f, dp0 = 0.f, dp1 = 0.f, dp2 = 0.f, last_dT = 0.f;
    uint32_t nc = 0;
    if (inside)
    {
        T = out_T[pid];
        nc = n_contrib[pid];
        dp0 = dL_dcolor[pid];
        dp1 = dL_dcolor[HW + pid];
        dp2 = dL_dcolor[2 * HW + pid];
        last_dT = dL_dT[pid] + bg * (dp0 + dp1 + dp2);
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k][w][lane] = 0.f;
    uint64_t qpix[8];
#pragma unroll
    for (uint32_t q = 0; q < 8; ++q) qpix[q] = __ballot(inside && nc > 0 && ray.pq == q);
    const int64_t rx = (int64_t)range.x;
    int64_t end = rx + wave_max_u32(nc);

    EntryRegs cur;
    uint32_t cur_o = 0;
    {
        const int64_t jl = end - 1 - lane;
        const bool valid = jl >= rx;
        load_entry(cur, valid, valid ? vox_list[jl] : 0u, bboxes, centers, lengths, geos, rgbs, sx0, sy0, sx1, sy1,
                   qpix);
        if (valid) cur_o = orig[jl];
    }
    while (end > rx)
    {
        wave_sync();
        stage_entry(st, w, lane, cur);
        uint64_t sync = __ballot(__popcll(cur.pm) >= SYNC_MIN);
        wave_sync();
        uint64_t hm = wave_transpose64(cur.pm, lane);
        // entry end-1-l is behind this pixel's last contributor when
        // end-1-l >= range.x + nc, i.e. l < end - range.x - nc
        const int64_t cut = end - rx - (int64_t)nc;
        if (cut >= 64) hm = 0ull;
        else if (cut > 0) hm &= ~((1ull << cut) - 1ull);

        while (true)
        {
            const int s = sync ? __builtin_ctzll(sync) : 64;
            uint64_t seg = (s == 64) ? hm : (hm & ((1ull << s) - 1ull));
            hm &= ~seg;
            while (seg)
            {
                const int l = __builtin_ctzll(seg);
                seg &= seg - 1ull;
                float gq[12];
                if (bwd_hit<NS, EXACT>(st, w, l, ray, T, last_dT, dp0, dp1, dp2, gq))
                {
#pragma unroll
                    for (int i = 0; i < 12; ++i) atomicAdd(&acc[i][w][l], gq[i]);
                }
            }
            if (s == 64) break;
            sync &= sync - 1ull;
            const bool mine = (hm >> s) & 1ull;
            hm &= ~(1ull << s);
            float gq[12];
#pragma unroll
            for (int i = 0; i < 12; ++i) gq[i] = 0.f;
            bool hit = false;
            if (mine) hit = bwd_hit<NS, EXACT>(st, w, s, ray, T, last_dT, dp0, dp1, dp2, gq);
            if (__ballot(hit) == 0ull) continue;
            if (!hit)
            {
#pragma unroll
                for (int i = 0; i < 12; ++i) gq[i] = 0.f;
            }
#pragma unroll
            for (int i = 0; i < 12; ++i)
            {
                const float r = wave_sum_f32(gq[i]);
                if (lane == 0) acc[i][w][s] += r;
            }
        }

        // prefetch the next chunk before this chunk's atomics are issued
        const uint32_t fo = cur_o;
        const bool fpm = cur.pm != 0ull;
        const bool fsole = fpm && single_subtile(cur.bb, tx * TILE_X, ty * TILE_Y, W, H);
        const int64_t nend = end - 64;
        {
            const int64_t jl = nend - 1 - lane;
            const bool valid = jl >= rx;
            load_entry(cur, valid, valid ? vox_list[jl] : 0u, bboxes, centers, lengths, geos, rgbs, sx0, sy0, sx1,
                       sy1, qpix);
            if (valid) cur_o = orig[jl];
        }
        wave_sync();
        // per-entry partial sums: an entry belongs to one tile, so at most the
        // four waves of this workgroup add into its row; when the entry's bbox
        // lies in this wave's sub-tile alone, the wave owns the (zeroed) row
        // and writes it with plain stores
        if (fsole)
        {
            float a[12];
            bool any = false;
#pragma unroll
            for (int i = 0; i < 12; ++i)
            {
                a[i] = acc[i][w][lane];
                any = any || (a[i] != 0.f);
                acc[i][w][lane] = 0.f;
            }
            if (any)
            {
                float4* row = reinterpret_cast<float4*>(partial + 12 * (size_t)fo);
                row[0] = make_float4(a[0], a[1], a[2], a[3]);
                row[1] = make_float4(a[4], a[5], a[6], a[7]);
                row[2] = make_float4(a[8], a[9], a[10], a[11]);
            }
        }
        else if (fpm)
        {
            float* row = partial + 12 * (size_t)fo;
#pragma unroll
            for (int i = 0; i < 12; ++i)
            {
                const float a = acc[i][w][lane];
                if (a != 0.f) unsafeAtomicAdd(row + i, a);
                acc[i][w][lane] = 0.f;
            }
        }
        end = nend;
    }
}

// Launched twice per frame, split by tiles like render_fwd_kernel.
template <int NS, int SYNC_MIN, bool EXACT>
__global__ __launch_bounds__(256) void render_bwd_kernel(
    int W, int H, int gx, int T_tiles, int per_xcd, float tanfovx, float tanfovy, float cx, float cy,
    const float* __restrict__ c2w, float bg, const uint2* __restrict__ ranges, const uint32_t* __restrict__ vox_list,
    const uint2* __restrict__ bboxes, const float* __restrict__ centers, const float* __restrict__ lengths,
    const float* __restrict__ geos, const float* __restrict__ rgbs, const float* __restrict__ out_T,
    const uint32_t* __restrict__ n_contrib, const float* __restrict__ dL_dcolor, const float* __restrict__ dL_dT,
    const uint32_t* __restrict__ orig, float* __restrict__ partial)
{
    __shared__ Stage st;
    __shared__ float acc[12][NWAVE][64];
    const int tile = xcd_tile(blockIdx.x, per_xcd);
    if (tile >= T_tiles) return;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tx = tile % gx, ty = tile / gx;
    const int sx0 = tx * TILE_X + (w & 1) * 8, sy0 = ty * TILE_Y + (w >> 1) * 8;
    const int sx1 = min(sx0 + 7, W - 1), sy1 = min(sy0 + 7, H - 1);
    const int px = sx0 + (lane & 7), py = sy0 + (lane >> 3);
    const bool inside = px < W && py < H;
    const PixRay ray = make_ray(px, py, W, H, tanfovx, tanfovy, cx, cy, c2w);
    const uint2 range = ranges[tile];
    const size_t HW = (size_t)H * W;
    const size_t pid = inside ? (size_t)py * W + px : 0;
    if ((__syncthreads_or(!ray_inv_finite(ray)) != 0) != EXACT) return;
    adjoint_subtile<NS, SYNC_MIN, EXACT>(st, acc, w, lane, W, H, tx, ty, sx0, sy0, sx1, sy1, inside, pid, HW, ray,
                                         range, bg, vox_list, bboxes, centers, lengths, geos, rgbs, out_T, n_contrib,
                                         dL_dcolor, dL_dT, orig, partial);
}

// Per-voxel gradient = sum of the partial rows of its entries, which are
// contiguous in emission order: [scan[v] - n_dup[v], scan[v]).  Voxels with
// more than BIG_DUP entries (listed in big_list by the duplication pass) are
// summed by a whole workgroup.

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

__global__ __launch_bounds__(256) void grad_reduce_kernel(int64_t N, const uint32_t* __restrict__ n_dup,
                                                          const uint32_t* __restrict__ scan,
                                                          const float* __restrict__ partial,
                                                          float* __restrict__ dgeos, float* __restrict__ drgbs,
                                                          float* __restrict__ dsub)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const uint32_t nd = n_dup[v];
    if (nd > BIG_DUP) return;
    float a[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) a[i] = 0.f;
    const uint32_t off = scan[v] - nd;
    for (uint32_t e = 0; e < nd; ++e)
    {
        const float4* r = reinterpret_cast<const float4*>(partial + 12 * (size_t)(off + e));
        const float4 p0 = r[0], p1 = r[1], p2 = r[2];
        a[0] += p0.x; a[1] += p0.y; a[2] += p0.z; a[3] += p0.w;
        a[4] += p1.x; a[5] += p1.y; a[6] += p1.z; a[7] += p1.w;
        a[8] += p2.x; a[9] += p2.y; a[10] += p2.z; a[11] += p2.w;
    }
    write_grad((size_t)v, a, dgeos, drgbs, dsub);
}

__global__ __launch_bounds__(256) void grad_reduce_big_kernel(const uint32_t* __restrict__ big_list,
                                                              const uint32_t* __restrict__ big_count,
                                                              const uint32_t* __restrict__ n_dup,
                                                              const uint32_t* __restrict__ scan,
                                                              const float* __restrict__ partial,
                                                              float* __restrict__ dgeos, float* __restrict__ drgbs,
                                                              float* __restrict__ dsub)
{
    __shared__ float red[12][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // one workgroup per voxel with many entries; blocks stride over the queue
    const uint32_t count = *big_count;
    for (uint32_t i = blockIdx.x; i < count; i += gridDim.x)
    {
        const uint32_t v = big_list[i];
        const uint32_t nd = n_dup[v];
        const uint32_t off = scan[v] - nd;
        float a[12];
#pragma unroll
        for (int i = 0; i < 12; ++i) a[i] = 0.f;
        for (uint32_t e = threadIdx.x; e < nd; e += blockDim.x)
        {
            const float4* r = reinterpret_cast<const float4*>(partial + 12 * (size_t)(off + e));
            const float4 p0 = r[0], p1 = r[1], p2 = r[2];
            a[0] += p0.x; a[1] += p0.y; a[2] += p0.z; a[3] += p0.w;
            a[4] += p1.x; a[5] += p1.y; a[6] += p1.z; a[7] += p1.w;
            a[8] += p2.x; a[9] += p2.y; a[10] += p2.z; a[11] += p2.w;
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
            write_grad((size_t)v, t, dgeos, drgbs, dsub);
        }
        __syncthreads();
    }
}

template <int NS>
int launch_fwd(int W, int H, float tanfovx, float tanfovy, float cx, float cy, const float* c2w, float bg,
               const GeomState& g, const BinningState& b, const ImageState& img, const float* centers,
               const float* lengths, const float* geos, const float* rgbs, float* out_color, float* out_T,
               hipStream_t stream)
{
    const int gx = tiles_x(W), gy = tiles_y(H);
    const int T = gx * gy;
    const int per = (T + 7) / 8;
    prof_begin(PROF_RENDER_FWD, stream);
    render_fwd_kernel<NS, false><<<(unsigned)(8 * per), 256, 0, stream>>>(
        W, H, gx, T, per, tanfovx, tanfovy, cx, cy, c2w, bg, img.ranges, b.vox_list, g.bbox, centers, lengths, geos,
        rgbs, img.tile_last, img.n_contrib, out_color, out_T);
    SVR_CHECK_LAUNCH();
    render_fwd_kernel<NS, true><<<(unsigned)(8 * per), 256, 0, stream>>>(
        W, H, gx, T, per, tanfovx, tanfovy, cx, cy, c2w, bg, img.ranges, b.vox_list, g.bbox, centers, lengths, geos,
        rgbs, img.tile_last, img.n_contrib, out_color, out_T);
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
    const int per = (T + 7) / 8;
    if (fill_u32(partial, 0u, 12 * R, stream) != SVR_OK) return SVR_ERR_HIP;
    prof_begin(PROF_RENDER_BWD, stream);
    render_bwd_kernel<NS, BWD_SYNC_MIN, false><<<(unsigned)(8 * per), 256, 0, stream>>>(
        W, H, gx, T, per, tanfovx, tanfovy, cx, cy, c2w, bg, img.ranges, b.vox_list, g.bbox, centers, lengths, geos,
        rgbs, out_T, img.n_contrib, dL_dcolor, dL_dT, b.orig, partial);
    SVR_CHECK_LAUNCH();
    render_bwd_kernel<NS, BWD_SYNC_MIN, true><<<(unsigned)(8 * per), 256, 0, stream>>>(
        W, H, gx, T, per, tanfovx, tanfovy, cx, cy, c2w, bg, img.ranges, b.vox_list, g.bbox, centers, lengths, geos,
        rgbs, out_T, img.n_contrib, dL_dcolor, dL_dT, b.orig, partial);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_RENDER_BWD, stream);
    prof_begin(PROF_GRAD_REDUCE, stream);
    grad_reduce_kernel<<<(unsigned)((N + 255) / 256), 256, 0, stream>>>(N, g.n_dup, g.scan, partial, dgeos, drgbs,
                                                                       dsub);
    SVR_CHECK_LAUNCH();
    grad_reduce_big_kernel<<<1024, 256, 0, stream>>>(g.big_list, g.counters + 2, g.n_dup, g.scan, partial, dgeos,
                                                     drgbs, dsub);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_GRAD_REDUCE, stream);
    return SVR_OK;
}

}  // namespace svr

using namespace svr;

extern "C" {

int svr_rasterize_forward(int32_t n_samp, int32_t W, int32_t H, float tanfovx, float tanfovy, float cx, float cy,
                          const float* c2w, float bg, int64_t N, const float* centers, const float* lengths,
                          const float* geos, const float* rgbs, const void* geom_state, const void* order_table,
                          int64_t R, void* binning_state, void* image_state, float* out_color, float* out_T,
                          svr_stream_t stream)
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
    GeomState g = GeomState::carve((void*)geom_state, N, nullptr);
    OrderTable ot = OrderTable::carve((void*)order_table, N, nullptr);
    BinningState b = BinningState::carve(binning_state, R, nullptr);
    ImageState img = ImageState::carve(image_state, (int64_t)W * H, T, nullptr);
    int rc = binning_forward(W, H, N, g, ot, R, b, img, stream);
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

size_t svr_backward_workspace_size(int64_t R)
{
    return align_up(sizeof(float) * 12 * (size_t)(R > 0 ? R : 0), 256);
}

int svr_rasterize_backward(int32_t n_samp, int32_t W, int32_t H, float tanfovx, float tanfovy, float cx, float cy,
                           const float* c2w, float bg, int64_t N, const float* centers, const float* lengths,
                           const float* geos, const float* rgbs, const void* geom_state, int64_t R,
                           const void* binning_state, const void* image_state, const float* out_T,
                           const float* dL_dcolor, const float* dL_dT, void* workspace, float* dgeos, float* drgbs,
                           float* dsub, svr_stream_t stream)
{
    SVR_REQUIRE(n_samp >= 1 && n_samp <= MAX_N_SAMP, SVR_ERR_INVALID, "n_samp_per_vox should be in range [1, %d].",
                MAX_N_SAMP);
    if (N == 0) return SVR_OK;
    SVR_REQUIRE(geom_state && image_state && out_T && dL_dcolor && dL_dT && dgeos && drgbs && dsub && c2w,
                SVR_ERR_INVALID, "svr_rasterize_backward: null pointer");
    SVR_REQUIRE(((uintptr_t)geos & 15u) == 0 && ((uintptr_t)dgeos & 15u) == 0, SVR_ERR_INVALID,
                "geos and dL_dgeos must be 16-byte aligned");
    if (R == 0)
    {
        if (fill_u32(dgeos, 0u, 8 * N, stream) != SVR_OK) return SVR_ERR_HIP;
        if (fill_u32(drgbs, 0u, 3 * N, stream) != SVR_OK) return SVR_ERR_HIP;
        if (fill_u32(dsub, 0u, N, stream) != SVR_OK) return SVR_ERR_HIP;
        return SVR_OK;
    }
    SVR_REQUIRE(binning_state && workspace, SVR_ERR_INVALID, "svr_rasterize_backward: null binning state or workspace");
    SVR_REQUIRE(((uintptr_t)workspace & 15u) == 0, SVR_ERR_INVALID, "backward workspace must be 16-byte aligned");
    const int64_t T = (int64_t)tiles_x(W) * tiles_y(H);
    GeomState g = GeomState::carve((void*)geom_state, N, nullptr);
    BinningState b = BinningState::carve((void*)binning_state, R, nullptr);
    ImageState img = ImageState::carve((void*)image_state, (int64_t)W * H, T, nullptr);
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
