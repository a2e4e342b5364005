// Device-side building blocks of the per-tile compositor and its adjoint,
// shared by the single-view kernels (render.hip) and the multi-view
// persistent kernels (batch.hip).
//
// Reference: backends/new_cuda/src/forward.cu:43-394 (renderCUDA),
// backward.cu:37-530 (renderCUDA backward).
#pragma once

#include "svr_device.h"

namespace svr {

// Sync-slot threshold of the backward: an entry used by at least this many
// pixels of a wave is summed across the wave by a butterfly instead of
// per-lane LDS atomics (measured on the 1080p scene with 6-chunk windows:
// 8 beats 6, 10, 12 and 16; in the batch kernel 8 also beats 4 and 16 with
// windows of 2 and 4 chunks on 16 views).
#ifndef BWD_SYNC_MIN_MASKED
#define BWD_SYNC_MIN_MASKED 8
#endif

// Binning tile of TX x TY pixels; each wave of its workgroup owns one 8x8
// sub-tile (wave w at column w % WX, row w / WX).  The single-view path uses
// the reference's 16x16 tiles (forward.cu BLOCK_X / BLOCK_Y); the batch path
// can also bin at 8x8 or 32x8 (per-pixel results do not depend on it).
template <int TX_, int TY_>
struct TileShape
{
    static constexpr int TX = TX_, TY = TY_;
    static constexpr int WX = TX_ / 8, WY = TY_ / 8;
    static constexpr int NW = WX * WY;
    static constexpr int THREADS = 64 * NW;
    static_assert(TX_ % 8 == 0 && TY_ % 8 == 0 && NW >= 1 && NW <= 16, "tile must be whole 8x8 sub-tiles");
};
using Tile16 = TileShape<16, 16>;
constexpr int NWAVE = Tile16::NW;

// Staged payload of the current chunk, one 64-entry slot row per wave: the
// view-constant slab bounds and sample origin of the entry (entry_bounds),
// 1 / length, the 8 densities and the colour.
template <int NW>
struct StageT
{
    float lo[3][NW][64];
    float hi[3][NW][64];
    float cm[3][NW][64];
    float linv[NW][64];
    float g[8][NW][64];
    float rgb[3][NW][64];
    uint64_t diag[NW][32];  // each wave's copy of the diagonal half-plane masks (init_diag)
};
using Stage = StageT<NWAVE>;

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

// Diagonal half-plane masks of an 8x8 sub-tile (bit 8 * row + column):
// entry k + 1 (k = -1 .. 14) holds the pixels with column + row <= k, entry
// 24 + k (k = -8 .. 7) those with column - row <= k.  Lanes 0..31 of a wave
// write one entry each into the wave's table.
__device__ inline void init_diag(uint64_t* diag, int lane)
{
    if (lane < 32)
    {
        const int k = lane < 16 ? lane - 1 : lane - 24;
        uint64_t m = 0ull;
#pragma unroll
        for (int r = 0; r < 8; ++r)
        {
            const int t = lane < 16 ? k - r : k + r;  // last column kept in row r
            const uint32_t row = t < 0 ? 0u : (t >= 7 ? 0xFFu : (2u << t) - 1u);
            m |= (uint64_t)row << (8 * r);
        }
        diag[lane] = m;
    }
}

// Pixels of the sub-tile whose centre lies in the projected octagon o
// (bounds of x + y and of x - y in the pixel frame, the sub-tile's first
// pixel at (sx0, sy0)): lo <= S0 + col + row <= hi and likewise for x - y.
__device__ inline uint64_t oct_pixel_mask(const uint64_t* diag, float4 o, int sx0, int sy0)
{
    const float s0 = (float)(sx0 + sy0), d0 = (float)(sx0 - sy0);
    const auto cl = [](float x) { return fminf(fmaxf(x, -16.f), 32.f); };
    const int a0 = (int)ceilf(cl(o.x - s0)), a1 = (int)floorf(cl(o.y - s0));
    const int b0 = (int)ceilf(cl(o.z - d0)), b1 = (int)floorf(cl(o.w - d0));
    const uint64_t sm = diag[min(max(a1, -1), 14) + 1] & ~diag[min(max(a0 - 1, -1), 14) + 1];
    const uint64_t dm = diag[min(max(b1, -8), 7) + 24] & ~diag[min(max(b0 - 1, -8), 7) + 24];
    return sm & dm;
}

__device__ inline uint64_t select_q(const uint64_t* qpix, uint32_t q)
{
    uint64_t s = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) s = (q == i) ? qpix[i] : s;
    return s;
}

// A value that is the same in every lane, moved to scalar registers.
__device__ inline uint64_t uniform_u64(uint64_t x)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
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
    float lo[3], hi[3], cm[3], linv;
    float4 g0, g1;
    float r0, r1, r2;
};

// RGB = false: the colour is not read (weights-only passes) and staged as 0.
// OCT: the candidate mask is also cut to the projected octagon.  Only for
// rays with finite 1 / direction: where a direction component is exactly
// zero, the ray-box test of the reference yields NaN slab bounds for boxes
// with a face in the ray's plane and counts them as hits (forward.cu's
// !(a > b)), whether or not the ray meets the box, so those pixels keep the
// bbox candidates.
// ro: the camera centre (the view's ray origin).
template <bool RGB = true, bool OCT = true>
__device__ inline void load_entry(EntryRegs& e, bool valid, uint32_t val, const uint2* __restrict__ bboxes, const OctBox* __restrict__ octs,
                                  const float* __restrict__ centers, const float* __restrict__ lengths,
                                  const float* __restrict__ geos, const float* __restrict__ rgbs, int sx0, int sy0,
                                  int sx1, int sy1, const uint64_t* qpix, const uint64_t* diag, const float* ro)
{
    e.pm = 0ull;
    e.v = val & VOX_ID_MASK;
    if (!valid) return;
    if constexpr (OCT)
    {
        const OctBox& ob = octs[e.v];
        const float4 o = ob.o;
        e.pm = bbox_pixel_mask(ob.tb, sx0, sy0, sx1, sy1, select_q(qpix, val >> 29)) & oct_pixel_mask(diag, o, sx0, sy0);
    }
    else
    {
        e.pm = bbox_pixel_mask(bboxes[e.v], sx0, sy0, sx1, sy1, select_q(qpix, val >> 29));
    }
    if (!e.pm) return;
    const size_t v = e.v;
    const float l = lengths[v];
    entry_bounds(centers[3 * v + 0], centers[3 * v + 1], centers[3 * v + 2], l, ro, e.lo, e.hi, e.cm);
    e.linv = 1.f / l;
    e.g0 = *reinterpret_cast<const float4*>(geos + 8 * v);
    e.g1 = *reinterpret_cast<const float4*>(geos + 8 * v + 4);
    if constexpr (RGB)
    {
        e.r0 = rgbs[3 * v + 0];
        e.r1 = rgbs[3 * v + 1];
        e.r2 = rgbs[3 * v + 2];
    }
    else
    {
        e.r0 = e.r1 = e.r2 = 0.f;
    }
}

template <int NW>
__device__ inline void stage_entry(StageT<NW>& st, int w, int lane, const EntryRegs& e)
{
    if (!e.pm) return;
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        st.lo[k][w][lane] = e.lo[k];
        st.hi[k][w][lane] = e.hi[k];
        st.cm[k][w][lane] = e.cm[k];
    }
    st.linv[w][lane] = e.linv;
    st.g[0][w][lane] = e.g0.x; st.g[1][w][lane] = e.g0.y; st.g[2][w][lane] = e.g0.z; st.g[3][w][lane] = e.g0.w;
    st.g[4][w][lane] = e.g1.x; st.g[5][w][lane] = e.g1.y; st.g[6][w][lane] = e.g1.z; st.g[7][w][lane] = e.g1.w;
    st.rgb[0][w][lane] = e.r0; st.rgb[1][w][lane] = e.r1; st.rgb[2][w][lane] = e.r2;
}

// Front-to-back compositing of one wave's 8x8 sub-tile over the tile's list
// (forward.cu:150-328 for colour and T).  Returns the pixel's n_contrib and
// leaves its transmittance and colour in T, C0..C2.  EXACT selects the
// ray-box min/max form (see ray_aabb); it is uniform over the wave.
// MAX_W: weights-only pass (forward.cu:322-325 with track_max_w): colours are
// not read, the voxel id of each staged entry goes to sv, and every hit raises
// max_w_bits[voxel] to its blending weight (a float stored as its bits:
// weights are positive, so their order is that of the bits).
// MASK: after each chunk c the lane stores its hit mask (bit l: entry l of the
// chunk contributed) at mask_out[c * 64 * NW].
template <int NS, bool EXACT, int NW, bool MAX_W = false, bool MASK = false>
__device__ inline uint32_t composite_subtile(StageT<NW>& st, int w, int lane, const PixRay& ray, bool inside, uint2 range,
                                             int sx0, int sy0, int sx1, int sy1,
                                             const uint32_t* __restrict__ vox_list, const uint2* __restrict__ bboxes, const OctBox* __restrict__ octs,
                                             const float* __restrict__ centers, const float* __restrict__ lengths,
                                             const float* __restrict__ geos, const float* __restrict__ rgbs,
                                             float& T, float& C0, float& C1, float& C2,
                                             uint32_t (*sv)[64] = nullptr, uint32_t* __restrict__ max_w_bits = nullptr,
                                             uint64_t* __restrict__ mask_out = nullptr)
{
    uint32_t last = 0;
    bool done = !inside;
    init_diag(st.diag[w], lane);
    wave_sync();

    // The next chunk is loaded (into registers) while the current one is
    // composited; its pixel masks use the quadrant masks of the lanes active
    // at load time, and lanes that finish meanwhile drop their candidates.
    // pixels of each ray quadrant (fixed), intersected with the active ones
    uint64_t qall[8], qpix[8];
#pragma unroll
    for (uint32_t q = 0; q < 8; ++q) qall[q] = __ballot(ray.pq == q);
    {
        const uint64_t act = __ballot(!done);
#pragma unroll
        for (uint32_t q = 0; q < 8; ++q) qpix[q] = qall[q] & act;
    }
    EntryRegs cur;
    uint32_t base = range.x;
    {
        const uint32_t j = base + (uint32_t)lane;
        const bool valid = j < range.y;
        load_entry<!MAX_W, !EXACT>(cur, valid, valid ? vox_list[j] : 0u, bboxes, octs, centers, lengths, geos, rgbs, sx0, sy0, sx1,
                           sy1, qpix, st.diag[w], ray.ro);
    }
    while (base < range.y)
    {
        wave_sync();
        stage_entry(st, w, lane, cur);
        if constexpr (MAX_W)
        {
            if (cur.pm) sv[w][lane] = cur.v;
        }
        wave_sync();
        uint64_t hm = wave_transpose64(cur.pm, lane);
        if (done) hm = 0ull;

        // prefetch the next chunk
        const uint32_t nbase = base + 64;
        if (nbase < range.y)
        {
            const uint64_t act = __ballot(!done);
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) qpix[q] = qall[q] & act;
            const uint32_t j = nbase + (uint32_t)lane;
            const bool valid = j < range.y;
            load_entry<!MAX_W, !EXACT>(cur, valid, valid ? vox_list[j] : 0u, bboxes, octs, centers, lengths, geos, rgbs, sx0, sy0,
                               sx1, sy1, qpix, st.diag[w], ray.ro);
        }

        uint64_t hits = 0ull;
        while (hm)
        {
            // a lane whose candidate misses the box tests its next ones in the same step (the slab test
            // is cheap), so that fewer lanes sit out the density and exponential part: 15 % of the
            // candidates of the 1080p scene miss, and this took the forward from 0.387 to 0.377 ms
            int l;
            float a, b;
            do
            {
                l = __builtin_ctzll(hm);
                hm &= hm - 1ull;
                ray_aabb_pre<EXACT>(st.lo[0][w][l], st.lo[1][w][l], st.lo[2][w][l], st.hi[0][w][l],
                                    st.hi[1][w][l], st.hi[2][w][l], ray, a, b);
            } while ((a > b) && hm);
            if (a > b) continue;
            float g[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) g[i] = st.g[i][w][l];
            const float vol = vox_vol_cm<NS, false>(st.cm[0][w][l], st.cm[1][w][l], st.cm[2][w][l], st.linv[w][l], g,
                                                    a, b, ray, nullptr);
            const float alpha = fminf(MAX_ALPHA, 1.f - svr_expf_core(-vol));
            if (alpha < MIN_ALPHA) continue;
            if constexpr (MASK) hits |= 1ull << l;
            const float pt_w = T * alpha;
            if constexpr (MAX_W)
            {
                atomicMax(max_w_bits + sv[w][l], __float_as_uint(pt_w));
            }
            else
            {
                C0 = __builtin_fmaf(pt_w, st.rgb[0][w][l], C0);
                C1 = __builtin_fmaf(pt_w, st.rgb[1][w][l], C1);
                C2 = __builtin_fmaf(pt_w, st.rgb[2][w][l], C2);
            }
            T = T * (1.f - alpha);
            last = base + (uint32_t)l - range.x + 1u;
            if (T < EARLY_STOP_T)
            {
                done = true;
                break;
            }
        }
        if constexpr (MASK) mask_out[(size_t)((base - range.x) >> 6) * (64 * NW)] = hits;
        base = nbase;
        if (__ballot(!done) == 0ull) break;
    }
    return last;
}

// Staged payload of one entry for the backward, 6 float4 (one 16-byte LDS
// read each): slab bounds lo (3) and hi (3), sample origin cm (3),
// 1 / length, colour (3), densities 0-3, densities 4-7 (entry_bounds).
struct EntryPay
{
    float4 q[6];
};

__device__ inline void make_pay(EntryPay& p, const float* lo, const float* hi, const float* cm, float linv, float r0,
                                float r1, float r2, float4 g0, float4 g1)
{
    p.q[0] = make_float4(lo[0], lo[1], lo[2], hi[0]);
    p.q[1] = make_float4(hi[1], hi[2], cm[0], cm[1]);
    p.q[2] = make_float4(cm[2], linv, r0, r1);
    p.q[3] = make_float4(r2, 0.f, 0.f, 0.f);
    p.q[4] = g0;
    p.q[5] = g1;
}

// Adjoint of one (pixel, entry) hit (backward.cu:259-527 without the
// regularizers): recovers T before the entry, updates the running dL/dT and
// writes the 12 gradient terms (8 densities, 3 colours, subdivision priority).
// Branch-free: the terms are computed for every lane, and the pixel's state
// only changes where `active` holds and the entry contributes (the ray meets
// the box and alpha >= MIN_ALPHA; always true for the entries of the
// forward's hit masks, which came from the same arithmetic).  Returns whether
// the terms count.  A lane with nothing to do passes any staged slot and
// active = false; its terms are garbage and must not be used.
template <int NS, bool EXACT>
__device__ inline bool bwd_hit(const EntryPay& p, const PixRay& ray, float& T, float& last_dT, float dp0, float dp1,
                               float dp2, float* gq, bool active)
{
    const float4 q0 = p.q[0], q1 = p.q[1], q2 = p.q[2], g0 = p.q[4], g1 = p.q[5];
    const float r2 = p.q[3].x;
    float a, b;
    ray_aabb_pre<EXACT>(q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, ray, a, b);
    const float g[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
    // NS == 1: dI[i] = dd * w[i] is folded into the gradient scale below
    float dI[8], dd = 1.f;
    const float vol = (NS == 1) ? vox_vol1_w_cm(q1.z, q1.w, q2.x, q2.y, g, a, b, ray, dI, dd)
                                : vox_vol_cm<NS, true>(q1.z, q1.w, q2.x, q2.y, g, a, b, ray, dI);
    const float e = svr_expf_core(-vol);
    const float alpha = fminf(MAX_ALPHA, 1.f - e);
    const bool hit = active && !(a > b) && !(alpha < MIN_ALPHA);
    const float Tn = div_refined(T, 1.f - alpha);
    const float pt_w = alpha * Tn;
    const float dpt = __builtin_fmaf(dp2, r2, __builtin_fmaf(dp1, q2.w, dp0 * q2.z));
    const float diff = dpt - last_dT;
    const float dalpha = Tn * diff;
    const float ldn = __builtin_fmaf(alpha, diff, last_dT);
    const float dIs = (dalpha * e) * dd;
#pragma unroll
    for (int i = 0; i < 8; ++i) gq[i] = dIs * dI[i];
    gq[8] = pt_w * dp0;
    gq[9] = pt_w * dp1;
    gq[10] = pt_w * dp2;
    gq[11] = fabsf(dalpha * alpha);
    T = hit ? Tn : T;
    last_dT = hit ? ldn : last_dT;
    return hit;
}

// A pixel of the plain adjoint: its ray and running state; one call per hit,
// in descending entry order.
template <int NS, bool EXACT>
struct PlainHit
{
    PixRay ray;
    float T, last_dT, dp0, dp1, dp2;

    // backward.cu:160-180: state after the pixel's last contributor
    __device__ inline uint32_t init(bool inside, size_t pid, size_t HW, float bg, const float* __restrict__ out_T,
                                    const uint32_t* __restrict__ n_contrib, const float* __restrict__ dL_dcolor,
                                    const float* __restrict__ dL_dT)
    {
        T = 0.f; dp0 = 0.f; dp1 = 0.f; dp2 = 0.f; last_dT = 0.f;
        if (!inside) return 0u;
        T = out_T[pid];
        dp0 = dL_dcolor[pid];
        dp1 = dL_dcolor[HW + pid];
        dp2 = dL_dcolor[2 * HW + pid];
        last_dT = (dL_dT ? dL_dT[pid] : 0.f) + bg * (dp0 + dp1 + dp2);  // dL_dT null: no loss on T
        return n_contrib[pid];
    }
    __device__ inline bool operator()(const EntryPay& p, float* gq, bool active)
    {
        return bwd_hit<NS, EXACT>(p, ray, T, last_dT, dp0, dp1, dp2, gq, active);
    }
};

// Workgroup-shared state of the windowed adjoint: for the KW chunks of a
// window, the staged payload, the per-entry gradient accumulators, each
// wave's mask of the entries its pixels used, and the entries' emission-order
// rows.
template <int KW, int NW = 4>
struct WindowLDS
{
    EntryPay pay[KW][64];
    float acc[KW][12][64];
    uint64_t usedw[NW][KW];
    uint32_t orow[KW][64];
};

// Chunks of 64 list entries that hold a tile's contributors, from the
// forward's tile_last (range.x + the largest n_contrib of the tile's pixels),
// never more than the tile's list has.
__device__ inline uint32_t tile_chunks(uint2 range, uint32_t tile_last)
{
    const uint32_t n = min(tile_last - range.x, range.y - range.x);
    return (n + 63u) >> 6;
}
// Chunks per window of the single-view backward.  On the 1 M-voxel 1080p
// scene after 410 training steps (sync threshold 8) render_bwd took 0.697,
// 0.653, 0.615, 0.601, 0.728, 0.725 and 0.720 ms with windows of 2 to 8
// chunks: 5 chunks (44 KB of LDS, three workgroups per CU) cover most tiles'
// contributors in one window; 4 (four workgroups per CU) need a second window
// more often; from 6 on only two workgroups fit on a CU.  The batch kernels
// take their depth from the batch (autotuned).
#ifndef BWD_WINDOW_CHUNKS
#define BWD_WINDOW_CHUNKS 5
#endif
constexpr int BWD_WINDOW = BWD_WINDOW_CHUNKS;

// Reverse traversal of a tile (backward.cu:175-527 for the colour and T
// losses), driven by the forward's hit masks (composite_subtile's MASK): for
// chunk c a lane reads the 64-bit mask of entries that contributed to its
// pixel, so it walks only real hits (no bbox, quadrant, ray-box or alpha
// rejections, no n_contrib cut).
//
// The tile's waves go through its C chunks (tile_chunks) together, KW at a
// time (a window), from the tile's last contributor down.  The first window's
// hit masks and list values are requested before the pixels' state (init,
// which loads it and returns the pixel's n_contrib), so that both round trips
// overlap:
//   1. each wave reads its masks for the window's chunks and marks the
//      entries its pixels used;
//   2. the workgroup stages the payload of the used entries, one entry per
//      thread, once for all its waves;
//   3. each wave walks its pixels' hits, entries in descending order.  Entries
//      used by at least SYNC_MIN pixels of the wave are "sync slots": every
//      lane reaches them at the same step, at their place in its own order,
//      and their 12 terms are summed across the wave by one butterfly
//      (wave_sum12).  Between sync slots each lane walks its own hits in
//      lock-step.  Both add into the window's shared accumulators with LDS
//      atomics (the waves of the tile meet there instead of in global memory);
//   4. the workgroup writes the 12-float row of every entry of the window
//      into partial[orig] with plain stores (zeros for entries no pixel
//      used): an entry belongs to one tile, so this is its only writer.
// After the last window, the entries past the tile's last contributor get
// zero rows, so partial needs no clearing beforehand.
// HIT holds the pixel's ray and running state and computes one hit's 12
// terms (PlainHit here; the adjoint with the optional terms has its own).
template <int SYNC_MIN, typename TS, int KW, typename HIT, typename INIT>
__device__ inline void adjoint_window(WindowLDS<KW, TS::NW>& sh, int w, int lane, uint32_t C, INIT&& init, HIT& hit_fn,
                                      uint2 range, const uint64_t* __restrict__ masks,
                                      const uint32_t* __restrict__ vox_list, const float* __restrict__ centers,
                                      const float* __restrict__ lengths, const float* __restrict__ geos,
                                      const float* __restrict__ rgbs, const uint32_t* __restrict__ orig,
                                      float* __restrict__ partial)
{
    constexpr size_t STRIDE = 64 * TS::NW;
    constexpr int PER = (KW * 64 + TS::THREADS - 1) / TS::THREADS;  // window entries per thread
    const int t = threadIdx.x;

    // Software pipeline over windows: the hit masks of the next window and the
    // list values / emission rows of its entries are loaded while the current
    // window is staged and walked, so a window starts with its masks in
    // registers and its payload loads need one round trip.  Masks of chunks
    // past the wave's last contributor were never written (lim: the wave's
    // chunk count once known); they are loaded for the first window, whose
    // loads go out before the count, and dropped when the window starts.
    uint64_t m_pf[KW];
    uint32_t val_pf[PER], orow_pf[PER];
    auto prefetch = [&](int c0, uint32_t lim) {
        const int nj = min(KW, c0 + 1);
#pragma unroll
        for (int j = 0; j < KW; ++j)
            m_pf[j] = (j < nj && (uint32_t)(c0 - j) < lim) ? masks[(size_t)(c0 - j) * STRIDE] : 0ull;
#pragma unroll
        for (int k = 0; k < PER; ++k)
        {
            const int i = t + k * TS::THREADS;
            const int j = i >> 6, l = i & 63;
            const uint32_t e = range.x + ((uint32_t)(c0 - j) << 6) + (uint32_t)l;
            const bool ok = i < KW * 64 && j < nj && e < range.y;
            val_pf[k] = ok ? vox_list[e] : 0u;
            orow_pf[k] = ok ? orig[e] : 0u;
        }
    };
    if (C > 0) prefetch((int)C - 1, ~0u);
    const uint32_t nc = init();
    const uint32_t my_chunks = __builtin_amdgcn_readfirstlane((wave_max_u32(nc) + 63u) >> 6);
    // cleared before the first walk (two barriers away)
    for (int i = t; i < KW * 12 * 64; i += TS::THREADS) (&sh.acc[0][0][0])[i] = 0.f;

    for (int c0 = (int)C - 1; c0 >= 0; c0 -= KW)
    {
        const int nj = min(KW, c0 + 1);
        // 1. this wave's masks, the entries they use, the sync slots
        uint64_t m[KW], syncm[KW];
#pragma unroll
        for (int j = 0; j < KW; ++j)
        {
            const bool mine = j < nj && (uint32_t)(c0 - j) < my_chunks;
            m[j] = mine ? m_pf[j] : 0ull;
            syncm[j] = 0ull;
            uint64_t um = 0ull;
            if (mine)
            {
                const uint64_t er = wave_transpose64(m[j], lane);
                um = __ballot(er != 0ull);
                syncm[j] = __ballot(__popcll(er) >= SYNC_MIN);
            }
            if (lane == 0) sh.usedw[w][j] = um;
        }
        __syncthreads();
        // 2. stage the used entries' payload
#pragma unroll
        for (int k = 0; k < PER; ++k)
        {
            const int i = t + k * TS::THREADS;
            const int j = i >> 6, l = i & 63;
            const bool inwin = i < KW * 64 && j < nj;
            if (inwin && range.x + ((uint32_t)(c0 - j) << 6) + (uint32_t)l < range.y) sh.orow[j][l] = orow_pf[k];
            uint64_t um = 0ull;
            if (inwin)
            {
#pragma unroll
                for (int v = 0; v < TS::NW; ++v) um |= sh.usedw[v][j];
            }
            if ((um >> l) & 1ull)
            {
                const size_t v = val_pf[k] & VOX_ID_MASK;
                const float len = lengths[v];
                float lo[3], hi[3], cm[3];
                entry_bounds(centers[3 * v + 0], centers[3 * v + 1], centers[3 * v + 2], len, hit_fn.ray.ro, lo, hi, cm);
                make_pay(sh.pay[j][l], lo, hi, cm, 1.f / len, rgbs[3 * v + 0], rgbs[3 * v + 1], rgbs[3 * v + 2],
                         *reinterpret_cast<const float4*>(geos + 8 * v),
                         *reinterpret_cast<const float4*>(geos + 8 * v + 4));
            }
        }
        if (c0 - KW >= 0) prefetch(c0 - KW, my_chunks);
        __syncthreads();
        // 3. walk this wave's hits
#pragma unroll
        for (int j = 0; j < KW; ++j)
        {
            if (j >= nj) break;
            uint64_t hm = m[j], sync = uniform_u64(syncm[j]);
            float (&acc)[12][64] = sh.acc[j];
            // One hit per lane and step, each lane's hits in descending order.
            // A lane takes its highest remaining hit h unless h is a sync slot;
            // the wave's highest remaining sync slot s is taken in the step where
            // every lane that uses it has reached it (h == s), by all of them at
            // once (wave_sum12), while the other lanes go on with their own hits.
            // sync is the same in every lane and stays in scalar registers; the
            // hit is computed in every lane (a lane with nothing to take passes
            // slot 0 and active = false) so that no branch splits the step.
            while (__ballot(hm != 0ull) != 0ull)
            {
                const int h = hm ? 63 - __builtin_clzll(hm) : -1;
                const bool h_sync = h >= 0 && ((sync >> h) & 1ull);
                const int s = sync ? 63 - __builtin_clzll(sync) : -1;
                const bool owner = s >= 0 && ((hm >> s) & 1ull);
                const bool ready = s >= 0 && __ballot(owner && h != s) == 0ull;
                const bool is_sync = ready && owner;
                const int l = is_sync ? s : (h_sync ? -1 : h);
                const bool act = l >= 0;
                const int ls = act ? l : 0;
                hm &= act ? ~(1ull << ls) : ~0ull;
                float gq[12];
                // the payload's byte offset by a 24-bit multiply (ls < 64), a full-rate instruction
                const EntryPay& pay = *reinterpret_cast<const EntryPay*>(
                    reinterpret_cast<const char*>(&sh.pay[j][0]) + __umul24((unsigned)ls, (unsigned)sizeof(EntryPay)));
                const bool hit = hit_fn(pay, gq, act);
                if (ready)
                {
                    sync = uniform_u64(sync & ~(1ull << s));
                    if (__ballot(is_sync && hit) != 0ull)
                    {
                        float gs[12];
#pragma unroll
                        for (int i = 0; i < 12; ++i) gs[i] = (is_sync && hit) ? gq[i] : 0.f;
                        const float r = wave_sum12(gs, lane);
                        const int ri = wave_sum12_index(lane);
                        if ((lane & 3) == 0 && ri >= 0) atomicAdd(&acc[ri][s], r);
                    }
                }
                if (hit && !is_sync)
                {
#pragma unroll
                    for (int i = 0; i < 12; ++i) atomicAdd(&acc[i][l], gq[i]);
                }
            }
        }
        __syncthreads();
        // 4. one row per entry (accumulators of unused entries are zero), then reset the window
        for (int i = t; i < KW * 64; i += TS::THREADS)
        {
            const int j = i >> 6, l = i & 63;
            if (j < nj && range.x + ((uint32_t)(c0 - j) << 6) + (uint32_t)l < range.y)
            {
                float a[12];
#pragma unroll
                for (int k = 0; k < 12; ++k)
                {
                    a[k] = sh.acc[j][k][l];
                    sh.acc[j][k][l] = 0.f;
                }
                float4* row = reinterpret_cast<float4*>(partial + 12 * (size_t)sh.orow[j][l]);
                row[0] = make_float4(a[0], a[1], a[2], a[3]);
                row[1] = make_float4(a[4], a[5], a[6], a[7]);
                row[2] = make_float4(a[8], a[9], a[10], a[11]);
            }
        }
        __syncthreads();
    }
    // entries past every pixel's last contributor
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    for (uint32_t e = range.x + (C << 6) + (uint32_t)t; e < range.y; e += TS::THREADS)
    {
        float4* row = reinterpret_cast<float4*>(partial + 12 * (size_t)orig[e]);
        row[0] = z;
        row[1] = z;
        row[2] = z;
    }
}

}  // namespace svr
