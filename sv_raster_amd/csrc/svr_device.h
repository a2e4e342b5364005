// Device-side constants and math shared by the rasterizer kernels (gfx950).
//
// Floating-point contract: every file is compiled with -ffp-contract=off and
// IEEE-correct fp32 division/sqrt, so the only fused operations are the
// explicit fmaf() calls written below and in the kernels.  exp() is
// svr_expf(), a Cody-Waite reduction + degree-7 Taylor polynomial built only
// from correctly rounded operations, so its result is the same on any IEEE
// machine; this is what keeps the integer outputs (bboxes, n_contrib,
// tile_last) reproducible.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svr {

// Per-voxel candidate bounds of the forward, in the pixel frame (pixel
// (x, y) centred at (x, y)): o = bounds of x + y and of x - y over the
// projected corners (the octagon), tb = the pixel box [ceil(lo), floor(hi)]
// of the projected x and y extents, packed like the bbox (x << 16 | y), both
// widened by a rounding margin.  32 bytes, so one voxel's bounds sit in one
// cache line.
struct alignas(16) OctBox
{
    float4 o;
    uint2 tb;
    uint2 pad;
};

// backends/new_cuda/src/config.h:14-34
constexpr int TILE_X = 16;
constexpr int TILE_Y = 16;
constexpr int TILE_PIX = TILE_X * TILE_Y;
constexpr int MAX_NUM_LEVELS = 21;
constexpr int ORDER_RANK_BITS = 3 * MAX_NUM_LEVELS;
constexpr float MAX_ALPHA = 0.99999f;
constexpr float MIN_ALPHA = 0.00001f;
constexpr float EARLY_STOP_T = 0.0001f;
constexpr float STEP_SZ_SCALE = 100.f;
constexpr int MAX_N_SAMP = 3;
constexpr int MAX_SORT_KEY_BITS = 128;
constexpr int MAX_PACKED_IMAGE_DIM = (1 << 16) - 1;
constexpr int MAX_TILE_GRID_X = (MAX_PACKED_IMAGE_DIM + TILE_X) / TILE_X;
constexpr int MAX_TILE_GRID_Y = (MAX_PACKED_IMAGE_DIM + TILE_Y) / TILE_Y;
constexpr int64_t MAX_RENDER_TILES = (int64_t)MAX_TILE_GRID_X * MAX_TILE_GRID_Y;
constexpr uint32_t VOX_ID_MASK = (1u << 29) - 1u;

__host__ __device__ constexpr uint64_t order_table(uint32_t q)
{
    // q repeated in all 21 three-bit groups (auxiliary.h:33-55): q * sum_{i<21} 8^i
    return (uint64_t)q * 0x1249249249249249ull;
}
static_assert(order_table(7) == 0x7FFFFFFFFFFFFFFFull, "order table must cover all 21 levels");
static_assert(order_table(1) == ((1ull << 63) - 1ull) / 7ull, "order table must repeat q 21 times");

__host__ __device__ inline uint32_t required_bits(uint64_t n)
{
    uint32_t b = 0;
    if (n <= 1) return 0;
    n -= 1;
    while (n) { n >>= 1; ++b; }
    return b;
}

// svr_expf without its overflow and NaN checks: the same value for every
// x <= 88.72283935546875 (above that svr_expf returns +inf and this one the
// ldexp overflow, +inf, up to the largest reduced arguments); a NaN input
// gives a NaN either way.  The compositors call it on -vol <= 0 and on the
// exp_linear_11 argument, which is <= 0.1 wherever its value is used.
__device__ inline float svr_expf_core(float x)
{
    const float kf = __builtin_rintf(x * 1.44269502162933349609375f);
    float r = __builtin_fmaf(-kf, 0.693145751953125f, x);
    r = __builtin_fmaf(-kf, 1.428606765330187045e-06f, r);
    float p = 1.98412701976485550403594970703125e-04f;
    p = __builtin_fmaf(p, r, 1.388888922519981861114501953125e-03f);
    p = __builtin_fmaf(p, r, 8.33333376795053482055664062500e-03f);
    p = __builtin_fmaf(p, r, 4.16666679084300994873046875000e-02f);
    p = __builtin_fmaf(p, r, 1.66666671633720397949218750000e-01f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    // exact scaling by 2^k with a single rounding (also into the denormal range)
    const float res = __builtin_ldexpf(p, (int)kf);
    return (x < -103.97208404541015625f) ? 0.f : res;
}

__device__ inline float svr_expf(float x)
{
    float res = svr_expf_core(x);
    res = (x > 88.72283935546875f) ? __uint_as_float(0x7f800000u) : res;
    res = (x != x) ? x : res;
    return res;
}

// auxiliary.h:348-356
__device__ inline float exp_linear_11(float x)
{
    return (x > 1.1f) ? x : svr_expf(__builtin_fmaf(0.909090909091f, x, -0.904689820196f));
}

__device__ inline float exp_linear_11_bw(float x)
{
    return (x > 1.1f) ? 1.0f : 0.909090909091f * svr_expf(__builtin_fmaf(0.909090909091f, x, -0.904689820196f));
}

__device__ inline float smin(float a, float b) { return (a < b) ? a : b; }
__device__ inline float smax(float a, float b) { return (a > b) ? a : b; }

__device__ inline float inv_nsamp(int n) { return n == 1 ? 1.f : (n == 2 ? 0.5f : 1.f / 3.f); }

struct PixRay
{
    float ro[3], rd[3], rd_inv[3];
    float rd_norm_inv;  // 1 / |camera-space direction| (depth is scaled by it)
    uint32_t pq;
};

// forward.cu:102-113; auxiliary.h:140-148, 294-302
__device__ inline PixRay make_ray(int px, int py, int W, int H, float tanfovx, float tanfovy,
                                  float cx, float cy, const float* __restrict__ c2w)
{
    PixRay r;
    const float pfx = (float)px, pfy = (float)py;
    const float dx = (pfx + 0.5f - cx) * 2.f * tanfovx / (float)W;
    const float dy = (pfy + 0.5f - cy) * 2.f * tanfovy / (float)H;
    const float dz = 1.f;
    const float nrm = __builtin_sqrtf(dx * dx + dy * dy + dz * dz);
    const float rx = c2w[0] * dx + c2w[1] * dy + c2w[2] * dz;
    const float ry = c2w[4] * dx + c2w[5] * dy + c2w[6] * dz;
    const float rz = c2w[8] * dx + c2w[9] * dy + c2w[10] * dz;
    const float ninv = 1.f / nrm;
    r.rd_norm_inv = ninv;
    r.ro[0] = c2w[3]; r.ro[1] = c2w[7]; r.ro[2] = c2w[11];
    r.rd[0] = rx * ninv; r.rd[1] = ry * ninv; r.rd[2] = rz * ninv;
    r.rd_inv[0] = 1.f / r.rd[0]; r.rd_inv[1] = 1.f / r.rd[1]; r.rd_inv[2] = 1.f / r.rd[2];
    r.pq = ((r.rd[0] < 0.f) << 2) | ((r.rd[1] < 0.f) << 1) | (r.rd[2] < 0.f);
    return r;
}

// auxiliary.h:310-323.  The reference's min/max are ternaries ((a < b) ? a : b).
// They differ from the hardware min/max only for NaN operands and in the sign
// of a zero result.  The sign of a zero a or b cannot change any output (the
// test a > b, the segment length b - a and the first sample point
// a + step / 2 come out the same), and a slab bound is NaN only when 1/rd is
// not finite (a zero or denormal direction component).  EXACT = false uses
// v_min/v_max; the kernels use it for waves whose rays all have finite 1/rd.
template <bool EXACT>
__device__ inline void ray_aabb(float cx_, float cy_, float cz_, float l, const PixRay& ray, float& a, float& b)
{
    const float r = 0.5f * l;
    const float d0 = cx_ - ray.ro[0], d1 = cy_ - ray.ro[1], d2 = cz_ - ray.ro[2];
    const float a0 = (d0 - r) * ray.rd_inv[0], b0 = (d0 + r) * ray.rd_inv[0];
    const float a1 = (d1 - r) * ray.rd_inv[1], b1 = (d1 + r) * ray.rd_inv[1];
    const float a2 = (d2 - r) * ray.rd_inv[2], b2 = (d2 + r) * ray.rd_inv[2];
    if constexpr (EXACT)
    {
        const float lo0 = smin(a0, b0), hi0 = smax(a0, b0);
        const float lo1 = smin(a1, b1), hi1 = smax(a1, b1);
        const float lo2 = smin(a2, b2), hi2 = smax(a2, b2);
        a = smax(smax(lo0, lo1), lo2);
        b = smin(smin(hi0, hi1), hi2);
    }
    else
    {
        // med3(x, y, -inf) = min(x, y) and med3(x, y, +inf) = max(x, y); written
        // this way the compiler emits plain v_min/v_max (and v_min3/v_max3)
        // without quieting the operands first
        const float ninf = -__builtin_inff(), pinf = __builtin_inff();
        const float lo0 = __builtin_amdgcn_fmed3f(a0, b0, ninf), hi0 = __builtin_amdgcn_fmed3f(a0, b0, pinf);
        const float lo1 = __builtin_amdgcn_fmed3f(a1, b1, ninf), hi1 = __builtin_amdgcn_fmed3f(a1, b1, pinf);
        const float lo2 = __builtin_amdgcn_fmed3f(a2, b2, ninf), hi2 = __builtin_amdgcn_fmed3f(a2, b2, pinf);
        a = __builtin_amdgcn_fmed3f(__builtin_amdgcn_fmed3f(lo0, lo1, pinf), lo2, pinf);
        b = __builtin_amdgcn_fmed3f(__builtin_amdgcn_fmed3f(hi0, hi1, ninf), hi2, ninf);
    }
}

// The view-constant part of ray_aabb and of the sample positions for one
// entry: every ray of a view starts at the camera centre ro, so
// lo_k = (c_k - ro_k) - r, hi_k = (c_k - ro_k) + r and cm_k = c_k - r
// (r = l / 2) can be computed once per entry.  The roundings are those of
// ray_aabb and vox_vol, so ray_aabb_pre and the *_cm forms below give the
// same bits as the forms that take the centre and length.
__device__ inline void entry_bounds(float c0, float c1, float c2, float l, const float* ro, float* lo, float* hi,
                                    float* cm)
{
    const float r = 0.5f * l;
    const float d0 = c0 - ro[0], d1 = c1 - ro[1], d2 = c2 - ro[2];
    lo[0] = d0 - r; lo[1] = d1 - r; lo[2] = d2 - r;
    hi[0] = d0 + r; hi[1] = d1 + r; hi[2] = d2 + r;
    cm[0] = c0 - r; cm[1] = c1 - r; cm[2] = c2 - r;
}

template <bool EXACT>
__device__ inline void ray_aabb_pre(float lo0, float lo1, float lo2, float hi0, float hi1, float hi2,
                                    const PixRay& ray, float& a, float& b)
{
    const float a0 = lo0 * ray.rd_inv[0], b0 = hi0 * ray.rd_inv[0];
    const float a1 = lo1 * ray.rd_inv[1], b1 = hi1 * ray.rd_inv[1];
    const float a2 = lo2 * ray.rd_inv[2], b2 = hi2 * ray.rd_inv[2];
    if constexpr (EXACT)
    {
        const float lo0_ = smin(a0, b0), hi0_ = smax(a0, b0);
        const float lo1_ = smin(a1, b1), hi1_ = smax(a1, b1);
        const float lo2_ = smin(a2, b2), hi2_ = smax(a2, b2);
        a = smax(smax(lo0_, lo1_), lo2_);
        b = smin(smin(hi0_, hi1_), hi2_);
    }
    else
    {
        const float ninf = -__builtin_inff(), pinf = __builtin_inff();
        const float lo0_ = __builtin_amdgcn_fmed3f(a0, b0, ninf), hi0_ = __builtin_amdgcn_fmed3f(a0, b0, pinf);
        const float lo1_ = __builtin_amdgcn_fmed3f(a1, b1, ninf), hi1_ = __builtin_amdgcn_fmed3f(a1, b1, pinf);
        const float lo2_ = __builtin_amdgcn_fmed3f(a2, b2, ninf), hi2_ = __builtin_amdgcn_fmed3f(a2, b2, pinf);
        a = __builtin_amdgcn_fmed3f(__builtin_amdgcn_fmed3f(lo0_, lo1_, pinf), lo2_, pinf);
        b = __builtin_amdgcn_fmed3f(__builtin_amdgcn_fmed3f(hi0_, hi1_, ninf), hi2_, ninf);
    }
}

// True when every component of 1/rd is finite (the ray-box slabs cannot be NaN).
__device__ inline bool ray_inv_finite(const PixRay& r)
{
    return __builtin_fabsf(r.rd_inv[0]) < __builtin_inff() && __builtin_fabsf(r.rd_inv[1]) < __builtin_inff() &&
           __builtin_fabsf(r.rd_inv[2]) < __builtin_inff();
}

// Volume integral of one voxel segment (forward.cu:224-253, backward.cu:283-328).
// When DI is true, dI[k] receives sum_s 100*step*exp_linear_11'(d_s)*w_k(s).
// cm_k = c_k - r (entry_bounds).
template <int NS, bool DI>
__device__ inline float vox_vol_cm(float cm0, float cm1, float cm2, float l_inv, const float* geo, float a, float b,
                                   const PixRay& ray, float* dI)
{
    const float step_sz = (b - a) * inv_nsamp(NS);
    const float t0 = __builtin_fmaf(0.5f, step_sz, a);
    float qt0 = (__builtin_fmaf(t0, ray.rd[0], ray.ro[0]) - cm0) * l_inv;
    float qt1 = (__builtin_fmaf(t0, ray.rd[1], ray.ro[1]) - cm1) * l_inv;
    float qt2 = (__builtin_fmaf(t0, ray.rd[2], ray.ro[2]) - cm2) * l_inv;
    const float qs0 = (step_sz * ray.rd[0]) * l_inv;
    const float qs1 = (step_sz * ray.rd[1]) * l_inv;
    const float qs2 = (step_sz * ray.rd[2]) * l_inv;
    const float scale = STEP_SZ_SCALE * step_sz;
    float vol = 0.f;
    if (DI)
    {
#pragma unroll
        for (int i = 0; i < 8; ++i) dI[i] = 0.f;
    }
#pragma unroll
    for (int k = 0; k < NS; ++k)
    {
        const float wx[2] = {1.f - qt0, qt0};
        const float wy[2] = {1.f - qt1, qt1};
        const float wz[2] = {1.f - qt2, qt2};
        float w[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) w[i] = (wx[(i >> 2) & 1] * wy[(i >> 1) & 1]) * wz[i & 1];
        float d = geo[0] * w[0];
#pragma unroll
        for (int i = 1; i < 8; ++i) d = __builtin_fmaf(geo[i], w[i], d);
        // exp_linear_11 and its derivative share one exponential
        const float ex = svr_expf_core(__builtin_fmaf(0.909090909091f, d, -0.904689820196f));
        const bool lin = d > 1.1f;
        // the first sample's term is the sum itself: 0 + x == x for every
        // x >= +0 or NaN (scale and the activation are never negative)
        vol = (k == 0) ? scale * (lin ? d : ex) : vol + scale * (lin ? d : ex);
        if (DI)
        {
            const float dd = scale * (lin ? 1.0f : 0.909090909091f * ex);
#pragma unroll
            for (int i = 0; i < 8; ++i) dI[i] = __builtin_fmaf(dd, w[i], dI[i]);
        }
        qt0 = qt0 + qs0; qt1 = qt1 + qs1; qt2 = qt2 + qs2;
    }
    return vol;
}

template <int NS, bool DI>
__device__ inline float vox_vol(float c0, float c1, float c2, float l, float l_inv, const float* geo, float a,
                                float b, const PixRay& ray, float* dI)
{
    const float r = 0.5f * l;
    return vox_vol_cm<NS, DI>(c0 - r, c1 - r, c2 - r, l_inv, geo, a, b, ray, dI);
}

// vox_vol<1, true> with the density derivative kept apart: w receives the 8
// trilinear weights and dd the factor 100*step*exp_linear_11'(d), so that
// dI[i] = dd * w[i] (the same vol; the backward scales w by dL/dvol * dd once).
__device__ inline float vox_vol1_w_cm(float cm0, float cm1, float cm2, float l_inv, const float* geo, float a,
                                      float b, const PixRay& ray, float* w, float& dd)
{
    const float step_sz = b - a;
    const float t0 = __builtin_fmaf(0.5f, step_sz, a);
    const float qt0 = (__builtin_fmaf(t0, ray.rd[0], ray.ro[0]) - cm0) * l_inv;
    const float qt1 = (__builtin_fmaf(t0, ray.rd[1], ray.ro[1]) - cm1) * l_inv;
    const float qt2 = (__builtin_fmaf(t0, ray.rd[2], ray.ro[2]) - cm2) * l_inv;
    const float scale = STEP_SZ_SCALE * step_sz;
    const float wx[2] = {1.f - qt0, qt0};
    const float wy[2] = {1.f - qt1, qt1};
    const float wz[2] = {1.f - qt2, qt2};
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = (wx[(i >> 2) & 1] * wy[(i >> 1) & 1]) * wz[i & 1];
    float d = geo[0] * w[0];
#pragma unroll
    for (int i = 1; i < 8; ++i) d = __builtin_fmaf(geo[i], w[i], d);
    const float ex = svr_expf_core(__builtin_fmaf(0.909090909091f, d, -0.904689820196f));
    const bool lin = d > 1.1f;
    dd = scale * (lin ? 1.0f : 0.909090909091f * ex);
    return scale * (lin ? d : ex);  // == vox_vol<1>: 0 + x is x for x >= +0
}

// x / y from the hardware reciprocal and one Newton correction (within an
// ulp or so of the IEEE quotient; y is a normal positive number here).  The
// backward recovers the transmittance before a hit with it; only gradients,
// held to a relative tolerance, depend on it.
__device__ inline float div_refined(float x, float y)
{
    const float r = __builtin_amdgcn_rcpf(y);
    const float q = x * r;
    return __builtin_fmaf(__builtin_fmaf(-q, y, x), r, q);
}

// Per-sample form of vox_vol (same vol and dI): lv[k] is sample k's volume
// term and, with DI, edI[k][i] = dd_k * w_k[i] its share of dI[i]
// (forward.cu:235-253 / backward.cu:300-320 keep them for the depth output).
template <int NS, bool DI>
__device__ inline float vox_samples_cm(float cm0, float cm1, float cm2, float l_inv, const float* geo, float a,
                                       float b, const PixRay& ray, float* lv, float (*edI)[8], float* dI)
{
    const float step_sz = (b - a) * inv_nsamp(NS);
    const float t0 = __builtin_fmaf(0.5f, step_sz, a);
    float qt0 = (__builtin_fmaf(t0, ray.rd[0], ray.ro[0]) - cm0) * l_inv;
    float qt1 = (__builtin_fmaf(t0, ray.rd[1], ray.ro[1]) - cm1) * l_inv;
    float qt2 = (__builtin_fmaf(t0, ray.rd[2], ray.ro[2]) - cm2) * l_inv;
    const float qs0 = (step_sz * ray.rd[0]) * l_inv;
    const float qs1 = (step_sz * ray.rd[1]) * l_inv;
    const float qs2 = (step_sz * ray.rd[2]) * l_inv;
    const float scale = STEP_SZ_SCALE * step_sz;
    float vol = 0.f;
    if (DI)
    {
#pragma unroll
        for (int i = 0; i < 8; ++i) dI[i] = 0.f;
    }
#pragma unroll
    for (int k = 0; k < NS; ++k)
    {
        const float wx[2] = {1.f - qt0, qt0};
        const float wy[2] = {1.f - qt1, qt1};
        const float wz[2] = {1.f - qt2, qt2};
        float w[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) w[i] = (wx[(i >> 2) & 1] * wy[(i >> 1) & 1]) * wz[i & 1];
        float d = geo[0] * w[0];
#pragma unroll
        for (int i = 1; i < 8; ++i) d = __builtin_fmaf(geo[i], w[i], d);
        const float ex = svr_expf_core(__builtin_fmaf(0.909090909091f, d, -0.904689820196f));
        const bool lin = d > 1.1f;
        const float term = scale * (lin ? d : ex);
        lv[k] = term;
        vol = vol + term;
        if (DI)
        {
            const float dd = scale * (lin ? 1.0f : 0.909090909091f * ex);
#pragma unroll
            for (int i = 0; i < 8; ++i)
            {
                edI[k][i] = dd * w[i];
                dI[i] = __builtin_fmaf(dd, w[i], dI[i]);
            }
        }
        qt0 = qt0 + qs0; qt1 = qt1 + qs1; qt2 = qt2 + qs2;
    }
    return vol;
}

// Transpose a 64x64 bit matrix held one 64-bit row per lane (row = lane):
// afterwards lane c holds bit r set iff lane r held bit c.
__device__ inline uint64_t wave_transpose64(uint64_t row, int lane)
{
    const uint64_t masks[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                               0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
    for (int i = 0; i < 6; ++i)
    {
        const int s = 32 >> i;
        const uint64_t m = masks[i];
        const uint32_t lo = __shfl_xor((uint32_t)row, s, 64);
        const uint32_t hi = __shfl_xor((uint32_t)(row >> 32), s, 64);
        const uint64_t other = ((uint64_t)hi << 32) | lo;
        row = (lane & s) ? ((row & ~m) | ((other & ~m) >> s)) : ((row & m) | ((other & m) << s));
    }
    return row;
}

template <int CTRL>
__device__ inline float dpp_f32(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sums of 12 values over the 64 lanes of a wave (all lanes must be active) by
// a halving butterfly: each exchange step keeps half of the remaining values
// in each lane and adds its partner's copy of them.  Lanes i and i ^ 32
// (v_permlane32_swap) pair value k with value k + 6, leaving 6 registers of 2
// values; lanes i and i ^ 16 (v_permlane16_swap) pair those registers two by
// two, leaving 3 registers whose 16-lane row r holds value
// 2k + (r & 1) + 6 (r >> 1) of register k.  Within the rows, a DPP row rotation
// (i ^ 8) adds registers 0 and 1 into one register (lanes with bit 3 clear
// keep register 0) and register 2 into another, a half-row mirror (i ^ 7)
// adds those two (lanes with bit 2 clear keep the first), and two DPP steps
// finish the sum.  About 28 instructions for all 12 sums instead of 12
// separate reductions.  Lane i returns the total of value
// wave_sum12_index(i), or a copy of another lane's total where that is -1.
__device__ inline float wave_sum12(const float (&g)[12], int lane)
{
    float u[6];
#pragma unroll
    for (int k = 0; k < 6; ++k)
    {
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(g[k]), __float_as_uint(g[k + 6]), false,
                                                        false);
        u[k] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    float w[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(u[2 * k]), __float_as_uint(u[2 * k + 1]),
                                                        false, false);
        w[k] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    const bool b3 = (lane >> 3) & 1, b2 = (lane >> 2) & 1;
    const float x0 = (b3 ? w[1] : w[0]) + dpp_f32<0x128>(b3 ? w[0] : w[1]);  // row_ror:8
    const float x1 = w[2] + dpp_f32<0x128>(w[2]);
    float y = (b2 ? x1 : x0) + dpp_f32<0x141>(b2 ? x0 : x1);  // row_half_mirror
    y = y + dpp_f32<0xB1>(y);  // quad_perm [1,0,3,2]
    y = y + dpp_f32<0x4E>(y);  // quad_perm [2,3,0,1]
    return y;
}

// The value whose total wave_sum12 returns in lane i (lanes 4g .. 4g + 3 hold
// the same one), or -1 where the lane holds a copy of register 2's total that
// the lanes 8 below it hold as well.
__device__ inline int wave_sum12_index(int lane)
{
    const int r = lane >> 4, h = (lane >> 3) & 1, q = (lane >> 2) & 1;
    return (q && h) ? -1 : (q ? 4 : 2 * h) + (r & 1) + 6 * (r >> 1);
}

// Sum over the 64 lanes of a wave (all lanes must be active); every lane gets
// the total.  DPP within each 16-lane row, then the four row sums by readlane.
__device__ inline float wave_sum_f32(float v)
{
    v = v + dpp_f32<0xB1>(v);   // quad_perm [1,0,3,2]
    v = v + dpp_f32<0x4E>(v);   // quad_perm [2,3,0,1]
    v = v + dpp_f32<0x141>(v);  // row_half_mirror
    v = v + dpp_f32<0x140>(v);  // row_mirror
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return (r0 + r1) + (r2 + r3);
}

__device__ inline uint64_t lanemask_lt(int lane) { return (lane == 0) ? 0ull : (~0ull >> (64 - lane)); }

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Scheduling bucket of a tile's list length: 8 sub-buckets per octave (the
// bit length and the next three bits), 0..255, larger for longer lists.  The
// compositors take tiles longest bucket first: per-tile work spans three
// orders of magnitude, and a long list that starts late is the frame's tail.
__device__ inline uint32_t length_bucket(uint32_t len)
{
    if (len == 0) return 0u;
    const uint32_t b = 32u - (uint32_t)__clz((int)len);
    const uint32_t top4 = b >= 4 ? (len >> (b - 4)) : (len << (4 - b));
    return 8u * (b - 1u) + (top4 & 7u);
}

}  // namespace svr
