// The compositor with its optional outputs and regularizer gradients: mean and
// median depth, the distortion sum, the rendered normal, the per-voxel maximum
// blending weight, and in the adjoint the depth / normal / distortion
// gradients plus the R_concen and ascending regularizers.
//
// Reference: backends/new_cuda/src/forward.cu:150-394 (need_depth,
// need_distortion, need_normal, track_max_w branches of renderCUDA),
// backward.cu:175-527 (need_depth, need_normal, R_concen, ascending,
// distortion branches), host sides forward.cu:792-879, backward.cu:654-772.
//
// Same tile mapping as the plain compositor (one 256-thread workgroup per
// 16x16 tile, one wave per 8x8 sub-tile, entries staged 64 at a time and
// dispatched to pixels through a transposed bit matrix); the per-pixel
// arithmetic follows the reference's order, so the forward outputs of a pixel
// do not depend on the mapping.  The kernels are used only when one of the
// options is on; the plain kernels keep the default path free of the extra
// registers.  The options are runtime flags, uniform over the launch.
#include "compositor.h"
#include "svr_host.h"

namespace svr {

enum : int
{
    EXT_DEPTH = 1,      // mean + median depth (out_depth channels 0, 2; backward: dL/d mean depth)
    EXT_DIST = 2,       // distortion sum (out_depth channel 1; backward: lambda_dist > 0)
    EXT_NORMAL = 4,     // rendered normal
    EXT_MAX_W = 8,      // per-voxel max blending weight (forward only)
    EXT_R_CONCEN = 16,  // backward only
    EXT_ASCENDING = 32, // backward only
};

// auxiliary.h:305-308, 368-371
__device__ inline float depth_contracted(float x) { return (x < 1.f) ? x : 2.f - 1.f / x; }
__device__ inline float safe_rnorm(float x, float y, float z)
{
    return 1.f / __builtin_sqrtf(x * x + y * y + z * z + 1e-15f);
}

// forward.cu:301-314: density gradient of the trilinear field (corner bit2 = x,
// bit1 = y, bit0 = z), summed in the reference's order.
__device__ inline void lin_normal(const float* g, float* n)
{
    n[0] = (g[4] + g[5] + g[6] + g[7]) - (g[0] + g[1] + g[2] + g[3]);
    n[1] = (g[2] + g[3] + g[6] + g[7]) - (g[0] + g[1] + g[4] + g[5]);
    n[2] = (g[1] + g[3] + g[5] + g[7]) - (g[0] + g[2] + g[4] + g[6]);
}

// forward.cu:262-299: expected depth of one entry along its own samples
// (alpha of each sample from its volume term lv[k]) or at the segment
// midpoint for one sample; dD[k] = d depth / d lv[k] up to the factor
// dL_dD * T of backward.cu:421-454.
template <int NS>
__device__ inline float entry_depth(float a, float b, float alpha, const float* lv, float* dD)
{
    if constexpr (NS == 3)
    {
        const float st = 0.3333333f * (b - a);
        const float a0 = fminf(MAX_ALPHA, 1.f - svr_expf_core(-lv[0]));
        const float a1 = fminf(MAX_ALPHA, 1.f - svr_expf_core(-lv[1]));
        const float a2 = fminf(MAX_ALPHA, 1.f - svr_expf_core(-lv[2]));
        const float t0 = a + 0.5f * st, t1 = a + 1.5f * st, t2 = a + 2.5f * st;
        if (dD)
        {
            dD[0] = (t0 + a1 * a2 * t2 - a1 * t1 - a2 * t2) * (1.f - a0);
            dD[1] = (t1 + a0 * a2 * t2 - a0 * t1 - a2 * t2) * (1.f - a1);
            dD[2] = (t2 + a0 * a1 * t2 - a0 * t2 - a1 * t2) * (1.f - a2);
        }
        return a0 * t0 + (1.f - a0) * a1 * t1 + (1.f - a0) * (1.f - a1) * a2 * t2;
    }
    else if constexpr (NS == 2)
    {
        const float st = 0.5f * (b - a);
        const float a0 = fminf(MAX_ALPHA, 1.f - svr_expf_core(-lv[0]));
        const float a1 = fminf(MAX_ALPHA, 1.f - svr_expf_core(-lv[1]));
        const float t0 = a + 0.5f * st, t1 = a + 1.5f * st;
        if (dD)
        {
            dD[0] = (t0 - a1 * t1) * (1.f - a0);
            dD[1] = (t1 - a0 * t1) * (1.f - a1);
        }
        return a0 * t0 + (1.f - a0) * a1 * t1;
    }
    else
    {
        const float t0 = 0.5f * (a + b);
        if (dD) dD[0] = t0 * (1.f - alpha);
        return alpha * t0;
    }
}

// Density at the normalized in-voxel point q (trilinear weights in w).
__device__ inline float density_at(const float* g, const float* q, float* w)
{
    const float wx[2] = {1.f - q[0], q[0]};
    const float wy[2] = {1.f - q[1], q[1]};
    const float wz[2] = {1.f - q[2], q[2]};
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = (wx[(i >> 2) & 1] * wy[(i >> 1) & 1]) * wz[i & 1];
    float d = g[0] * w[0];
#pragma unroll
    for (int i = 1; i < 8; ++i) d = __builtin_fmaf(g[i], w[i], d);
    return d;
}

// Running sums of one pixel in the forward.
struct ExtAcc
{
    float D, Dd, N0, N1, N2, med_T;
    int med_v;
};

// composite_subtile with the optional outputs (forward.cu:150-328).
template <int NS, bool EXACT, int NW>
__device__ inline uint32_t composite_ext(StageT<NW>& st, uint32_t (&sv)[NW][64], int w, int lane, const PixRay& ray,
                                         bool inside, uint2 range, int sx0, int sy0, int sx1, int sy1,
                                         const uint32_t* __restrict__ vox_list, const uint2* __restrict__ bboxes, const OctBox* __restrict__ octs,
                                         const float* __restrict__ centers, const float* __restrict__ lengths,
                                         const float* __restrict__ geos, const float* __restrict__ rgbs, int flags,
                                         float* __restrict__ max_w, float& T, float& C0, float& C1, float& C2,
                                         ExtAcc& x, uint64_t* __restrict__ mask_out)
{
    uint32_t last = 0;
    bool done = !inside;
    init_diag(st.diag[w], lane);
    wave_sync();
    uint64_t qpix[8];
#pragma unroll
    for (uint32_t q = 0; q < 8; ++q) qpix[q] = __ballot(!done && ray.pq == q);
    EntryRegs cur;
    uint32_t base = range.x;
    {
        const uint32_t j = base + (uint32_t)lane;
        const bool valid = j < range.y;
        load_entry<true, !EXACT>(cur, valid, valid ? vox_list[j] : 0u, bboxes, octs, centers, lengths, geos, rgbs, sx0, sy0, sx1, sy1,
                   qpix, st.diag[w], ray.ro);
    }
    while (base < range.y)
    {
        wave_sync();
        stage_entry(st, w, lane, cur);
        if (cur.pm) sv[w][lane] = cur.v;
        wave_sync();
        uint64_t hm = wave_transpose64(cur.pm, lane);
        if (done) hm = 0ull;

        const uint32_t nbase = base + 64;
        if (nbase < range.y)
        {
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) qpix[q] = __ballot(!done && ray.pq == q);
            const uint32_t j = nbase + (uint32_t)lane;
            const bool valid = j < range.y;
            load_entry<true, !EXACT>(cur, valid, valid ? vox_list[j] : 0u, bboxes, octs, centers, lengths, geos, rgbs, sx0, sy0, sx1,
                       sy1, qpix, st.diag[w], ray.ro);
        }

        uint64_t hits = 0ull;
        while (hm)
        {
            // a lane whose candidate misses the box tests its next ones in the same step, as in the
            // plain forward
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
            float g[8], lv[NS];
#pragma unroll
            for (int i = 0; i < 8; ++i) g[i] = st.g[i][w][l];
            const float vol = vox_samples_cm<NS, false>(st.cm[0][w][l], st.cm[1][w][l], st.cm[2][w][l],
                                                        st.linv[w][l], g, a, b, ray, lv, nullptr, nullptr);
            const float alpha = fminf(MAX_ALPHA, 1.f - svr_expf_core(-vol));
            if (alpha < MIN_ALPHA) continue;
            hits |= 1ull << l;
            const float pt_w = T * alpha;
            C0 = __builtin_fmaf(pt_w, st.rgb[0][w][l], C0);
            C1 = __builtin_fmaf(pt_w, st.rgb[1][w][l], C1);
            C2 = __builtin_fmaf(pt_w, st.rgb[2][w][l], C2);
            if (flags & EXT_DEPTH)
            {
                x.D = x.D + T * entry_depth<NS>(a, b, alpha, lv, nullptr);
                if (T > 0.5f)
                {
                    x.med_v = (int)sv[w][l];
                    x.med_T = T;
                }
            }
            if (flags & EXT_DIST)
            {
                const float m = 0.5f * (depth_contracted(a) + depth_contracted(b));
                x.Dd = x.Dd + pt_w * m;
            }
            if (flags & EXT_NORMAL)
            {
                float n[3];
                lin_normal(g, n);
                const float rl = safe_rnorm(n[0], n[1], n[2]);
                x.N0 = x.N0 + pt_w * rl * n[0];
                x.N1 = x.N1 + pt_w * rl * n[1];
                x.N2 = x.N2 + pt_w * rl * n[2];
            }
            T = T * (1.f - alpha);
            last = base + (uint32_t)l - range.x + 1u;
            // pt_w > 0: the order of non-negative floats is that of their bits
            if (flags & EXT_MAX_W) atomicMax(reinterpret_cast<unsigned int*>(max_w) + sv[w][l], __float_as_uint(pt_w));
            if (T < EARLY_STOP_T)
            {
                done = true;
                break;
            }
        }
        // the hit masks of the plain backward (a forward with options may be
        // followed by a backward without them)
        mask_out[(size_t)((base - range.x) >> 6) * (64 * NW)] = hits;
        base = nbase;
        if (__ballot(!done) == 0ull) break;
    }
    return last;
}

// forward.cu:330-366: the median depth is refined with 16 samples inside the
// last voxel entered with T > 0.5, stepping until T falls to 0.5.
template <bool EXACT>
__device__ inline float median_refine(const PixRay& ray, int v, float med_T, const float* __restrict__ centers,
                                      const float* __restrict__ lengths, const float* __restrict__ geos)
{
    const size_t mv = (size_t)v;
    const float c[3] = {centers[3 * mv + 0], centers[3 * mv + 1], centers[3 * mv + 2]};
    const float l = lengths[mv];
    float g[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) g[i] = geos[8 * mv + i];
    float a, b;
    ray_aabb<EXACT>(c[0], c[1], c[2], l, ray, a, b);
    const float l_inv = 1.f / l;
    const float step_sz = (b - a) * (1.f / 16.f);
    const float t0 = __builtin_fmaf(0.5f, step_sz, a);
    const float r = 0.5f * l;
    float qt[3], qs[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        qt[i] = (__builtin_fmaf(t0, ray.rd[i], ray.ro[i]) - (c[i] - r)) * l_inv;
        qs[i] = (step_sz * ray.rd[i]) * l_inv;
    }
    float Dmed = a - 0.5f * step_sz;
    const float scale = STEP_SZ_SCALE * step_sz;
    for (int k = 0; k < 16 && med_T > 0.5f; ++k)
    {
        Dmed = Dmed + step_sz;
        float w[8];
        const float d = density_at(g, qt, w);
        med_T = med_T * svr_expf(-(scale * exp_linear_11(d)));
#pragma unroll
        for (int i = 0; i < 3; ++i) qt[i] = qt[i] + qs[i];
    }
    return Dmed;
}

constexpr unsigned EXT_EXACT_GRID = 64;

template <int NS, bool EXACT>
__global__ __launch_bounds__(256) void render_fwd_ext_kernel(
    int W, int H, int gx, const uint32_t* __restrict__ tile_order, float tanfovx, float tanfovy, float cx, float cy,
    const float* __restrict__ c2w, float bg, const uint2* __restrict__ ranges, const uint32_t* __restrict__ vox_list,
    const uint2* __restrict__ bboxes, const OctBox* __restrict__ octs, const float* __restrict__ centers, const float* __restrict__ lengths,
    const float* __restrict__ geos, const float* __restrict__ rgbs, int flags, uint32_t* __restrict__ tile_last,
    uint32_t* __restrict__ n_contrib, float* __restrict__ out_color, float* __restrict__ out_T,
    float* __restrict__ out_depth, float* __restrict__ out_normal, float* __restrict__ max_w,
    uint64_t* __restrict__ hit_masks, uint32_t* __restrict__ exact_ctr, uint32_t* __restrict__ exact_list)
{
    __shared__ Stage st;
    __shared__ uint32_t sv[NWAVE][64];
    __shared__ uint32_t s_last[NWAVE];
    // EXACT = false: workgroup b takes tile order[b] and lists it for the
    // exact launch if one of its rays has a non-finite 1/rd; EXACT = true: a
    // few workgroups loop over that list (render.hip's protocol)
    const uint32_t count = EXACT ? exact_ctr[0] : 1u;
    for (uint32_t it = EXACT ? blockIdx.x : 0u; it < count; it += EXACT ? gridDim.x : 1u)
    {
        const int tile = EXACT ? (int)exact_list[it] : (int)tile_order[blockIdx.x];
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int tx = tile % gx, ty = tile / gx;
        const int sx0 = tx * TILE_X + (w & 1) * 8, sy0 = ty * TILE_Y + (w >> 1) * 8;
        const int sx1 = min(sx0 + 7, W - 1), sy1 = min(sy0 + 7, H - 1);
        const int px = sx0 + (lane & 7), py = sy0 + (lane >> 3);
        const bool inside = px < W && py < H;
        const PixRay ray = make_ray(px, py, W, H, tanfovx, tanfovy, cx, cy, c2w);
        if (!EXACT && __syncthreads_or(!ray_inv_finite(ray)) != 0)
        {
            if (threadIdx.x == 0) exact_list[atomicAdd(exact_ctr, 1u)] = (uint32_t)tile;
            return;
        }
        const uint2 range = ranges[tile];

        float T = 1.f, C0 = 0.f, C1 = 0.f, C2 = 0.f;
        ExtAcc x{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, -1};
        uint64_t* mask_out = hit_masks + (((size_t)(range.x >> 6) + tile) * NWAVE + w) * 64 + lane;
        const uint32_t last = composite_ext<NS, EXACT>(st, sv, w, lane, ray, inside, range, sx0, sy0, sx1, sy1, vox_list,
                                                       bboxes, octs, centers, lengths, geos, rgbs, flags, max_w, T, C0, C1, C2,
                                                       x, mask_out);
        if (inside)
        {
            const size_t pid = (size_t)py * W + px;
            const size_t HW = (size_t)H * W;
            n_contrib[pid] = last;
            out_color[pid] = C0 + T * bg;
            out_color[HW + pid] = C1 + T * bg;
            out_color[2 * HW + pid] = C2 + T * bg;
            out_T[pid] = T;
            if (out_depth)
            {
                float Dmed = 0.f;
                if ((flags & EXT_DEPTH) && x.med_v >= 0)
                    Dmed = median_refine<EXACT>(ray, x.med_v, x.med_T, centers, lengths, geos);
                out_depth[pid] = (flags & EXT_DEPTH) ? x.D * ray.rd_norm_inv : 0.f;
                out_depth[HW + pid] = (flags & EXT_DIST) ? x.Dd : 0.f;
                out_depth[2 * HW + pid] = (flags & EXT_DEPTH) ? Dmed * ray.rd_norm_inv : 0.f;
            }
            if (flags & EXT_NORMAL)
            {
                out_normal[pid] = x.N0;
                out_normal[HW + pid] = x.N1;
                out_normal[2 * HW + pid] = x.N2;
            }
        }
        const uint32_t wm = wave_max_u32(inside ? last : 0u);
        if (lane == 0) s_last[w] = wm;
        __syncthreads();
        if (threadIdx.x == 0)
            tile_last[tile] = range.x + max(max(s_last[0], s_last[1]), max(s_last[2], s_last[3]));
        __syncthreads();
    }
}

// Per-pixel inputs and running sums of the adjoint.
struct ExtPix
{
    float dL_dD, dN0, dN1, dN2, gt0, gt1, gt2;
    float prefix_wm, suffix_wm, prefix_w, suffix_w;
};

struct ExtCoef
{
    int flags;
    float w_R, w_asc, w_dist;  // lambda / (W * H)
};

// Adjoint of one (pixel, entry) hit with the optional terms
// (backward.cu:259-527, in the reference's order of accumulation).
template <int NS, bool EXACT>
__device__ inline bool bwd_hit_ext(const EntryPay& pe, const PixRay& ray, const ExtCoef& k, ExtPix& p, float& T,
                                   float& last_dT, float dp0, float dp1, float dp2, float* gq)
{
    const float4 q0 = pe.q[0], q1 = pe.q[1], q2 = pe.q[2], g0 = pe.q[4], g1 = pe.q[5];
    float a, b;
    ray_aabb_pre<EXACT>(q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, ray, a, b);
    if (a > b) return false;
    const float g[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
    float dI[8], lv[NS], edI[NS][8];
    const float cm[3] = {q1.z, q1.w, q2.x};
    const float linv = q2.y;
    const float vol = vox_samples_cm<NS, true>(cm[0], cm[1], cm[2], linv, g, a, b, ray, lv, edI, dI);
    const float e = svr_expf_core(-vol);
    const float alpha = fminf(MAX_ALPHA, 1.f - e);
    if (alpha < MIN_ALPHA) return false;
    T = T / (1.f - alpha);
    const float pt_w = alpha * T;
    const float col[3] = {q2.z, q2.w, pe.q[3].x};
    float dpt = __builtin_fmaf(dp2, col[2], __builtin_fmaf(dp1, col[1], dp0 * col[0]));
    float dg[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) dg[i] = 0.f;
    if (k.flags & EXT_DIST)
    {
        // backward.cu:347-361 (DVGO's distortion gradient from prefix / suffix sums)
        const float ad = depth_contracted(a), bd = depth_contracted(b);
        const float now_m = 0.5f * (ad + bd);
        const float now_wm = now_m * pt_w;
        p.prefix_wm = p.prefix_wm - now_wm;
        p.prefix_w = p.prefix_w - pt_w;
        const float uni = 0.6666666f * pt_w * (bd - ad);
        const float bi = 2.f * (now_m * (p.prefix_w - p.suffix_w) - (p.prefix_wm - p.suffix_wm));
        dpt = dpt + k.w_dist * (uni + bi);
        p.suffix_wm = p.suffix_wm + now_wm;
        p.suffix_w = p.suffix_w + pt_w;
    }
    if (k.flags & EXT_NORMAL)
    {
        // backward.cu:366-400
        float n[3];
        lin_normal(g, n);
        const float rl = safe_rnorm(n[0], n[1], n[2]);
        const float sn[3] = {rl * n[0], rl * n[1], rl * n[2]};
        dpt = dpt + (p.dN0 * sn[0] + p.dN1 * sn[1] + p.dN2 * sn[2]);
        const float ds[3] = {pt_w * p.dN0, pt_w * p.dN1, pt_w * p.dN2};
        const float dsn = ds[0] * n[0] + ds[1] * n[1] + ds[2] * n[2];
        float dl[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) dl[i] = rl * (ds[i] - dsn * rl * sn[i]);
#pragma unroll
        for (int i = 0; i < 8; ++i)
            dg[i] = ((i & 4) ? dl[0] : -dl[0]) + ((i & 2) ? dl[1] : -dl[1]) + ((i & 1) ? dl[2] : -dl[2]);
    }
    const float diff = dpt - last_dT;
    const float dalpha = T * diff;
    last_dT = __builtin_fmaf(alpha, diff, last_dT);
    const float dIs = dalpha * e;
#pragma unroll
    for (int i = 0; i < 8; ++i) dg[i] = dg[i] + dIs * dI[i];
    if (k.flags & EXT_DEPTH)
    {
        // backward.cu:421-473
        float dD[NS];
        const float dval = entry_depth<NS>(a, b, alpha, lv, dD);
        last_dT = last_dT + p.dL_dD * dval;
        const float f = p.dL_dD * T;
        if constexpr (NS == 1)
        {
#pragma unroll
            for (int i = 0; i < 8; ++i) dg[i] = dg[i] + (f * dD[0]) * dI[i];
        }
        else
        {
#pragma unroll
            for (int i = 0; i < 8; ++i)
            {
                float s2 = (f * dD[0]) * edI[0][i];
#pragma unroll
                for (int m = 1; m < NS; ++m) s2 = s2 + (f * dD[m]) * edI[m][i];
                dg[i] = dg[i] + s2;
            }
        }
    }
    if (k.flags & EXT_ASCENDING)
    {
        // backward.cu:476-498: penalise density decreasing along the ray
        // (weights and the segment ends held fixed)
        float qa[3], qb[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            qa[i] = (__builtin_fmaf(a, ray.rd[i], ray.ro[i]) - cm[i]) * linv;
            qb[i] = (__builtin_fmaf(b, ray.rd[i], ray.ro[i]) - cm[i]) * linv;
        }
        float wa[8], wb[8];
        const float da = density_at(g, qa, wa);
        const float db = density_at(g, qb, wb);
        const float rw = k.w_asc * pt_w * (da > db ? 1.f : 0.f);
#pragma unroll
        for (int i = 0; i < 8; ++i) dg[i] = dg[i] + rw * (wa[i] - wb[i]);
    }
    float dr0 = pt_w * dp0, dr1 = pt_w * dp1, dr2 = pt_w * dp2;
    if (k.flags & EXT_R_CONCEN)
    {
        // backward.cu:500-510: colour concentration toward the target colour
        dr0 = dr0 + k.w_R * pt_w * 2.f * (col[0] - p.gt0);
        dr1 = dr1 + k.w_R * pt_w * 2.f * (col[1] - p.gt1);
        dr2 = dr2 + k.w_R * pt_w * 2.f * (col[2] - p.gt2);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) gq[i] = dg[i];
    gq[8] = dr0;
    gq[9] = dr1;
    gq[10] = dr2;
    gq[11] = fabsf(dalpha * alpha);
    return true;
}

// A pixel of the adjoint with the optional terms (the per-hit work of
// bwd_hit_ext; traversal, sync slots and row writes are adjoint_window's).
template <int NS, bool EXACT>
struct ExtHit
{
    PixRay ray;
    ExtCoef k;
    ExtPix p;
    float T, last_dT, dp0, dp1, dp2;

    // backward.cu:160-211: state after the pixel's last contributor
    __device__ inline uint32_t init(bool inside, size_t pid, size_t HW, float bg, const float* __restrict__ out_T,
                                    const uint32_t* __restrict__ n_contrib, const float* __restrict__ dL_dcolor,
                                    const float* __restrict__ dL_dT, const float* __restrict__ dL_ddepth,
                                    const float* __restrict__ dL_dnormal, const float* __restrict__ gt,
                                    const float* __restrict__ out_depth)
    {
        T = 0.f; dp0 = 0.f; dp1 = 0.f; dp2 = 0.f; last_dT = 0.f;
        p = ExtPix{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (!inside) return 0u;
        T = out_T[pid];
        dp0 = dL_dcolor[pid];
        dp1 = dL_dcolor[HW + pid];
        dp2 = dL_dcolor[2 * HW + pid];
        last_dT = dL_dT[pid] + bg * (dp0 + dp1 + dp2);
        if (k.flags & EXT_DEPTH) p.dL_dD = dL_ddepth[pid] * ray.rd_norm_inv;
        if (k.flags & EXT_NORMAL)
        {
            p.dN0 = dL_dnormal[pid];
            p.dN1 = dL_dnormal[HW + pid];
            p.dN2 = dL_dnormal[2 * HW + pid];
        }
        if (k.flags & EXT_R_CONCEN)
        {
            p.gt0 = gt[pid];
            p.gt1 = gt[HW + pid];
            p.gt2 = gt[2 * HW + pid];
        }
        if (k.flags & EXT_DIST)
        {
            p.prefix_wm = out_depth[HW + pid];
            p.prefix_w = 1.f - T;
        }
        return n_contrib[pid];
    }
    __device__ inline bool operator()(const EntryPay& pe, float* gq, bool active)
    {
        if (!active) return false;
        return bwd_hit_ext<NS, EXACT>(pe, ray, k, p, T, last_dT, dp0, dp1, dp2, gq);
    }
};

template <int NS, bool EXACT>
__global__ __launch_bounds__(256) void render_bwd_ext_kernel(
    int W, int H, int gx, const uint32_t* __restrict__ tile_order, float tanfovx, float tanfovy, float cx, float cy,
    const float* __restrict__ c2w, float bg, ExtCoef k, const uint2* __restrict__ ranges,
    const uint32_t* __restrict__ vox_list, const float* __restrict__ centers, const float* __restrict__ lengths,
    const float* __restrict__ geos, const float* __restrict__ rgbs, const float* __restrict__ out_T,
    const uint32_t* __restrict__ n_contrib, const float* __restrict__ dL_dcolor, const float* __restrict__ dL_dT,
    const float* __restrict__ dL_ddepth, const float* __restrict__ dL_dnormal, const float* __restrict__ gt,
    const float* __restrict__ out_depth, const uint32_t* __restrict__ orig, const uint64_t* __restrict__ hit_masks,
    const uint32_t* __restrict__ tile_last, float* __restrict__ partial, const uint32_t* __restrict__ exact_ctr,
    const uint32_t* __restrict__ exact_list)
{
    __shared__ WindowLDS<BWD_WINDOW> sh;
    // tiles split between the two launches as in render_fwd_ext_kernel (the
    // forward of the frame built exact_list)
    const uint32_t count = EXACT ? exact_ctr[0] : 1u;
    for (uint32_t it = EXACT ? blockIdx.x : 0u; it < count; it += EXACT ? gridDim.x : 1u)
    {
        const int tile = EXACT ? (int)exact_list[it] : (int)tile_order[blockIdx.x];
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int tx = tile % gx, ty = tile / gx;
        const int px = tx * TILE_X + (w & 1) * 8 + (lane & 7), py = ty * TILE_Y + (w >> 1) * 8 + (lane >> 3);
        const bool inside = px < W && py < H;
        ExtHit<NS, EXACT> hit;
        hit.ray = make_ray(px, py, W, H, tanfovx, tanfovy, cx, cy, c2w);
        hit.k = k;
        const uint2 range = ranges[tile];
        const size_t HW = (size_t)H * W;
        const size_t pid = inside ? (size_t)py * W + px : 0;
        if (!EXACT && __syncthreads_or(!ray_inv_finite(hit.ray)) != 0) return;
        const uint32_t C = tile_chunks(range, tile_last[tile]);
        const uint64_t* masks = hit_masks + (((size_t)(range.x >> 6) + tile) * NWAVE + w) * 64 + lane;
        adjoint_window<BWD_SYNC_MIN_MASKED, Tile16, BWD_WINDOW>(
            sh, w, lane, C,
            [&]() {
                return hit.init(inside, pid, HW, bg, out_T, n_contrib, dL_dcolor, dL_dT, dL_ddepth, dL_dnormal, gt,
                                out_depth);
            },
            hit, range, masks, vox_list, centers, lengths, geos, rgbs, orig, partial);
        __syncthreads();
    }
}

template <int NS>
int launch_fwd_ext(int W, int H, float tanfovx, float tanfovy, float cx, float cy, const float* c2w, float bg,
                   const GeomState& g, const BinningState& b, const ImageState& img, const float* centers,
                   const float* lengths, const float* geos, const float* rgbs, int flags, float* out_color,
                   float* out_T, float* out_depth, float* out_normal, float* max_w, hipStream_t stream)
{
    const int gx = tiles_x(W), gy = tiles_y(H);
    const int T = gx * gy;
    prof_begin(PROF_RENDER_FWD, stream);
    render_fwd_ext_kernel<NS, false><<<(unsigned)T, 256, 0, stream>>>(
        W, H, gx, img.order, tanfovx, tanfovy, cx, cy, c2w, bg, img.ranges, b.vox_list, g.bbox, g.oct, centers, lengths, geos,
        rgbs, flags, img.tile_last, img.n_contrib, out_color, out_T, out_depth, out_normal, max_w,
        img.hit_masks, img.ctr, img.exact_list);
    SVR_CHECK_LAUNCH();
    render_fwd_ext_kernel<NS, true><<<EXT_EXACT_GRID, 256, 0, stream>>>(
        W, H, gx, img.order, tanfovx, tanfovy, cx, cy, c2w, bg, img.ranges, b.vox_list, g.bbox, g.oct, centers, lengths, geos,
        rgbs, flags, img.tile_last, img.n_contrib, out_color, out_T, out_depth, out_normal, max_w,
        img.hit_masks, img.ctr, img.exact_list);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_RENDER_FWD, stream);
    return SVR_OK;
}

template <int NS>
int launch_bwd_ext(int W, int H, float tanfovx, float tanfovy, float cx, float cy, const float* c2w, float bg,
                   const ExtCoef& k, int64_t N, int64_t R, const GeomState& g, const BinningState& b,
                   const ImageState& img, const float* centers, const float* lengths, const float* geos,
                   const float* rgbs, const float* out_T, const float* dL_dcolor, const float* dL_dT,
                   const float* dL_ddepth, const float* dL_dnormal, const float* gt, const float* out_depth,
                   float* partial, float* dgeos, float* drgbs, float* dsub, hipStream_t stream)
{
    const int gx = tiles_x(W), gy = tiles_y(H);
    const int T = gx * gy;
    prof_begin(PROF_RENDER_BWD, stream);
    render_bwd_ext_kernel<NS, false><<<(unsigned)T, 256, 0, stream>>>(
        W, H, gx, img.order, tanfovx, tanfovy, cx, cy, c2w, bg, k, img.ranges, b.vox_list, centers, lengths, geos, rgbs,
        out_T, img.n_contrib, dL_dcolor, dL_dT, dL_ddepth, dL_dnormal, gt, out_depth, b.orig, img.hit_masks,
        img.tile_last, partial, img.ctr, img.exact_list);
    SVR_CHECK_LAUNCH();
    render_bwd_ext_kernel<NS, true><<<EXT_EXACT_GRID, 256, 0, stream>>>(
        W, H, gx, img.order, tanfovx, tanfovy, cx, cy, c2w, bg, k, img.ranges, b.vox_list, centers, lengths, geos, rgbs,
        out_T, img.n_contrib, dL_dcolor, dL_dT, dL_ddepth, dL_dnormal, gt, out_depth, b.orig, img.hit_masks,
        img.tile_last, partial, img.ctr, img.exact_list);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_RENDER_BWD, stream);
    prof_begin(PROF_GRAD_REDUCE, stream);
    const int rc = grad_reduce_launch(N, g, partial, dgeos, drgbs, dsub, huge_rows(partial, R), stream);
    prof_end(PROF_GRAD_REDUCE, stream);
    return rc;
}

}  // namespace svr

using namespace svr;

extern "C" {

int svr_rasterize_forward_ext(int32_t n_samp, int32_t W, int32_t H, float tanfovx, float tanfovy, float cx, float cy,
                              const float* c2w, float bg, int32_t need_depth, int32_t need_distortion,
                              int32_t need_normal, int32_t track_max_w, int64_t N, const float* centers,
                              const float* lengths, const float* geos, const float* rgbs, const void* geom_state,
                              const void* order_table, int64_t R, void* binning_state, void* image_state,
                              float* out_color, float* out_T, float* out_depth, float* out_normal, float* max_w,
                              int32_t fwd_flags, svr_stream_t stream)
{
    SVR_REQUIRE(!(need_depth || need_distortion) || out_depth, SVR_ERR_INVALID,
                "svr_rasterize_forward_ext: need_depth / need_distortion without out_depth");
    SVR_REQUIRE(!need_normal || out_normal, SVR_ERR_INVALID, "svr_rasterize_forward_ext: need_normal without out_normal");
    SVR_REQUIRE(!track_max_w || N == 0 || max_w, SVR_ERR_INVALID,
                "svr_rasterize_forward_ext: track_max_w without max_w");
    GeomState g;
    BinningState b;
    ImageState img;
    int rc = forward_setup(n_samp, W, H, c2w, N, centers, lengths, geos, rgbs, geom_state, order_table, R,
                           binning_state, image_state, out_color, out_T, g, b, img, stream, fwd_flags);
    if (rc != SVR_OK) return rc;
    const int flags = (need_depth ? EXT_DEPTH : 0) | (need_distortion ? EXT_DIST : 0) |
                      (need_normal ? EXT_NORMAL : 0) | (track_max_w && N > 0 ? EXT_MAX_W : 0);
    if (flags & EXT_MAX_W)
        if (fill_u32(max_w, 0u, N, stream) != SVR_OK) return SVR_ERR_HIP;
    float* depth = (need_depth || need_distortion) ? out_depth : nullptr;
    switch (n_samp)
    {
    case 1: return launch_fwd_ext<1>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, g, b, img, centers, lengths, geos, rgbs,
                                     flags, out_color, out_T, depth, out_normal, max_w, stream);
    case 2: return launch_fwd_ext<2>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, g, b, img, centers, lengths, geos, rgbs,
                                     flags, out_color, out_T, depth, out_normal, max_w, stream);
    default: return launch_fwd_ext<3>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, g, b, img, centers, lengths, geos,
                                      rgbs, flags, out_color, out_T, depth, out_normal, max_w, stream);
    }
}

int svr_rasterize_backward_ext(int32_t n_samp, int32_t W, int32_t H, float tanfovx, float tanfovy, float cx, float cy,
                               const float* c2w, float bg, int64_t N, const float* centers, const float* lengths,
                               const float* geos, const float* rgbs, const void* geom_state, int64_t R,
                               const void* binning_state, const void* image_state, const float* out_T,
                               const float* dL_dcolor, const float* dL_ddepth, const float* dL_dnormal,
                               const float* dL_dT, float lambda_R_concen, const float* gt_color,
                               float lambda_ascending, float lambda_dist, int32_t need_depth, int32_t need_normal,
                               const float* out_depth, void* workspace, float* dgeos, float* drgbs, float* dsub,
                               svr_stream_t stream)
{
    SVR_REQUIRE(!need_depth || dL_ddepth, SVR_ERR_INVALID, "svr_rasterize_backward_ext: need_depth without dL_ddepth");
    SVR_REQUIRE(!need_normal || dL_dnormal, SVR_ERR_INVALID,
                "svr_rasterize_backward_ext: need_normal without dL_dnormal");
    SVR_REQUIRE(!(lambda_R_concen > 0.f) || gt_color, SVR_ERR_INVALID,
                "svr_rasterize_backward_ext: lambda_R_concen > 0 without gt_color");
    SVR_REQUIRE(!(lambda_dist > 0.f) || out_depth, SVR_ERR_INVALID,
                "svr_rasterize_backward_ext: lambda_dist > 0 without the forward's out_depth");
    GeomState g;
    BinningState b;
    ImageState img;
    const int rc = backward_setup(n_samp, W, H, c2w, N, geos, geom_state, R, binning_state, image_state, out_T,
                                  dL_dcolor, dL_dT, workspace, dgeos, drgbs, dsub, g, b, img, stream);
    if (rc != SVR_OK) return rc == 1 ? SVR_OK : rc;
    // backward.cu:654-772: the regularizer weights are lambda / (W * H)
    const float WH_inv = 1.f / (float)(W * H);
    ExtCoef k;
    k.flags = (need_depth ? EXT_DEPTH : 0) | (need_normal ? EXT_NORMAL : 0) | (lambda_dist > 0.f ? EXT_DIST : 0) |
              (lambda_R_concen > 0.f ? EXT_R_CONCEN : 0) | (lambda_ascending > 0.f ? EXT_ASCENDING : 0);
    k.w_R = lambda_R_concen * WH_inv;
    k.w_asc = lambda_ascending * WH_inv;
    k.w_dist = lambda_dist * WH_inv;
    float* partial = (float*)workspace;
    switch (n_samp)
    {
    case 1: return launch_bwd_ext<1>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, k, N, R, g, b, img, centers, lengths,
                                     geos, rgbs, out_T, dL_dcolor, dL_dT, dL_ddepth, dL_dnormal, gt_color, out_depth,
                                     partial, dgeos, drgbs, dsub, stream);
    case 2: return launch_bwd_ext<2>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, k, N, R, g, b, img, centers, lengths,
                                     geos, rgbs, out_T, dL_dcolor, dL_dT, dL_ddepth, dL_dnormal, gt_color, out_depth,
                                     partial, dgeos, drgbs, dsub, stream);
    default: return launch_bwd_ext<3>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, k, N, R, g, b, img, centers, lengths,
                                      geos, rgbs, out_T, dL_dcolor, dL_dT, dL_ddepth, dL_dnormal, gt_color,
                                      out_depth, partial, dgeos, drgbs, dsub, stream);
    }
}

}  // extern "C"
