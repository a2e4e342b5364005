// Per-voxel device helpers shared by the single-view and multi-view paths:
// the preprocess of one voxel for one camera, the tile rectangle of a packed
// bbox and the real spherical-harmonics basis.
//
// Reference: backends/new_cuda/src/preprocess.cu:24-132, auxiliary.h:150-170,
// sh_compute.cu:33-106.
#pragma once

#include "svr_device.h"

namespace svr {

// Widening of the projected octagon, in pixels: a constant part plus a part
// proportional to the magnitudes in play (projected coordinates and focal
// lengths), far above the float rounding of the projection and of the
// ray-box test, so that a pixel whose ray meets the box is never cut.
constexpr float OCT_MARGIN = 0.05f;
constexpr float OCT_MARGIN_REL = 4e-6f;

// preprocess.cu:24-132 for one voxel (centre c0..c2, half size r) seen by one
// camera: near-sphere cull, 8-corner projection, packed pixel bbox, quadrant
// bitset and the number of list entries (tiles x quadrants).  The arithmetic
// order matches the oracle exactly (no contraction), so the packed bboxes are
// bit-identical.
// TX x TY: the binning tile (16 x 16 for the reference's lists).
template <int TX = TILE_X, int TY = TILE_Y>
__device__ inline void preprocess_voxel(float c0, float c1, float c2, float r, int W, int H, int gx, int gy,
                                        float focal_x, float focal_y, float cx, float cy, const float* w2c,
                                        const float* c2w, float near_plane, uint32_t& nd_out, uint2& bb_out,
                                        uint32_t& bits_out, OctBox& oct_out)
{
    const float ro0 = c2w[3], ro1 = c2w[7], ro2 = c2w[11];
    const float* m = w2c;
    uint32_t nd = 0, bits = 0;
    uint2 bb = make_uint2(0u, 0u);
    // octagon of the projection in the pixel frame (pixel (px, py) centred
    // at (px, py)): bounds of x + y and x - y over the projected corners,
    // widened by OCT_MARGIN pixels; if a corner is behind the near plane (the
    // projection of the box is then not their hull) the candidates are the
    // bbox's pixels, as in the reference
    float os0 = 1e30f, os1 = -1e30f, od0 = 1e30f, od1 = -1e30f;
    float pmax = 0.f, ox0 = 1e30f, ox1 = -1e30f, oy0 = 1e30f, oy1 = -1e30f;
    bool all_front = true;
    const float cxh_o = cx - 0.5f, cyh_o = cy - 0.5f;
    const float rel0 = c0 - ro0, rel1 = c1 - ro1, rel2 = c2 - ro2;
    if (!(rel0 * rel0 + rel1 * rel1 + rel2 * rel2 < near_plane * near_plane))
    {
        float umin = 1e9f, vmin = 1e9f, umax = -1e9f, vmax = -1e9f;
#pragma unroll
        for (int i = 0; i < 8; ++i)
        {
            const float sx = (float)(((i & 4) >> 2) * 2 - 1);
            const float sy = (float)(((i & 2) >> 1) * 2 - 1);
            const float sz = (float)((i & 1) * 2 - 1);
            const float p0 = c0 + r * sx, p1 = c1 + r * sy, p2 = c2 + r * sz;
            const float camx = m[0] * p0 + m[1] * p1 + m[2] * p2 + m[3];
            const float camy = m[4] * p0 + m[5] * p1 + m[6] * p2 + m[7];
            const float camz = m[8] * p0 + m[9] * p1 + m[10] * p2 + m[11];
            if (camz < near_plane)
            {
                all_front = false;
                continue;
            }
            const float iz = 1.0f / camz;
            const float u = camx * iz, vv = camy * iz;
            {
                const float up = __builtin_fmaf(focal_x, u, cxh_o), vp = __builtin_fmaf(focal_y, vv, cyh_o);
                os0 = fminf(os0, up + vp); os1 = fmaxf(os1, up + vp);
                od0 = fminf(od0, up - vp); od1 = fmaxf(od1, up - vp);
                pmax = fmaxf(pmax, fabsf(up) + fabsf(vp));
                ox0 = fminf(ox0, up); ox1 = fmaxf(ox1, up);
                oy0 = fminf(oy0, vp); oy1 = fmaxf(oy1, vp);
            }
            const uint32_t q = ((p0 < ro0) << 2) | ((p1 < ro1) << 1) | (p2 < ro2);
            umin = smin(umin, u); vmin = smin(vmin, vv);
            umax = smax(umax, u); vmax = smax(vmax, vv);
            bits |= 1u << q;
        }
        const float cxh = cx - 0.5f, cyh = cy - 0.5f;
        const float bminx = fmaxf(focal_x * umin + cxh, 0.0f);
        const float bminy = fmaxf(focal_y * vmin + cyh, 0.0f);
        const float bmaxx = fminf(focal_x * umax + cxh, (float)W);
        const float bmaxy = fminf(focal_y * vmax + cyh, (float)H);
        if (!(bminx > bmaxx || bminy > bmaxy))
        {
            const uint32_t xmin = (uint32_t)(int)__builtin_rintf(bminx);
            const uint32_t ymin = (uint32_t)(int)__builtin_rintf(bminy);
            const uint32_t xmax = (uint32_t)(int)__builtin_rintf(bmaxx);
            const uint32_t ymax = (uint32_t)(int)__builtin_rintf(bmaxy);
            bb = make_uint2((xmin << 16) | ymin, (xmax << 16) | ymax);
            // auxiliary.h:150-170
            const int txmin = min((int)(xmin / TX), gx - 1), tymin = min((int)(ymin / TY), gy - 1);
            const int txmax = min((int)(xmax / TX), gx - 1), tymax = min((int)(ymax / TY), gy - 1);
            nd = (uint32_t)((1 + tymax - tymin) * (1 + txmax - txmin)) * (uint32_t)__popc(bits);
        }
        else
        {
            bits = 0;
        }
    }
    else
    {
        bits = 0;
    }
    nd_out = nd;
    bb_out = bb;
    bits_out = bits;
    const float mg = OCT_MARGIN + OCT_MARGIN_REL * (pmax + focal_x + focal_y);
    if (all_front && nd > 0)
    {
        const auto px16 = [](float x) { return (uint32_t)fminf(fmaxf(x, 0.f), 65535.f); };
        oct_out.o = make_float4(os0 - mg, os1 + mg, od0 - mg, od1 + mg);
        // the pixel box, kept inside the bbox: the candidates never leave the
        // reference's pix_in_bbox set, whatever the margin
        const uint32_t x0 = max(px16(ceilf(ox0 - mg)), bb.x >> 16), y0 = max(px16(ceilf(oy0 - mg)), bb.x & 0xFFFFu);
        const uint32_t x1 = min(px16(floorf(ox1 + mg)), bb.y >> 16), y1 = min(px16(floorf(oy1 + mg)), bb.y & 0xFFFFu);
        oct_out.tb = make_uint2((x0 << 16) | y0, (x1 << 16) | y1);
    }
    else
    {
        // a corner behind the near plane: the projection is not the hull of
        // the projected corners, so only the reference's own per-pixel test
        // applies, the packed bbox of the corners in front (forward.cu
        // pix_in_bbox); the octagon does not cut
        oct_out.o = make_float4(-1e30f, 1e30f, -1e30f, 1e30f);
        oct_out.tb = bb;
    }
    oct_out.pad = make_uint2(0u, 0u);
}

// Tile rectangle of a packed bbox (auxiliary.h:150-170) for TX x TY tiles.
template <int TX = TILE_X, int TY = TILE_Y>
__device__ inline void rect_of(uint2 bb, int gx, int gy, int& txmin, int& tymin, int& txmax, int& tymax)
{
    const uint32_t xmin = bb.x >> 16, ymin = bb.x & 0xFFFFu, xmax = bb.y >> 16, ymax = bb.y & 0xFFFFu;
    txmin = min((int)(xmin / TX), gx - 1);
    tymin = min((int)(ymin / TY), gy - 1);
    txmax = min((int)(xmax / TX), gx - 1);
    tymax = min((int)(ymax / TY), gy - 1);
}

// Tiles along an image side of n pixels for tiles of t pixels.
__host__ __device__ inline int tiles_along(int n, int t) { return (n + t - 1) / t; }

// sh_compute.cu:33-50 (real SH normalisations with their signs)
constexpr float SH_C2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                            -1.0925484305920792f, 0.5462742152960396f};
constexpr float SH_C3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                            0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                            -0.5900435899266435f};
constexpr float SH_C0 = 0.28209479177387814f;
constexpr float SH_C1 = 0.4886025119029199f;

__device__ inline int sh_basis(int deg, float x, float y, float z, float* b)
{
    b[0] = SH_C0;
    if (deg < 1) return 1;
    b[1] = -SH_C1 * y; b[2] = SH_C1 * z; b[3] = -SH_C1 * x;
    if (deg < 2) return 4;
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    b[4] = SH_C2[0] * xy;
    b[5] = SH_C2[1] * yz;
    b[6] = SH_C2[2] * (2.f * zz - xx - yy);
    b[7] = SH_C2[3] * xz;
    b[8] = SH_C2[4] * (xx - yy);
    if (deg < 3) return 9;
    b[9] = SH_C3[0] * y * (3.f * xx - yy);
    b[10] = SH_C3[1] * xy * z;
    b[11] = SH_C3[2] * y * (4.f * zz - xx - yy);
    b[12] = SH_C3[3] * z * (2.f * zz - 3.f * xx - 3.f * yy);
    b[13] = SH_C3[4] * x * (4.f * zz - xx - yy);
    b[14] = SH_C3[5] * z * (xx - yy);
    b[15] = SH_C3[6] * x * (xx - 3.f * yy);
    return 16;
}

// Linear staging of a block's 64 coefficient rows (degree 3): s0 holds the
// 64 x 3 sh0 values, s1 the 64 x 45 shs values, row k at 3 k and 45 k (odd
// strides: lane k touching row k is bank-conflict free), so that a block's
// contiguous spans move between global memory and LDS as 16-byte copies.
struct ShStage
{
    alignas(16) float s0[64 * 3 + 4];
    alignas(16) float s1[64 * 45 + 4];
};

// Loads the coefficient rows of voxels v0 .. v0 + nv - 1 into st: 16-byte
// loads (all in flight before the LDS stores) for a full block of 64 with
// 16-byte aligned arrays, single loads otherwise.  One wave.  In the same
// batch of loads lane l (l < nv) reads the centre of voxel v0 + l into c (so
// that no second memory round trip follows) and, with `word`, the voxel's
// word[word_stride l] (the first word of its rgb row with stride 3, its entry
// count with stride 1), which it returns.
__device__ inline uint32_t sh_stage_rows(ShStage& st, const float* __restrict__ sh0, const float* __restrict__ shs,
                                         const float* __restrict__ centers, int64_t v0, int nv, int lane, float* c,
                                         const uint32_t* __restrict__ word = nullptr, int word_stride = 3)
{
    uint32_t wv = 0u;
    const int64_t vc = v0 + min(lane, nv - 1);
    c[0] = centers[3 * vc + 0];
    c[1] = centers[3 * vc + 1];
    c[2] = centers[3 * vc + 2];
    if (nv == 64 && ((((uintptr_t)sh0) | ((uintptr_t)shs)) & 15u) == 0)
    {
        // unconditional loads (indices clamped into the span), so that the
        // compiler issues all of them before the first wait; named registers,
        // not an array, so that nothing goes through private memory
        const float4* i1 = reinterpret_cast<const float4*>(shs + 45 * v0);
#define SVR_SH_LD(j) const float4 r##j = i1[min(lane + 64 * j, 719)];
        SVR_SH_LD(0) SVR_SH_LD(1) SVR_SH_LD(2) SVR_SH_LD(3) SVR_SH_LD(4) SVR_SH_LD(5)
        SVR_SH_LD(6) SVR_SH_LD(7) SVR_SH_LD(8) SVR_SH_LD(9) SVR_SH_LD(10) SVR_SH_LD(11)
#undef SVR_SH_LD
        const float4 q0 = reinterpret_cast<const float4*>(sh0 + 3 * v0)[min(lane, 47)];
        if (word) wv = word[word_stride * lane];
        __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the stores
        // 720 = 11 x 64 + 16: eleven full rounds and a 16-lane tail
        float4* s1 = reinterpret_cast<float4*>(st.s1);
#define SVR_SH_ST(j) s1[lane + 64 * j] = r##j;
        SVR_SH_ST(0) SVR_SH_ST(1) SVR_SH_ST(2) SVR_SH_ST(3) SVR_SH_ST(4) SVR_SH_ST(5)
        SVR_SH_ST(6) SVR_SH_ST(7) SVR_SH_ST(8) SVR_SH_ST(9) SVR_SH_ST(10)
#undef SVR_SH_ST
        if (lane < 16) s1[lane + 704] = r11;
        if (lane < 48) reinterpret_cast<float4*>(st.s0)[lane] = q0;
    }
    else
    {
        if (word && lane < nv) wv = word[word_stride * lane];
        for (int e = lane; e < nv * 45; e += 64) st.s1[e] = shs[45 * v0 + e];
        for (int e = lane; e < nv * 3; e += 64) st.s0[e] = sh0[3 * v0 + e];
    }
    return wv;
}

// rgb of one voxel from its staged rows (sh_compute.cu:55-106 arithmetic).
__device__ inline void sh_eval_rows(int nb, const float* b, float x, float y, float z, const float* r0,
                                    const float* r1, float* res)
{
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
    {
        float a = SH_C0 * r0[ch];
        if (nb > 1) a = a - (SH_C1 * y) * r1[ch] + (SH_C1 * z) * r1[3 + ch] - (SH_C1 * x) * r1[6 + ch];
        for (int k = 4; k < nb; ++k) a = a + b[k] * r1[3 * (k - 1) + ch];
        a = a + 0.5f;
        res[ch] = a > 0.f ? a : 0.f;
    }
}

__device__ inline void sh_dir_c(const float* c, const float* __restrict__ cam, float& x, float& y, float& z)
{
    const float dx = c[0] - cam[0], dy = c[1] - cam[1], dz = c[2] - cam[2];
    const float rn = 1.f / __builtin_sqrtf(dx * dx + dy * dy + dz * dz);
    x = dx * rn; y = dy * rn; z = dz * rn;
}

__device__ inline void sh_dir(const float* __restrict__ centers, int64_t v, const float* __restrict__ cam, float& x,
                              float& y, float& z)
{
    const float dx = centers[3 * v] - cam[0], dy = centers[3 * v + 1] - cam[1], dz = centers[3 * v + 2] - cam[2];
    const float rn = 1.f / __builtin_sqrtf(dx * dx + dy * dy + dz * dz);
    x = dx * rn; y = dy * rn; z = dz * rn;
}

}  // namespace svr
