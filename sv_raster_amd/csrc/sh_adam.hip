// View-dependent colour (real SH, degree <= 3), the Adam update and the octree
// codec utilities.
//
// Reference: backends/new_cuda/src/sh_compute.cu:33-299, adam_step.cu:21-164,
// utils.cu:106-189.
#include <math.h>

#include "voxel_ops.h"
#include "svr_host.h"

namespace svr {

// sh_compute.cu:55-106, 180-208
__global__ __launch_bounds__(256) void sh_fwd_kernel(int deg, int64_t cnt, const int64_t* __restrict__ idx,
                                                     const float* __restrict__ centers, const float* __restrict__ cam,
                                                     const float* __restrict__ sh0, const float* __restrict__ shs,
                                                     int M, float* __restrict__ rgbs)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int64_t v = idx ? idx[t] : t;
    float x, y, z, b[16];
    sh_dir(centers, v, cam, x, y, z);
    const int nb = sh_basis(deg, x, y, z, b);
    const float* s = shs + (size_t)v * (M - 1) * 3;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
    {
        float res = SH_C0 * sh0[3 * v + ch];
        if (nb > 1) res = res - (SH_C1 * y) * s[ch] + (SH_C1 * z) * s[3 + ch] - (SH_C1 * x) * s[6 + ch];
        for (int k = 4; k < nb; ++k) res = res + b[k] * s[(k - 1) * 3 + ch];
        res = res + 0.5f;
        rgbs[3 * v + ch] = res > 0.f ? res : 0.f;
    }
}

// sh_compute.cu:110-176, 212-240: no clamp mask on the incoming gradient.
__global__ __launch_bounds__(256) void sh_bwd_kernel(int deg, int64_t cnt, const int64_t* __restrict__ idx,
                                                     const float* __restrict__ centers, const float* __restrict__ cam,
                                                     const float* __restrict__ dL_drgbs, int M,
                                                     float* __restrict__ dsh0, float* __restrict__ dshs)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int64_t v = idx ? idx[t] : t;
    float x, y, z, b[16];
    sh_dir(centers, v, cam, x, y, z);
    const int nb = sh_basis(deg, x, y, z, b);
    float* d = dshs + (size_t)v * (M - 1) * 3;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
    {
        const float g = dL_drgbs[3 * v + ch];
        dsh0[3 * v + ch] = SH_C0 * g;
        for (int k = 1; k < nb; ++k) d[(k - 1) * 3 + ch] = b[k] * g;
    }
}

// Coalesced variants for the full degree-3 layout (M = 16): one wave stages
// the 48 coefficients (sh0 + shs) of its 64 voxels through LDS with
// contiguous loads (rows padded to 49 floats so that lane k reading row k is
// bank-conflict free), then each lane evaluates its own voxel.  Same
// arithmetic as sh_fwd_kernel / sh_bwd_kernel.
constexpr int SH_F = 48;
constexpr int SH_PAD = 49;


__global__ __launch_bounds__(64) void sh_fwd16_kernel(int deg, int64_t cnt, const int64_t* __restrict__ idx,
                                                      const float* __restrict__ centers,
                                                      const float* __restrict__ cam, const float* __restrict__ sh0,
                                                      const float* __restrict__ shs, float* __restrict__ rgbs)
{
    __shared__ float buf[64 * SH_PAD];
    __shared__ int64_t sv[64];
    const int lane = threadIdx.x;
    const int64_t t0 = (int64_t)blockIdx.x * 64;
    const int nv = (int)min((int64_t)64, cnt - t0);
    int64_t myv = 0;
    if (lane < nv) myv = idx ? idx[t0 + lane] : t0 + lane;
    sv[lane] = myv;
    wave_sync();
    // element e = lane + 64 j of the block's 64 x 48 coefficients: all 48
    // loads are issued before the LDS stores, so they are in flight together
    float r[SH_F];
#pragma unroll
    for (int j = 0; j < SH_F; ++j)
    {
        const int e = lane + 64 * j;
        const int k = e / SH_F, f = e - k * SH_F;
        float x = 0.f;
        if (k < nv)
        {
            const int64_t v = sv[k];
            x = (f < 3) ? sh0[3 * v + f] : shs[v * 45 + (f - 3)];
        }
        r[j] = x;
    }
#pragma unroll
    for (int j = 0; j < SH_F; ++j)
    {
        const int e = lane + 64 * j;
        const int k = e / SH_F, f = e - k * SH_F;
        buf[k * SH_PAD + f] = r[j];
    }
    wave_sync();
    if (lane >= nv) return;
    const int64_t v = myv;
    float x, y, z, b[16];
    sh_dir(centers, v, cam, x, y, z);
    const int nb = sh_basis(deg, x, y, z, b);
    const float* row = buf + lane * SH_PAD;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
    {
        float res = SH_C0 * row[ch];
        if (nb > 1) res = res - (SH_C1 * y) * row[3 + ch] + (SH_C1 * z) * row[6 + ch] - (SH_C1 * x) * row[9 + ch];
        for (int k = 4; k < nb; ++k) res = res + b[k] * row[3 * k + ch];
        res = res + 0.5f;
        rgbs[3 * v + ch] = res > 0.f ? res : 0.f;
    }
}

__global__ __launch_bounds__(64) void sh_bwd16_kernel(int deg, int64_t cnt, const int64_t* __restrict__ idx,
                                                      const float* __restrict__ centers,
                                                      const float* __restrict__ cam,
                                                      const float* __restrict__ dL_drgbs, float* __restrict__ dsh0,
                                                      float* __restrict__ dshs)
{
    __shared__ float buf[64 * SH_PAD];
    __shared__ int64_t sv[64];
    const int lane = threadIdx.x;
    const int64_t t0 = (int64_t)blockIdx.x * 64;
    const int nv = (int)min((int64_t)64, cnt - t0);
    int64_t myv = 0;
    if (lane < nv) myv = idx ? idx[t0 + lane] : t0 + lane;
    sv[lane] = myv;
    if (lane < nv)
    {
        float x, y, z, b[16];
        sh_dir(centers, myv, cam, x, y, z);
        const int nb = sh_basis(deg, x, y, z, b);
        float* row = buf + lane * SH_PAD;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch)
        {
            const float g = dL_drgbs[3 * myv + ch];
            row[ch] = SH_C0 * g;
            for (int k = 1; k < 16; ++k) row[3 * k + ch] = (k < nb) ? b[k] * g : 0.f;
        }
    }
    wave_sync();
    for (int e = lane; e < nv * SH_F; e += 64)
    {
        const int k = e / SH_F, f = e - k * SH_F;
        const int64_t v = sv[k];
        const float val = buf[k * SH_PAD + f];
        if (f < 3) dsh0[3 * v + f] = val;
        else dshs[v * 45 + (f - 3)] = val;
    }
}

// The degree-3 backward over every voxel, in blocks of 64 consecutive voxels:
// rows of voxels listed in idx (marked by sh_mark_kernel with MARK_BITS in
// the first word of their dL_dsh0 row) get the SH gradient, the others zeros,
// and each block writes its contiguous 64 x 3 and 64 x 45 output spans with
// 16-byte stores.  The same values as sh_bwd16_kernel over idx after zeroing
// both outputs (rows listed twice are written once: the reference writes, it
// does not accumulate), without the separate 12 N + 180 N byte zero fill.
constexpr uint32_t MARK_BITS = 0x7FC0DEADu;  // a NaN payload no gradient or colour produces

__global__ __launch_bounds__(256) void sh_mark_kernel(int64_t cnt, const int64_t* __restrict__ idx,
                                                      float* __restrict__ dsh0)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) reinterpret_cast<uint32_t*>(dsh0)[3 * idx[i]] = MARK_BITS;
}

// Which rows a dense SH pass evaluates: every voxel, the rows sh_mark_kernel
// marked, or the voxels with entries in a preprocess geometry state (n_dup > 0:
// exactly the in-frustum list svr_visible_voxels returns).
enum ShList
{
    SH_LIST_ALL = 0,
    SH_LIST_MARKS = 1,
    SH_LIST_NDUP = 2,
};

// The degree-3 forward over every voxel, in blocks of 64 consecutive voxels
// (for lists that cover most of the voxels): each block loads its contiguous
// 64 x 45 and 64 x 3 coefficient spans with 16-byte loads (all in flight
// before the LDS stores), each lane evaluates its voxel if it is listed (LIST)
// and the block writes its 64 x 3 output span with 16-byte stores, zeros for
// the rows not listed.  Same arithmetic as sh_fwd16_kernel.
template <int LIST>
__global__ __launch_bounds__(64) void sh_fwd16_dense_kernel(int deg, int64_t N, const float* __restrict__ centers,
                                                            const float* __restrict__ cam,
                                                            const float* __restrict__ sh0,
                                                            const float* __restrict__ shs, float* __restrict__ rgbs,
                                                            const uint32_t* __restrict__ n_dup)
{
    __shared__ ShStage st;
    __shared__ alignas(16) float out[64 * 3 + 4];
    const int lane = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.x * 64;
    const int nv = (int)min((int64_t)64, N - v0);
    const int64_t v = v0 + lane;
    // the listed words (the marks in the first word of each rgbs row, or the
    // entry counts) and the coefficients in one batch of loads
    float c[3];
    const uint32_t* word = LIST == SH_LIST_MARKS ? reinterpret_cast<const uint32_t*>(rgbs + 3 * v0)
                           : LIST == SH_LIST_NDUP ? n_dup + v0 : nullptr;
    const uint32_t mark = sh_stage_rows(st, sh0, shs, centers, v0, nv, lane, c, word, LIST == SH_LIST_NDUP ? 1 : 3);
    const bool listed = lane < nv && (LIST == SH_LIST_ALL || (LIST == SH_LIST_MARKS ? mark == MARK_BITS : mark != 0u));
    wave_sync();
    float res[3] = {0.f, 0.f, 0.f};
    if (listed)
    {
        float x, y, z, b[16];
        sh_dir_c(c, cam, x, y, z);
        const int nb = sh_basis(deg, x, y, z, b);
        sh_eval_rows(nb, b, x, y, z, st.s0 + 3 * lane, st.s1 + 45 * lane, res);
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) out[3 * lane + ch] = res[ch];
    wave_sync();
    if (nv == 64)
    {
        if (lane < 48) reinterpret_cast<float4*>(rgbs + 3 * v0)[lane] = reinterpret_cast<const float4*>(out)[lane];
    }
    else
    {
        for (int e = lane; e < nv * 3; e += 64) rgbs[3 * v0 + e] = out[e];
    }
}

// n_dup null: the rows sh_mark_kernel marked; otherwise the voxels with
// n_dup > 0 (SH_LIST_NDUP).
__global__ __launch_bounds__(64) void sh_bwd16_dense_kernel(int deg, int64_t N, const float* __restrict__ centers,
                                                            const float* __restrict__ cam,
                                                            const float* __restrict__ dL_drgbs,
                                                            float* __restrict__ dsh0, float* __restrict__ dshs,
                                                            const uint32_t* __restrict__ n_dup)
{
    __shared__ ShStage st;
    const int lane = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.x * 64;
    const int nv = (int)min((int64_t)64, N - v0);
    const int64_t v = v0 + lane;
    float* r0 = st.s0 + 3 * lane;
    float* r1 = st.s1 + 45 * lane;
    if (lane < nv)
    {
        const bool marked = n_dup ? n_dup[v] != 0u : reinterpret_cast<const uint32_t*>(dsh0)[3 * v] == MARK_BITS;
        if (marked)
        {
            float x, y, z, b[16];
            sh_dir(centers, v, cam, x, y, z);
            const int nb = sh_basis(deg, x, y, z, b);
#pragma unroll
            for (int ch = 0; ch < 3; ++ch)
            {
                const float g = dL_drgbs[3 * v + ch];
                r0[ch] = SH_C0 * g;
                for (int k = 1; k < 16; ++k) r1[3 * (k - 1) + ch] = (k < nb) ? b[k] * g : 0.f;
            }
        }
        else
        {
            for (int f = 0; f < 3; ++f) r0[f] = 0.f;
            for (int f = 0; f < 45; ++f) r1[f] = 0.f;
        }
    }
    wave_sync();
    if (nv == 64)
    {
        // 64 rows starting at a multiple of 64 voxels: 16-byte aligned spans
        float4* o0 = reinterpret_cast<float4*>(dsh0 + 3 * v0);
        for (int e = lane; e < 48; e += 64) o0[e] = reinterpret_cast<const float4*>(st.s0)[e];
        float4* o1 = reinterpret_cast<float4*>(dshs + 45 * v0);
        for (int e = lane; e < 720; e += 64) o1[e] = reinterpret_cast<const float4*>(st.s1)[e];
    }
    else
    {
        for (int e = lane; e < nv * 3; e += 64) dsh0[3 * v0 + e] = st.s0[e];
        for (int e = lane; e < nv * 45; e += 64) dshs[45 * v0 + e] = st.s1[e];
    }
}

// adam_step.cu:21-64 (unbiased) and 67-102 (biased); lerp(v0, v1, t) = fma(t, v1, fma(-t, v0, v0))
// With coef (the device-step form), step_size and rb2 come from coef[0..1].
inline unsigned grid_for(int64_t n, int64_t cap = 1 << 20)
{
    int64_t b = (n + 255) / 256;
    if (b > cap) b = cap;
    return (unsigned)(b < 1 ? 1 : b);
}

template <bool SPARSE, bool BIASED>
__device__ inline void adam_elem(float& p, float& m, float& v, float gv, float step_size, float beta1, float beta2,
                                 float rb2, float eps)
{
    if (SPARSE && gv == 0.0f) return;
    const float mv = __builtin_fmaf(beta1, m, __builtin_fmaf(-beta1, gv, gv));
    const float gg = gv * gv;
    const float vv = __builtin_fmaf(beta2, v, __builtin_fmaf(-beta2, gg, gg));
    float denom;
    if (BIASED)
        denom = __builtin_sqrtf(vv) + eps;
    else
        denom = __builtin_fmaf(__builtin_sqrtf(vv), rb2, eps);
    p -= step_size * (mv / denom);
    m = mv;
    v = vv;
}

template <bool SPARSE, bool BIASED>
__global__ __launch_bounds__(256) void adam_kernel(int64_t n, float* __restrict__ p, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, float step_size,
                                                   float beta1, float beta2, float rb2, float eps,
                                                   const float* __restrict__ coef)
{
    if (coef)
    {
        if (coef[2] != 0.f) return;
        step_size = coef[0];
        rb2 = coef[1];
    }
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        const float gv = g[i];
        if (SPARSE && gv == 0.0f) continue;
        float pv = p[i], mv = m[i], vv = v[i];
        adam_elem<false, BIASED>(pv, mv, vv, gv, step_size, beta1, beta2, rb2, eps);
        p[i] = pv;
        m[i] = mv;
        v[i] = vv;
    }
}

// Same update on 16-byte vectors (all four arrays 16-byte aligned): one
// float4 gradient load per thread and iteration; in the sparse form a vector
// of zero gradients is skipped whole, and inside a mixed vector the elements
// with a zero gradient keep their values (they are stored back unchanged).
// The last n % 4 elements are updated one per thread.
// 16-byte form: each thread takes ADAM_UNROLL vectors per pass (a grid
// stride apart), reads all their gradients, then the parameters and moments of
// those with a nonzero gradient, so that a wave keeps several round trips of
// loads in flight instead of one gradient load and one row load at a time.
constexpr int ADAM_UNROLL = 4;

// Parameters and moments are written once per step and not read again in it:
// non-temporal stores (3-4 % faster on the 1 M-voxel Adam than plain ones).
typedef float adam_v4f __attribute__((ext_vector_type(4)));
__device__ inline void adam_store(float4* dst, const float4& x)
{
    __builtin_nontemporal_store(*reinterpret_cast<const adam_v4f*>(&x), reinterpret_cast<adam_v4f*>(dst));
}

template <bool SPARSE, bool BIASED>
__global__ __launch_bounds__(256) void adam4_kernel(int64_t n, float* __restrict__ p, const float* __restrict__ g,
                                                    float* __restrict__ m, float* __restrict__ v, float step_size,
                                                    float beta1, float beta2, float rb2, float eps,
                                                    const float* __restrict__ coef)
{
    if (coef)
    {
        if (coef[2] != 0.f) return;
        step_size = coef[0];
        rb2 = coef[1];
    }
    const int64_t n4 = n >> 2;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float4* __restrict__ P = reinterpret_cast<float4*>(p);
    const float4* __restrict__ G = reinterpret_cast<const float4*>(g);
    float4* __restrict__ M = reinterpret_cast<float4*>(m);
    float4* __restrict__ V = reinterpret_cast<float4*>(v);
    for (int64_t base = gid; base < n4; base += ADAM_UNROLL * stride)
    {
        float4 g4[ADAM_UNROLL], p4[ADAM_UNROLL], m4[ADAM_UNROLL], v4[ADAM_UNROLL];
        bool on[ADAM_UNROLL];
#pragma unroll
        for (int u = 0; u < ADAM_UNROLL; ++u)
        {
            const int64_t i = base + u * stride;
            g4[u] = i < n4 ? G[i] : make_float4(0.f, 0.f, 0.f, 0.f);
            on[u] = i < n4;
        }
#pragma unroll
        for (int u = 0; u < ADAM_UNROLL; ++u)
        {
            if (SPARSE) on[u] = on[u] && !(g4[u].x == 0.f && g4[u].y == 0.f && g4[u].z == 0.f && g4[u].w == 0.f);
            if (on[u])
            {
                const int64_t i = base + u * stride;
                p4[u] = P[i];
                m4[u] = M[i];
                v4[u] = V[i];
            }
        }
#pragma unroll
        for (int u = 0; u < ADAM_UNROLL; ++u)
        {
            if (!on[u]) continue;
            const int64_t i = base + u * stride;
            adam_elem<SPARSE, BIASED>(p4[u].x, m4[u].x, v4[u].x, g4[u].x, step_size, beta1, beta2, rb2, eps);
            adam_elem<SPARSE, BIASED>(p4[u].y, m4[u].y, v4[u].y, g4[u].y, step_size, beta1, beta2, rb2, eps);
            adam_elem<SPARSE, BIASED>(p4[u].z, m4[u].z, v4[u].z, g4[u].z, step_size, beta1, beta2, rb2, eps);
            adam_elem<SPARSE, BIASED>(p4[u].w, m4[u].w, v4[u].w, g4[u].w, step_size, beta1, beta2, rb2, eps);
            adam_store(&P[i], p4[u]);
            adam_store(&M[i], m4[u]);
            adam_store(&V[i], v4[u]);
        }
    }
    const int64_t t = 4 * n4 + gid;
    if (t < n)
    {
        float pv = p[t], mv = m[t], vv = v[t];
        adam_elem<SPARSE, BIASED>(pv, mv, vv, g[t], step_size, beta1, beta2, rb2, eps);
        p[t] = pv;
        m[t] = mv;
        v[t] = vv;
    }
}

// Several parameter tensors in one launch (SparseAdam.step over its groups):
// workgroups [wg0[k], wg0[k + 1]) update tensor k with the body of
// adam4_kernel and the tensor's own step size, bias term, betas and eps.  All
// arrays 16-byte aligned (the caller sends other tensors through adam_launch).
constexpr int ADAM_MULTI_MAX = 8;
// below 2^31 vectors, so that a thread's indices (at most the tensor's vector
// count plus ADAM_UNROLL grid strides of at most 2^22) stay below 2^32
constexpr int64_t ADAM_MULTI_MAX_VECTORS = 1ll << 31;
struct AdamMulti
{
    int count;
    int64_t n[ADAM_MULTI_MAX];
    int64_t wg0[ADAM_MULTI_MAX + 1];
    float* p[ADAM_MULTI_MAX];
    const float* g[ADAM_MULTI_MAX];
    float* m[ADAM_MULTI_MAX];
    float* v[ADAM_MULTI_MAX];
    float step_size[ADAM_MULTI_MAX], rb2[ADAM_MULTI_MAX], beta1[ADAM_MULTI_MAX], beta2[ADAM_MULTI_MAX],
        eps[ADAM_MULTI_MAX];
};

template <bool SPARSE, bool BIASED>
__global__ __launch_bounds__(256) void adam4_multi_kernel(AdamMulti a)
{
    int k = 0;
    while (k + 1 < a.count && (int64_t)blockIdx.x >= a.wg0[k + 1]) ++k;
    const int64_t b = (int64_t)blockIdx.x - a.wg0[k];
    const int64_t n = a.n[k];
    const int64_t n4 = n >> 2;
    const int64_t stride = (a.wg0[k + 1] - a.wg0[k]) * 256;
    const int64_t gid = b * 256 + threadIdx.x;
    const float step_size = a.step_size[k], rb2 = a.rb2[k], beta1 = a.beta1[k], beta2 = a.beta2[k], eps = a.eps[k];
    float* __restrict__ p = a.p[k];
    const float* __restrict__ g = a.g[k];
    float* __restrict__ m = a.m[k];
    float* __restrict__ v = a.v[k];
    float4* __restrict__ P = reinterpret_cast<float4*>(p);
    const float4* __restrict__ G = reinterpret_cast<const float4*>(g);
    float4* __restrict__ M = reinterpret_cast<float4*>(m);
    float4* __restrict__ V = reinterpret_cast<float4*>(v);
    // 32-bit vector indices (fewer address registers: 83 VGPRs instead of 112,
    // five waves per SIMD instead of four; tensors of ADAM_MULTI_MAX_VECTORS
    // vectors or more go to adam4_kernel): 0.2456 -> 0.2416 ms at 1 M voxels
    const uint32_t n4u = (uint32_t)n4, strideu = (uint32_t)stride;
    for (uint32_t base = (uint32_t)gid; base < n4u; base += ADAM_UNROLL * strideu)
    {
        float4 g4[ADAM_UNROLL], p4[ADAM_UNROLL], m4[ADAM_UNROLL], v4[ADAM_UNROLL];
        bool on[ADAM_UNROLL];
#pragma unroll
        for (int u = 0; u < ADAM_UNROLL; ++u)
        {
            const uint32_t i = base + u * strideu;
            g4[u] = i < n4u ? G[i] : make_float4(0.f, 0.f, 0.f, 0.f);
            on[u] = i < n4u;
        }
#pragma unroll
        for (int u = 0; u < ADAM_UNROLL; ++u)
        {
            if (SPARSE) on[u] = on[u] && !(g4[u].x == 0.f && g4[u].y == 0.f && g4[u].z == 0.f && g4[u].w == 0.f);
            if (on[u])
            {
                const uint32_t i = base + u * strideu;
                p4[u] = P[i];
                m4[u] = M[i];
                v4[u] = V[i];
            }
        }
#pragma unroll
        for (int u = 0; u < ADAM_UNROLL; ++u)
        {
            if (!on[u]) continue;
            const uint32_t i = base + u * strideu;
            adam_elem<SPARSE, BIASED>(p4[u].x, m4[u].x, v4[u].x, g4[u].x, step_size, beta1, beta2, rb2, eps);
            adam_elem<SPARSE, BIASED>(p4[u].y, m4[u].y, v4[u].y, g4[u].y, step_size, beta1, beta2, rb2, eps);
            adam_elem<SPARSE, BIASED>(p4[u].z, m4[u].z, v4[u].z, g4[u].z, step_size, beta1, beta2, rb2, eps);
            adam_elem<SPARSE, BIASED>(p4[u].w, m4[u].w, v4[u].w, g4[u].w, step_size, beta1, beta2, rb2, eps);
            adam_store(&P[i], p4[u]);
            adam_store(&M[i], m4[u]);
            adam_store(&V[i], v4[u]);
        }
    }
    const int64_t t = 4 * n4 + gid;
    if (t < n)
    {
        float pv = p[t], mv = m[t], vv = v[t];
        adam_elem<SPARSE, BIASED>(pv, mv, vv, g[t], step_size, beta1, beta2, rb2, eps);
        p[t] = pv;
        m[t] = mv;
        v[t] = vv;
    }
}

template <bool BIASED>
void adam_launch(bool sparse, int64_t n, float* p, const float* g, float* m, float* v, float step_size, float beta1,
                 float beta2, float rb2, float eps, const float* coef, hipStream_t stream)
{
    const bool vec = ((((uintptr_t)p) | ((uintptr_t)g) | ((uintptr_t)m) | ((uintptr_t)v)) & 15u) == 0;
    if (vec)
    {
        const int64_t n4 = n >> 2;
        const unsigned grid = grid_for(n4 > 0 ? (n4 + ADAM_UNROLL - 1) / ADAM_UNROLL : 1, 16384);
        if (sparse)
            adam4_kernel<true, BIASED><<<grid, 256, 0, stream>>>(n, p, g, m, v, step_size, beta1, beta2, rb2, eps,
                                                                 coef);
        else
            adam4_kernel<false, BIASED><<<grid, 256, 0, stream>>>(n, p, g, m, v, step_size, beta1, beta2, rb2, eps,
                                                                  coef);
    }
    else
    {
        const unsigned grid = grid_for(n, 8192);
        if (sparse)
            adam_kernel<true, BIASED><<<grid, 256, 0, stream>>>(n, p, g, m, v, step_size, beta1, beta2, rb2, eps,
                                                                coef);
        else
            adam_kernel<false, BIASED><<<grid, 256, 0, stream>>>(n, p, g, m, v, step_size, beta1, beta2, rb2, eps,
                                                                 coef);
    }
}

// Device-step form of the bias terms (adam_step.cu:119-124): t = ++step and
// coef = {lr / (1 - beta1^t), 1 / sqrt(1 - beta2^t), skip}.  The terms come
// from a table the host filled with the host form (svr_adam_bias_terms), so
// they are bit-equal to the host-step update: table = {int32 first step,
// int32 count, then count (step_size, rb2) float pairs}.  A step outside the
// table falls back to pow() in double on the device (the caller refills the
// table before that happens).  *skip != 0 (a frame over its entry capacity)
// leaves the step count alone and sets coef[2], which makes the update kernels
// return without touching any element.
__global__ void adam_coef_kernel(int32_t* __restrict__ step, const int32_t* __restrict__ table,
                                 const uint32_t* __restrict__ skip, double lr, double beta1, double beta2,
                                 float* __restrict__ coef)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (skip && skip[0] != 0u)
    {
        coef[2] = 1.f;
        return;
    }
    const int32_t t = step[0] + 1;
    step[0] = t;
    coef[2] = 0.f;
    const int32_t i = table ? t - table[0] : -1;
    if (table && i >= 0 && i < table[1])
    {
        const float* terms = reinterpret_cast<const float*>(table + 2);
        coef[0] = terms[2 * i];
        coef[1] = terms[2 * i + 1];
        return;
    }
    const double bc1 = 1.0 - pow(beta1, (double)t);
    const double bc2 = 1.0 - pow(beta2, (double)t);
    coef[0] = (float)(lr / bc1);
    coef[1] = (float)(1.0 / sqrt(bc2));
}

__global__ void order_rank_kernel(int64_t n, const int64_t* __restrict__ paths, int64_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t p = (uint64_t)paths[i];
#pragma unroll
    for (uint32_t q = 0; q < 8; ++q) out[q * n + i] = (int64_t)(p ^ order_table(q));
}

__global__ void ijk2path_kernel(int64_t n, const int64_t* __restrict__ ijk, const int8_t* __restrict__ lv,
                                int64_t* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint64_t i = (uint64_t)ijk[3 * t], j = (uint64_t)ijk[3 * t + 1], k = (uint64_t)ijk[3 * t + 2];
    uint64_t path = 0;
    for (int l = 0; l < MAX_NUM_LEVELS; ++l)
        path |= ((((i >> l) & 1ull) << 2) | (((j >> l) & 1ull) << 1) | ((k >> l) & 1ull)) << (3 * l);
    out[t] = (int64_t)(path << (3 * (MAX_NUM_LEVELS - lv[t])));
}

__global__ void path2ijk_kernel(int64_t n, const int64_t* __restrict__ paths, const int8_t* __restrict__ lv,
                                int64_t* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint64_t path = (uint64_t)paths[t] >> (3 * (MAX_NUM_LEVELS - lv[t]));
    int64_t i = 0, j = 0, k = 0;
    for (int l = 0; l < MAX_NUM_LEVELS; ++l)
    {
        const uint64_t bits = (path >> (3 * l)) & 7ull;
        i |= (int64_t)((bits >> 2) & 1ull) << l;
        j |= (int64_t)((bits >> 1) & 1ull) << l;
        k |= (int64_t)(bits & 1ull) << l;
    }
    out[3 * t] = i; out[3 * t + 1] = j; out[3 * t + 2] = k;
}


}  // namespace svr

using namespace svr;

extern "C" {

int svr_sh_compute(int32_t deg, int64_t n_idx, const int64_t* idx, int64_t N, const float* centers,
                   const float* cam_pos, const float* sh0, const float* shs, int32_t M, float* rgbs,
                   svr_stream_t stream)
{
    SVR_REQUIRE(deg >= 0 && deg <= 3 && M >= (deg + 1) * (deg + 1), SVR_ERR_INVALID,
                "sh degree %d needs %d coefficients, got %d", deg, (deg + 1) * (deg + 1), M);
    if (N == 0) return SVR_OK;
    const bool aligned = ((((uintptr_t)sh0) | ((uintptr_t)shs) | ((uintptr_t)rgbs)) & 15u) == 0;
    if (M == 16 && aligned && (n_idx == 0 || 2 * n_idx >= N))
    {
        // a list covering at least half the voxels: every voxel's coefficients are read
        // contiguously, the listed rows marked first
        prof_begin(PROF_SH_FWD, stream);
        if (n_idx == 0)
        {
            sh_fwd16_dense_kernel<SH_LIST_ALL><<<(unsigned)((N + 63) / 64), 64, 0, stream>>>(deg, N, centers, cam_pos,
                                                                                             sh0, shs, rgbs, nullptr);
        }
        else
        {
            if (fill_u32(rgbs, 0u, 3 * N, stream) != SVR_OK) return SVR_ERR_HIP;
            sh_mark_kernel<<<(unsigned)((n_idx + 255) / 256), 256, 0, stream>>>(n_idx, idx, rgbs);
            SVR_CHECK_LAUNCH();
            sh_fwd16_dense_kernel<SH_LIST_MARKS><<<(unsigned)((N + 63) / 64), 64, 0, stream>>>(deg, N, centers, cam_pos,
                                                                                               sh0, shs, rgbs, nullptr);
        }
        SVR_CHECK_LAUNCH();
        prof_end(PROF_SH_FWD, stream);
        return SVR_OK;
    }
    if (fill_u32(rgbs, 0u, 3 * N, stream) != SVR_OK) return SVR_ERR_HIP;
    const int64_t cnt = n_idx ? n_idx : N;
    if (cnt == 0) return SVR_OK;
    prof_begin(PROF_SH_FWD, stream);
    if (M == 16)
        sh_fwd16_kernel<<<(unsigned)((cnt + 63) / 64), 64, 0, stream>>>(deg, cnt, n_idx ? idx : nullptr, centers,
                                                                         cam_pos, sh0, shs, rgbs);
    else
        sh_fwd_kernel<<<grid_for(cnt, 1ll << 30), 256, 0, stream>>>(deg, cnt, n_idx ? idx : nullptr, centers,
                                                                     cam_pos, sh0, shs, M, rgbs);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_SH_FWD, stream);
    return SVR_OK;
}

int svr_sh_compute_bw(int32_t deg, int32_t M, int64_t n_idx, const int64_t* idx, int64_t N, const float* centers,
                      const float* cam_pos, const float* dL_drgbs, float* dsh0, float* dshs, svr_stream_t stream)
{
    SVR_REQUIRE(deg >= 0 && deg <= 3 && M >= (deg + 1) * (deg + 1), SVR_ERR_INVALID,
                "sh degree %d needs %d coefficients, got %d", deg, (deg + 1) * (deg + 1), M);
    if (N == 0) return SVR_OK;
    if (M == 16 && n_idx > 0 && ((uintptr_t)dsh0 & 15u) == 0 && ((uintptr_t)dshs & 15u) == 0)
    {
        if (fill_u32(dsh0, 0u, 3 * N, stream) != SVR_OK) return SVR_ERR_HIP;
        prof_begin(PROF_SH_BWD, stream);
        sh_mark_kernel<<<(unsigned)((n_idx + 255) / 256), 256, 0, stream>>>(n_idx, idx, dsh0);
        SVR_CHECK_LAUNCH();
        sh_bwd16_dense_kernel<<<(unsigned)((N + 63) / 64), 64, 0, stream>>>(deg, N, centers, cam_pos, dL_drgbs, dsh0,
                                                                           dshs, nullptr);
        SVR_CHECK_LAUNCH();
        prof_end(PROF_SH_BWD, stream);
        return SVR_OK;
    }
    if (fill_u32(dsh0, 0u, 3 * N, stream) != SVR_OK) return SVR_ERR_HIP;
    if (M > 1 && fill_u32(dshs, 0u, 3 * N * (int64_t)(M - 1), stream) != SVR_OK) return SVR_ERR_HIP;
    const int64_t cnt = n_idx ? n_idx : N;
    if (cnt == 0) return SVR_OK;
    prof_begin(PROF_SH_BWD, stream);
    if (M == 16)
        sh_bwd16_kernel<<<(unsigned)((cnt + 63) / 64), 64, 0, stream>>>(deg, cnt, n_idx ? idx : nullptr, centers,
                                                                         cam_pos, dL_drgbs, dsh0, dshs);
    else
        sh_bwd_kernel<<<grid_for(cnt, 1ll << 30), 256, 0, stream>>>(deg, cnt, n_idx ? idx : nullptr, centers,
                                                                     cam_pos, dL_drgbs, M, dsh0, dshs);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_SH_BWD, stream);
    return SVR_OK;
}

// The listed voxels are those with entries in a preprocess geometry state
// (n_dup > 0), i.e. exactly svr_visible_voxels' list: no index list, no zero
// fill and no marking pass before the dense kernel.  Degree-3 layout
// (M == 16) and 16-byte aligned coefficient and output arrays only.
int svr_sh_compute_visible(int32_t deg, int64_t N, const void* geom_state, const float* centers, const float* cam_pos,
                           const float* sh0, const float* shs, int32_t M, float* rgbs, svr_stream_t stream)
{
    SVR_REQUIRE(deg >= 0 && deg <= 3 && M == 16, SVR_ERR_INVALID,
                "svr_sh_compute_visible: degree %d with %d coefficients (needs the 16-coefficient layout)", deg, M);
    SVR_REQUIRE(((((uintptr_t)sh0) | ((uintptr_t)shs) | ((uintptr_t)rgbs)) & 15u) == 0, SVR_ERR_INVALID,
                "svr_sh_compute_visible: coefficient and output arrays must be 16-byte aligned");
    if (N == 0) return SVR_OK;
    SVR_REQUIRE(geom_state, SVR_ERR_INVALID, "svr_sh_compute_visible: null geometry state");
    const GeomState g = GeomState::carve((void*)geom_state, N, nullptr);
    prof_begin(PROF_SH_FWD, stream);
    sh_fwd16_dense_kernel<SH_LIST_NDUP><<<(unsigned)((N + 63) / 64), 64, 0, stream>>>(deg, N, centers, cam_pos, sh0,
                                                                                      shs, rgbs, g.n_dup);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_SH_FWD, stream);
    return SVR_OK;
}

int svr_sh_compute_bw_visible(int32_t deg, int32_t M, int64_t N, const void* geom_state, const float* centers,
                              const float* cam_pos, const float* dL_drgbs, float* dsh0, float* dshs,
                              svr_stream_t stream)
{
    SVR_REQUIRE(deg >= 0 && deg <= 3 && M == 16, SVR_ERR_INVALID,
                "svr_sh_compute_bw_visible: degree %d with %d coefficients (needs the 16-coefficient layout)", deg, M);
    SVR_REQUIRE(((((uintptr_t)dsh0) | ((uintptr_t)dshs)) & 15u) == 0, SVR_ERR_INVALID,
                "svr_sh_compute_bw_visible: gradient arrays must be 16-byte aligned");
    if (N == 0) return SVR_OK;
    SVR_REQUIRE(geom_state, SVR_ERR_INVALID, "svr_sh_compute_bw_visible: null geometry state");
    const GeomState g = GeomState::carve((void*)geom_state, N, nullptr);
    prof_begin(PROF_SH_BWD, stream);
    sh_bwd16_dense_kernel<<<(unsigned)((N + 63) / 64), 64, 0, stream>>>(deg, N, centers, cam_pos, dL_drgbs, dsh0,
                                                                       dshs, g.n_dup);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_SH_BWD, stream);
    return SVR_OK;
}

int svr_adam_bias_terms(double step, double lr, double beta1, double beta2, float* step_size_host,
                        float* rb2_host)
{
    SVR_REQUIRE(step_size_host && rb2_host, SVR_ERR_INVALID, "svr_adam_bias_terms: null output");
    // adam_step.cu:119-124 (bias terms in double on the host)
    const double bc1 = 1.0 - pow(beta1, step);
    const double bc2 = 1.0 - pow(beta2, step);
    *step_size_host = (float)(lr / bc1);
    *rb2_host = (float)(1.0 / sqrt(bc2));
    return SVR_OK;
}

int svr_unbiased_adam_step(int32_t sparse, int64_t n, float* p, const float* g, float* m, float* v, double step,
                           double lr, double beta1, double beta2, float eps, svr_stream_t stream)
{
    if (n == 0) return SVR_OK;
    float step_size = 0.f, rb2 = 0.f;
    svr_adam_bias_terms(step, lr, beta1, beta2, &step_size, &rb2);
    prof_begin(PROF_ADAM, stream);
    adam_launch<false>(sparse != 0, n, p, g, m, v, step_size, (float)beta1, (float)beta2, rb2, eps, nullptr, stream);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_ADAM, stream);
    return SVR_OK;
}

int svr_unbiased_adam_step_device(int32_t sparse, int64_t n, float* p, const float* g, float* m, float* v,
                                  int32_t* step_dev, const void* bias_table, const uint32_t* skip_dev, float* coef_dev,
                                  double lr, double beta1, double beta2, float eps, svr_stream_t stream)
{
    SVR_REQUIRE(step_dev && coef_dev, SVR_ERR_INVALID, "svr_unbiased_adam_step_device: null step or coefficients");
    prof_begin(PROF_ADAM, stream);
    adam_coef_kernel<<<1, 64, 0, stream>>>(step_dev, (const int32_t*)bias_table, skip_dev, lr, beta1, beta2,
                                           coef_dev);
    SVR_CHECK_LAUNCH();
    if (n > 0)
    {
        adam_launch<false>(sparse != 0, n, p, g, m, v, 0.f, (float)beta1, (float)beta2, 0.f, eps, coef_dev, stream);
        SVR_CHECK_LAUNCH();
    }
    prof_end(PROF_ADAM, stream);
    return SVR_OK;
}

int svr_adam_step_multi(int32_t biased, int32_t sparse, int32_t count, const int64_t* numels, float* const* params,
                        const float* const* grads, float* const* exp_avgs, float* const* exp_avg_sqs,
                        const double* steps, const double* lrs, const double* beta1s, const double* beta2s,
                        const float* eps, svr_stream_t stream)
{
    SVR_REQUIRE(count >= 0, SVR_ERR_INVALID, "svr_adam_step_multi: negative tensor count %d", (int)count);
    if (count == 0) return SVR_OK;
    SVR_REQUIRE(numels && params && grads && exp_avgs && exp_avg_sqs && steps && lrs && beta1s && beta2s && eps,
                SVR_ERR_INVALID, "svr_adam_step_multi: null argument array");
    prof_begin(PROF_ADAM, stream);
    AdamMulti a{};
    auto flush = [&]() {
        if (a.count == 0) return;
        const unsigned grid = (unsigned)a.wg0[a.count];
        if (biased)
        {
            if (sparse) adam4_multi_kernel<true, true><<<grid, 256, 0, stream>>>(a);
            else adam4_multi_kernel<false, true><<<grid, 256, 0, stream>>>(a);
        }
        else
        {
            if (sparse) adam4_multi_kernel<true, false><<<grid, 256, 0, stream>>>(a);
            else adam4_multi_kernel<false, false><<<grid, 256, 0, stream>>>(a);
        }
        a = AdamMulti{};
    };
    for (int32_t i = 0; i < count; ++i)
    {
        const int64_t n = numels[i];
        if (n <= 0) continue;
        float* p = params[i];
        const float* g = grads[i];
        float* m = exp_avgs[i];
        float* v = exp_avg_sqs[i];
        SVR_REQUIRE(p && g && m && v, SVR_ERR_INVALID, "svr_adam_step_multi: null array of tensor %d", (int)i);
        float step_size = 0.f, rb2 = 1.f;
        if (biased)
            step_size = (float)lrs[i];
        else
            svr_adam_bias_terms(steps[i], lrs[i], beta1s[i], beta2s[i], &step_size, &rb2);
        const bool vec = ((((uintptr_t)p) | ((uintptr_t)g) | ((uintptr_t)m) | ((uintptr_t)v)) & 15u) == 0;
        // the multi-tensor kernel indexes vectors with 32 bits
        if (!vec || (n >> 2) >= ADAM_MULTI_MAX_VECTORS)
        {
            if (biased)
                adam_launch<true>(sparse != 0, n, p, g, m, v, step_size, (float)beta1s[i], (float)beta2s[i], rb2,
                                  eps[i], nullptr, stream);
            else
                adam_launch<false>(sparse != 0, n, p, g, m, v, step_size, (float)beta1s[i], (float)beta2s[i], rb2,
                                   eps[i], nullptr, stream);
            continue;
        }
        const int k = a.count;
        const int64_t n4 = n >> 2;
        a.n[k] = n;
        a.p[k] = p; a.g[k] = g; a.m[k] = m; a.v[k] = v;
        a.step_size[k] = step_size; a.rb2[k] = rb2;
        a.beta1[k] = (float)beta1s[i]; a.beta2[k] = (float)beta2s[i]; a.eps[k] = eps[i];
        a.wg0[k + 1] = a.wg0[k] + grid_for(n4 > 0 ? (n4 + ADAM_UNROLL - 1) / ADAM_UNROLL : 1, 16384);
        a.count = k + 1;
        if (a.count == ADAM_MULTI_MAX) flush();
    }
    flush();
    SVR_CHECK_LAUNCH();
    prof_end(PROF_ADAM, stream);
    return SVR_OK;
}

int svr_biased_adam_step(int32_t sparse, int64_t n, float* p, const float* g, float* m, float* v, float lr,
                         float beta1, float beta2, float eps, svr_stream_t stream)
{
    if (n == 0) return SVR_OK;
    prof_begin(PROF_ADAM, stream);
    adam_launch<true>(sparse != 0, n, p, g, m, v, lr, beta1, beta2, 1.f, eps, nullptr, stream);
    SVR_CHECK_LAUNCH();
    prof_end(PROF_ADAM, stream);
    return SVR_OK;
}

int svr_voxel_order_rank(int64_t n, const int64_t* paths, int64_t* out, svr_stream_t stream)
{
    if (n == 0) return SVR_OK;
    order_rank_kernel<<<grid_for(n, 1ll << 30), 256, 0, stream>>>(n, paths, out);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

int svr_ijk_2_octpath(int64_t n, const int64_t* ijk, const int8_t* lv, int64_t* out, svr_stream_t stream)
{
    if (n == 0) return SVR_OK;
    ijk2path_kernel<<<grid_for(n, 1ll << 30), 256, 0, stream>>>(n, ijk, lv, out);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

int svr_octpath_2_ijk(int64_t n, const int64_t* paths, const int8_t* lv, int64_t* out, svr_stream_t stream)
{
    if (n == 0) return SVR_OK;
    path2ijk_kernel<<<grid_for(n, 1ll << 30), 256, 0, stream>>>(n, paths, lv, out);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

}  // extern "C"
