// ORACLE — test infrastructure, not product code.
//
// A scalar CPU restatement of the sparse-voxel rasterizer hot path of the
// reference backend (backends/new_cuda).  It exists to check the HIP kernels
// and to give the CPU baseline timing; only tests/, the smoke check and the
// benchmark's cpu_baseline leg load it.  Nothing in sv_raster_amd/ links it.
//
// Every function cites the reference lines it restates.  Floating-point
// contract shared with the HIP kernels (both sides compile with
// -ffp-contract=off, so nothing is fused unless written as fma here):
//   * explicit fma at: trilinear density sum, sample point, colour
//     accumulation, backward dot / last_dL_dT update, Adam lerp;
//   * exp() is svr_ref_expf below (Cody-Waite reduction + degree-7 Taylor),
//     the same method and constants as the device exp, so that integer
//     outputs (bboxes, n_contrib, tile_last) are bit-identical;
//   * divisions and square roots are IEEE correctly rounded on both sides.
//
// Templated on the real type: float is the parity path, double is used by
// the finite-difference gradient checks.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// NC/src/config.h:14-34
constexpr int BLOCK_X = 16;
constexpr int BLOCK_Y = 16;
constexpr int MAX_NUM_LEVELS = 21;
constexpr float MAX_ALPHA = 0.99999f;
constexpr float MIN_ALPHA = 0.00001f;
constexpr float EARLY_STOP_T = 0.0001f;
constexpr float STEP_SZ_SCALE = 100.f;

// exp(x) by Cody & Waite argument reduction (x = k*ln2 + r, ln2 split in a
// high part with trailing zero bits and a low correction) followed by the
// degree-7 Taylor polynomial of e^r on |r| <= ln2/2 in Horner form, and an
// exact scaling by 2^k (ldexp, one rounding).  Error <= 2 ulp.
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

float svr_ref_expf(float x)
{
    if (x != x) return x;
    if (x > 88.72283935546875f) return u2f(0x7f800000u);
    if (x < -103.97208404541015625f) return 0.f;
    const float kf = std::rint(x * 1.44269502162933349609375f);
    float r = std::fma(-kf, 0.693145751953125f, x);
    r = std::fma(-kf, 1.428606765330187045e-06f, r);
    float p = 1.98412701976485550403594970703125e-04f;      // 1/5040
    p = std::fma(p, r, 1.388888922519981861114501953125e-03f); // 1/720
    p = std::fma(p, r, 8.33333376795053482055664062500e-03f);  // 1/120
    p = std::fma(p, r, 4.16666679084300994873046875000e-02f);  // 1/24
    p = std::fma(p, r, 1.66666671633720397949218750000e-01f);  // 1/6
    p = std::fma(p, r, 0.5f);
    p = std::fma(p, r, 1.0f);
    p = std::fma(p, r, 1.0f);
    // exact scaling by 2^k with a single rounding (also into the denormal range)
    return std::ldexp(p, (int)kf);
}

template <typename R> inline R rexp(R x);
template <> inline float rexp<float>(float x) { return svr_ref_expf(x); }
template <> inline double rexp<double>(double x) { return std::exp(x); }

template <typename R> inline R rfma(R a, R b, R c) { return std::fma(a, b, c); }

// NC/src/auxiliary.h:104-112: the ternary min/max (NaN-order sensitive)
template <typename R> inline R smin(R a, R b) { return (a < b) ? a : b; }
template <typename R> inline R smax(R a, R b) { return (a > b) ? a : b; }

// NC/src/auxiliary.h:33-60: octant order tables (q repeated 21 times)
inline uint64_t order_table(uint32_t q)
{
    uint64_t v = 0;
    for (int i = 0; i < MAX_NUM_LEVELS; ++i) v = (v << 3) | q;
    return v;
}

// NC/src/auxiliary.h:150-170
inline void tile_rect(uint32_t bx, uint32_t by, int gx, int gy,
                      int& txmin, int& tymin, int& txmax, int& tymax)
{
    const uint32_t xmin = bx >> 16, ymin = bx & 0xFFFFu;
    const uint32_t xmax = by >> 16, ymax = by & 0xFFFFu;
    txmin = std::min((int)(xmin / BLOCK_X), gx - 1);
    tymin = std::min((int)(ymin / BLOCK_Y), gy - 1);
    txmax = std::min((int)(xmax / BLOCK_X), gx - 1);
    tymax = std::min((int)(ymax / BLOCK_Y), gy - 1);
}

// NC/src/auxiliary.h:172-179
inline bool pix_in_bbox(uint32_t px, uint32_t py, uint32_t bx, uint32_t by)
{
    return px >= (bx >> 16) && py >= (bx & 0xFFFFu) && px <= (by >> 16) && py <= (by & 0xFFFFu);
}

template <typename R>
struct Ray
{
    R ro[3], rd[3], rd_inv[3];
    uint32_t pq;
};

// NC/src/forward.cu:102-113 (ray set-up), auxiliary.h:140-148, 294-302
template <typename R>
Ray<R> make_ray(int px, int py, int W, int H, R tanfovx, R tanfovy, R cx, R cy, const R* c2w)
{
    Ray<R> ray;
    const R pfx = (R)px, pfy = (R)py;
    const R dx = (pfx + R(0.5f) - cx) * R(2) * tanfovx / (R)W;
    const R dy = (pfy + R(0.5f) - cy) * R(2) * tanfovy / (R)H;
    const R dz = R(1);
    const R nrm = std::sqrt(dx * dx + dy * dy + dz * dz);
    const R rx = c2w[0] * dx + c2w[1] * dy + c2w[2] * dz;
    const R ry = c2w[4] * dx + c2w[5] * dy + c2w[6] * dz;
    const R rz = c2w[8] * dx + c2w[9] * dy + c2w[10] * dz;
    const R ninv = R(1) / nrm;
    ray.ro[0] = c2w[3]; ray.ro[1] = c2w[7]; ray.ro[2] = c2w[11];
    ray.rd[0] = rx * ninv; ray.rd[1] = ry * ninv; ray.rd[2] = rz * ninv;
    for (int i = 0; i < 3; ++i) ray.rd_inv[i] = R(1) / ray.rd[i];
    ray.pq = ((ray.rd[0] < 0) << 2) | ((ray.rd[1] < 0) << 1) | (ray.rd[2] < 0);
    return ray;
}

// NC/src/auxiliary.h:310-323
template <typename R>
inline void ray_aabb(const R* c, R l, const Ray<R>& ray, R& a, R& b)
{
    const R r = R(0.5f) * l;
    R lo[3], hi[3];
    for (int i = 0; i < 3; ++i)
    {
        const R dir = c[i] - ray.ro[i];
        const R c0 = (dir - r) * ray.rd_inv[i];
        const R c1 = (dir + r) * ray.rd_inv[i];
        lo[i] = smin(c0, c1);
        hi[i] = smax(c0, c1);
    }
    a = smax(smax(lo[0], lo[1]), lo[2]);
    b = smin(smin(hi[0], hi[1]), hi[2]);
}

// NC/src/auxiliary.h:348-356
template <typename R> inline R exp_linear_11(R x)
{
    return (x > R(1.1f)) ? x : rexp<R>(rfma<R>(R(0.909090909091f), x, R(-0.904689820196f)));
}
template <typename R> inline R exp_linear_11_bw(R x)
{
    return (x > R(1.1f)) ? R(1) : R(0.909090909091f) * rexp<R>(rfma<R>(R(0.909090909091f), x, R(-0.904689820196f)));
}

// NC/src/auxiliary.h:378-391 (corner order: bit2 = x, bit1 = y, bit0 = z)
template <typename R> inline void tri_w(const R* qt, R* w)
{
    const R wx[2] = {R(1) - qt[0], qt[0]};
    const R wy[2] = {R(1) - qt[1], qt[1]};
    const R wz[2] = {R(1) - qt[2], qt[2]};
    for (int i = 0; i < 8; ++i) w[i] = (wx[(i >> 2) & 1] * wy[(i >> 1) & 1]) * wz[i & 1];
}

template <typename R> inline R inv_nsamp(int n)
{
    return n == 1 ? R(1) : (n == 2 ? R(0.5f) : R(1.f / 3.f));
}

// Per-entry volume integral (NC/src/forward.cu:224-253; backward.cu:283-328).
// Returns vol; fills dI/dgeo if requested.
template <typename R>
inline R vox_vol(const R* c, R l, const R* geo, R a, R b, const Ray<R>& ray, int n_samp, R* dI)
{
    const R l_inv = R(1) / l;
    const R step_sz = (b - a) * inv_nsamp<R>(n_samp);
    const R t0 = rfma<R>(R(0.5f), step_sz, a);
    const R r = R(0.5f) * l;
    R qt[3], qs[3];
    for (int i = 0; i < 3; ++i)
    {
        const R pt = rfma<R>(t0, ray.rd[i], ray.ro[i]);
        qt[i] = (pt - (c[i] - r)) * l_inv;
        qs[i] = (step_sz * ray.rd[i]) * l_inv;
    }
    if (dI) for (int i = 0; i < 8; ++i) dI[i] = R(0);
    R vol = R(0);
    const R scale = R(STEP_SZ_SCALE) * step_sz;
    for (int k = 0; k < n_samp; ++k)
    {
        R w[8];
        tri_w<R>(qt, w);
        R d = geo[0] * w[0];
        for (int i = 1; i < 8; ++i) d = rfma<R>(geo[i], w[i], d);
        vol = vol + scale * exp_linear_11<R>(d);
        if (dI)
        {
            const R dd = scale * exp_linear_11_bw<R>(d);
            for (int i = 0; i < 8; ++i) dI[i] = rfma<R>(dd, w[i], dI[i]);
        }
        for (int i = 0; i < 3; ++i) qt[i] = qt[i] + qs[i];
    }
    return vol;
}

template <typename R>
void render_fwd(int W, int H, R tanfovx, R tanfovy, R cx, R cy, const R* c2w, R bg, int n_samp,
                const uint32_t* ranges, const uint32_t* vox_list, const uint32_t* bbox,
                const R* centers, const R* lengths, const R* geos, const R* rgbs,
                R* out_color, R* out_T, uint32_t* n_contrib, uint32_t* tile_last)
{
    const int gx = (W + BLOCK_X - 1) / BLOCK_X, gy = (H + BLOCK_Y - 1) / BLOCK_Y;
    const int T_tiles = gx * gy;
    const size_t HW = (size_t)H * W;
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < T_tiles; ++t)
    {
        const int tx = t % gx, ty = t / gx;
        const uint32_t rx = ranges[2 * t], ry = ranges[2 * t + 1];
        uint32_t tl = rx;  // NC/src/forward.cu:126-127, 392
        for (int py = ty * BLOCK_Y; py < std::min((ty + 1) * BLOCK_Y, H); ++py)
            for (int px = tx * BLOCK_X; px < std::min((tx + 1) * BLOCK_X, W); ++px)
            {
                const Ray<R> ray = make_ray<R>(px, py, W, H, tanfovx, tanfovy, cx, cy, c2w);
                R T = R(1), C[3] = {R(0), R(0), R(0)};
                uint32_t last = 0;
                // NC/src/forward.cu:156-327, flattened over batches
                for (uint32_t j = rx; j < ry; ++j)
                {
                    const uint32_t val = vox_list[j];
                    const uint32_t v = val & 0x1FFFFFFFu, q = val >> 29;
                    if (!pix_in_bbox(px, py, bbox[2 * v], bbox[2 * v + 1]) || q != ray.pq) continue;
                    const R* c = centers + 3 * (size_t)v;
                    R a, b;
                    ray_aabb<R>(c, lengths[v], ray, a, b);
                    if (a > b) continue;
                    const R vol = vox_vol<R>(c, lengths[v], geos + 8 * (size_t)v, a, b, ray, n_samp, nullptr);
                    const R alpha = std::min(R(MAX_ALPHA), R(1) - rexp<R>(-vol));
                    if (alpha < R(MIN_ALPHA)) continue;
                    const R pt_w = T * alpha;
                    for (int i = 0; i < 3; ++i) C[i] = rfma<R>(pt_w, rgbs[3 * (size_t)v + i], C[i]);
                    T = T * (R(1) - alpha);
                    last = j - rx + 1;
                    if (T < R(EARLY_STOP_T)) break;
                }
                const size_t pid = (size_t)py * W + px;
                n_contrib[pid] = last;
                for (int i = 0; i < 3; ++i) out_color[i * HW + pid] = C[i] + T * bg;
                out_T[pid] = T;
                tl = std::max(tl, rx + last);
            }
        tile_last[t] = tl;
    }
}

// ---- compositor with depth / distortion / normal / max_w and the
// regularizer gradients (NC/src/forward.cu:259-366, backward.cu:346-510) ----

// NC/src/auxiliary.h:305-308, 368-371
template <typename R> inline R depth_contracted(R x) { return (x < R(1)) ? x : R(2) - R(1) / x; }
template <typename R> inline R safe_rnorm(R x, R y, R z) { return R(1) / std::sqrt(x * x + y * y + z * z + R(1e-15f)); }

// Per-sample form of vox_vol: same vol and dI, plus each sample's volume term
// lv[k] and its dI contribution edI[k][i] = dd_k * w_k[i].
template <typename R>
inline R vox_samples(const R* c, R l, const R* geo, R a, R b, const Ray<R>& ray, int n_samp, R* lv, R (*edI)[8],
                     R* dI)
{
    const R l_inv = R(1) / l;
    const R step_sz = (b - a) * inv_nsamp<R>(n_samp);
    const R t0 = rfma<R>(R(0.5f), step_sz, a);
    const R r = R(0.5f) * l;
    R qt[3], qs[3];
    for (int i = 0; i < 3; ++i)
    {
        const R pt = rfma<R>(t0, ray.rd[i], ray.ro[i]);
        qt[i] = (pt - (c[i] - r)) * l_inv;
        qs[i] = (step_sz * ray.rd[i]) * l_inv;
    }
    if (dI) for (int i = 0; i < 8; ++i) dI[i] = R(0);
    R vol = R(0);
    const R scale = R(STEP_SZ_SCALE) * step_sz;
    for (int k = 0; k < n_samp; ++k)
    {
        R w[8];
        tri_w<R>(qt, w);
        R d = geo[0] * w[0];
        for (int i = 1; i < 8; ++i) d = rfma<R>(geo[i], w[i], d);
        const R term = scale * exp_linear_11<R>(d);
        lv[k] = term;
        vol = vol + term;
        if (dI)
        {
            const R dd = scale * exp_linear_11_bw<R>(d);
            for (int i = 0; i < 8; ++i)
            {
                edI[k][i] = dd * w[i];
                dI[i] = rfma<R>(dd, w[i], dI[i]);
            }
        }
        for (int i = 0; i < 3; ++i) qt[i] = qt[i] + qs[i];
    }
    return vol;
}

// NC/src/forward.cu:262-299: expected depth of one entry from its samples'
// alphas (n_samp 2, 3) or its midpoint (n_samp 1); dD[k] = d dval / d lv[k]
// scaled as in backward.cu:421-454 (without the dL_dD * T factor).
template <typename R>
inline R entry_depth(int n_samp, R a, R b, R alpha, const R* lv, R* dD)
{
    if (n_samp == 3)
    {
        const R st = R(0.3333333f) * (b - a);
        const R a0 = std::min(R(MAX_ALPHA), R(1) - rexp<R>(-lv[0]));
        const R a1 = std::min(R(MAX_ALPHA), R(1) - rexp<R>(-lv[1]));
        const R a2 = std::min(R(MAX_ALPHA), R(1) - rexp<R>(-lv[2]));
        const R t0 = a + R(0.5f) * st, t1 = a + R(1.5f) * st, t2 = a + R(2.5f) * st;
        if (dD)
        {
            dD[0] = (t0 + a1 * a2 * t2 - a1 * t1 - a2 * t2) * (R(1) - a0);
            dD[1] = (t1 + a0 * a2 * t2 - a0 * t1 - a2 * t2) * (R(1) - a1);
            dD[2] = (t2 + a0 * a1 * t2 - a0 * t2 - a1 * t2) * (R(1) - a2);
        }
        return a0 * t0 + (R(1) - a0) * a1 * t1 + (R(1) - a0) * (R(1) - a1) * a2 * t2;
    }
    if (n_samp == 2)
    {
        const R st = R(0.5f) * (b - a);
        const R a0 = std::min(R(MAX_ALPHA), R(1) - rexp<R>(-lv[0]));
        const R a1 = std::min(R(MAX_ALPHA), R(1) - rexp<R>(-lv[1]));
        const R t0 = a + R(0.5f) * st, t1 = a + R(1.5f) * st;
        if (dD)
        {
            dD[0] = (t0 - a1 * t1) * (R(1) - a0);
            dD[1] = (t1 - a0 * t1) * (R(1) - a1);
        }
        return a0 * t0 + (R(1) - a0) * a1 * t1;
    }
    const R t0 = R(0.5f) * (a + b);
    if (dD) dD[0] = t0 * (R(1) - alpha);
    return alpha * t0;
}

// NC/src/forward.cu:301-314 (sums in the reference's order)
template <typename R>
inline void lin_normal(const R* g, R* n)
{
    n[0] = (g[4] + g[5] + g[6] + g[7]) - (g[0] + g[1] + g[2] + g[3]);
    n[1] = (g[2] + g[3] + g[6] + g[7]) - (g[0] + g[1] + g[4] + g[5]);
    n[2] = (g[1] + g[3] + g[5] + g[7]) - (g[0] + g[2] + g[4] + g[6]);
}

template <typename R>
void render_fwd_ext(int W, int H, R tanfovx, R tanfovy, R cx, R cy, const R* c2w, R bg, int n_samp,
                    const uint32_t* ranges, const uint32_t* vox_list, const uint32_t* bbox, const R* centers,
                    const R* lengths, const R* geos, const R* rgbs, int need_depth, int need_dist, int need_normal,
                    int track_max_w, R* out_color, R* out_T, uint32_t* n_contrib, uint32_t* tile_last,
                    R* out_depth, R* out_normal, R* max_w, R* dist_loss)
{
    const int gx = (W + BLOCK_X - 1) / BLOCK_X, gy = (H + BLOCK_Y - 1) / BLOCK_Y;
    const int T_tiles = gx * gy;
    const size_t HW = (size_t)H * W;
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < T_tiles; ++t)
    {
        const int tx = t % gx, ty = t / gx;
        const uint32_t rx = ranges[2 * t], ry = ranges[2 * t + 1];
        uint32_t tl = rx;
        for (int py = ty * BLOCK_Y; py < std::min((ty + 1) * BLOCK_Y, H); ++py)
            for (int px = tx * BLOCK_X; px < std::min((tx + 1) * BLOCK_X, W); ++px)
            {
                const Ray<R> ray = make_ray<R>(px, py, W, H, tanfovx, tanfovy, cx, cy, c2w);
                const R dx = ((R)px + R(0.5f) - cx) * R(2) * tanfovx / (R)W;
                const R dy = ((R)py + R(0.5f) - cy) * R(2) * tanfovy / (R)H;
                const R rd_norm_inv = R(1) / std::sqrt(dx * dx + dy * dy + R(1) * R(1));
                R T = R(1), C[3] = {R(0), R(0), R(0)}, D = R(0), Dd = R(0), Nn[3] = {R(0), R(0), R(0)};
                int med_v = -1;
                R med_T = R(0);
                std::vector<R> ws, ms, spans;  // for the distortion loss value
                uint32_t last = 0;
                for (uint32_t j = rx; j < ry; ++j)
                {
                    const uint32_t val = vox_list[j];
                    const uint32_t v = val & 0x1FFFFFFFu, q = val >> 29;
                    if (!pix_in_bbox(px, py, bbox[2 * v], bbox[2 * v + 1]) || q != ray.pq) continue;
                    const R* c = centers + 3 * (size_t)v;
                    R a, b;
                    ray_aabb<R>(c, lengths[v], ray, a, b);
                    if (a > b) continue;
                    const R* g = geos + 8 * (size_t)v;
                    R lv[3];
                    const R vol = vox_samples<R>(c, lengths[v], g, a, b, ray, n_samp, lv, nullptr, nullptr);
                    const R alpha = std::min(R(MAX_ALPHA), R(1) - rexp<R>(-vol));
                    if (alpha < R(MIN_ALPHA)) continue;
                    const R pt_w = T * alpha;
                    for (int i = 0; i < 3; ++i) C[i] = rfma<R>(pt_w, rgbs[3 * (size_t)v + i], C[i]);
                    if (need_depth)
                    {
                        D = D + T * entry_depth<R>(n_samp, a, b, alpha, lv, nullptr);
                        if (T > R(0.5f))
                        {
                            med_v = (int)v;
                            med_T = T;
                        }
                    }
                    const R m = R(0.5f) * (depth_contracted<R>(a) + depth_contracted<R>(b));
                    if (need_dist) Dd = Dd + pt_w * m;
                    if (dist_loss)
                    {
                        ws.push_back(pt_w);
                        ms.push_back(m);
                        spans.push_back(depth_contracted<R>(b) - depth_contracted<R>(a));
                    }
                    if (need_normal)
                    {
                        R n[3];
                        lin_normal<R>(g, n);
                        const R rl = safe_rnorm<R>(n[0], n[1], n[2]);
                        for (int i = 0; i < 3; ++i) Nn[i] = Nn[i] + pt_w * rl * n[i];
                    }
                    T = T * (R(1) - alpha);
                    last = j - rx + 1;
                    if (track_max_w)
                    {
#pragma omp critical(oracle_max_w)
                        max_w[v] = std::max(max_w[v], pt_w);
                    }
                    if (T < R(EARLY_STOP_T)) break;
                }
                const size_t pid = (size_t)py * W + px;
                R Dmed = R(0);
                if (need_depth && med_v >= 0)
                {
                    // NC/src/forward.cu:330-366: 16-sample refinement inside the median voxel
                    const R* c = centers + 3 * (size_t)med_v;
                    const R l = lengths[med_v];
                    const R* g = geos + 8 * (size_t)med_v;
                    R a, b;
                    ray_aabb<R>(c, l, ray, a, b);
                    const R l_inv = R(1) / l;
                    const R step_sz = (b - a) * (R(1) / R(16));
                    const R t0 = rfma<R>(R(0.5f), step_sz, a);
                    const R r = R(0.5f) * l;
                    R qt[3], qs[3];
                    for (int i = 0; i < 3; ++i)
                    {
                        qt[i] = (rfma<R>(t0, ray.rd[i], ray.ro[i]) - (c[i] - r)) * l_inv;
                        qs[i] = (step_sz * ray.rd[i]) * l_inv;
                    }
                    Dmed = a - R(0.5f) * step_sz;
                    const R scale = R(STEP_SZ_SCALE) * step_sz;
                    for (int k = 0; k < 16 && med_T > R(0.5f); ++k)
                    {
                        Dmed = Dmed + step_sz;
                        R w[8];
                        tri_w<R>(qt, w);
                        R d = g[0] * w[0];
                        for (int i = 1; i < 8; ++i) d = rfma<R>(g[i], w[i], d);
                        med_T = med_T * rexp<R>(-(scale * exp_linear_11<R>(d)));
                        for (int i = 0; i < 3; ++i) qt[i] = qt[i] + qs[i];
                    }
                }
                n_contrib[pid] = last;
                for (int i = 0; i < 3; ++i) out_color[i * HW + pid] = C[i] + T * bg;
                out_T[pid] = T;
                if (need_depth)
                {
                    out_depth[pid] = D * rd_norm_inv;
                    out_depth[2 * HW + pid] = Dmed * rd_norm_inv;
                }
                if (need_dist) out_depth[HW + pid] = Dd;
                if (need_normal)
                    for (int i = 0; i < 3; ++i) out_normal[i * HW + pid] = Nn[i];
                if (dist_loss)
                {
                    // sum_ij w_i w_j |m_i - m_j| + 1/3 sum_i w_i^2 (contracted b_i - contracted a_i)
                    R L = R(0);
                    for (size_t i = 0; i < ws.size(); ++i)
                    {
                        for (size_t k = 0; k < ws.size(); ++k) L += ws[i] * ws[k] * std::fabs(ms[i] - ms[k]);
                        L += ws[i] * ws[i] * spans[i] / R(3);
                    }
                    dist_loss[pid] = L;
                }
                tl = std::max(tl, rx + last);
            }
        tile_last[t] = tl;
    }
}

template <typename R>
inline void atomic_add(double* p, double v)
{
#pragma omp atomic
    *p += v;
}

// NC/src/backward.cu:105-527 (no depth / normal / distortion / regularizers)
template <typename R>
void render_bwd(int W, int H, R tanfovx, R tanfovy, R cx, R cy, const R* c2w, R bg, int n_samp,
                const uint32_t* ranges, const uint32_t* vox_list, const uint32_t* bbox,
                const R* centers, const R* lengths, const R* geos, const R* rgbs,
                const R* out_T, const uint32_t* n_contrib, const uint32_t* tile_last,
                const R* dL_dcolor, const R* dL_dT, int N, double* acc /* N*12, zeroed */)
{
    const int gx = (W + BLOCK_X - 1) / BLOCK_X, gy = (H + BLOCK_Y - 1) / BLOCK_Y;
    const int T_tiles = gx * gy;
    const size_t HW = (size_t)H * W;
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < T_tiles; ++t)
    {
        const int tx = t % gx, ty = t / gx;
        const uint32_t rx = ranges[2 * t], rend = tile_last[t];
        for (int py = ty * BLOCK_Y; py < std::min((ty + 1) * BLOCK_Y, H); ++py)
            for (int px = tx * BLOCK_X; px < std::min((tx + 1) * BLOCK_X, W); ++px)
            {
                const Ray<R> ray = make_ray<R>(px, py, W, H, tanfovx, tanfovy, cx, cy, c2w);
                const size_t pid = (size_t)py * W + px;
                R T = out_T[pid];
                const uint32_t nc = n_contrib[pid];
                const R dpix[3] = {dL_dcolor[pid], dL_dcolor[HW + pid], dL_dcolor[2 * HW + pid]};
                R last = dL_dT[pid] + bg * (dpix[0] + dpix[1] + dpix[2]);
                for (uint32_t jj = rend; jj > rx; --jj)
                {
                    const uint32_t j = jj - 1;
                    if (j - rx >= nc) continue;
                    const uint32_t val = vox_list[j];
                    const uint32_t v = val & 0x1FFFFFFFu, q = val >> 29;
                    if (!pix_in_bbox(px, py, bbox[2 * v], bbox[2 * v + 1]) || q != ray.pq) continue;
                    const R* c = centers + 3 * (size_t)v;
                    R a, b;
                    ray_aabb<R>(c, lengths[v], ray, a, b);
                    if (a > b) continue;
                    R dI[8];
                    const R vol = vox_vol<R>(c, lengths[v], geos + 8 * (size_t)v, a, b, ray, n_samp, dI);
                    const R e = rexp<R>(-vol);
                    const R alpha = std::min(R(MAX_ALPHA), R(1) - e);
                    if (alpha < R(MIN_ALPHA)) continue;
                    T = T / (R(1) - alpha);
                    const R pt_w = alpha * T;
                    const R* col = rgbs + 3 * (size_t)v;
                    const R dpt = rfma<R>(dpix[2], col[2], rfma<R>(dpix[1], col[1], dpix[0] * col[0]));
                    const R diff = dpt - last;
                    const R dalpha = T * diff;
                    last = rfma<R>(alpha, diff, last);
                    const R dI_scale = dalpha * e;
                    double* g = acc + 12 * (size_t)v;
                    for (int i = 0; i < 8; ++i) atomic_add<R>(g + i, (double)(dI_scale * dI[i]));
                    for (int i = 0; i < 3; ++i) atomic_add<R>(g + 8 + i, (double)(pt_w * dpix[i]));
                    atomic_add<R>(g + 11, (double)std::fabs(dalpha * alpha));
                }
            }
    }
}

// NC/src/backward.cu:105-527 with depth (mean), distortion, normal and the
// R_concen / ascending regularizers.  dL_ddepth: [3, H, W] (channel 0 used),
// dL_dnormal [3, H, W], gt [3, H, W], out_depth [3, H, W] of the forward
// (channel 1 = distortion sum) -- each may be null when its feature is off.
template <typename R>
void render_bwd_ext(int W, int H, R tanfovx, R tanfovy, R cx, R cy, const R* c2w, R bg, int n_samp,
                    const uint32_t* ranges, const uint32_t* vox_list, const uint32_t* bbox, const R* centers,
                    const R* lengths, const R* geos, const R* rgbs, const R* out_T, const uint32_t* n_contrib,
                    const uint32_t* tile_last, const R* dL_dcolor, const R* dL_dT, int need_depth, int need_normal,
                    const R* dL_ddepth, const R* dL_dnormal, R lambda_R, const R* gt, R lambda_asc, R lambda_dist,
                    const R* out_depth, int N, double* acc)
{
    const int gx = (W + BLOCK_X - 1) / BLOCK_X, gy = (H + BLOCK_Y - 1) / BLOCK_Y;
    const int T_tiles = gx * gy;
    const size_t HW = (size_t)H * W;
    const R WH_inv = R(1) / (R)(W * H);
    const R w_R = lambda_R * WH_inv, w_asc = lambda_asc * WH_inv, w_dist = lambda_dist * WH_inv;
    const bool need_dist = lambda_dist > R(0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < T_tiles; ++t)
    {
        const int tx = t % gx, ty = t / gx;
        const uint32_t rx = ranges[2 * t], rend = tile_last[t];
        for (int py = ty * BLOCK_Y; py < std::min((ty + 1) * BLOCK_Y, H); ++py)
            for (int px = tx * BLOCK_X; px < std::min((tx + 1) * BLOCK_X, W); ++px)
            {
                const Ray<R> ray = make_ray<R>(px, py, W, H, tanfovx, tanfovy, cx, cy, c2w);
                const R ddx = ((R)px + R(0.5f) - cx) * R(2) * tanfovx / (R)W;
                const R ddy = ((R)py + R(0.5f) - cy) * R(2) * tanfovy / (R)H;
                const R rd_norm_inv = R(1) / std::sqrt(ddx * ddx + ddy * ddy + R(1) * R(1));
                const size_t pid = (size_t)py * W + px;
                const R T_final = out_T[pid];
                R T = T_final;
                const uint32_t nc = n_contrib[pid];
                const R dpix[3] = {dL_dcolor[pid], dL_dcolor[HW + pid], dL_dcolor[2 * HW + pid]};
                R last = dL_dT[pid] + bg * (dpix[0] + dpix[1] + dpix[2]);
                const R dL_dD = need_depth ? dL_ddepth[pid] * rd_norm_inv : R(0);
                R dN[3] = {R(0), R(0), R(0)};
                if (need_normal)
                    for (int i = 0; i < 3; ++i) dN[i] = dL_dnormal[i * HW + pid];
                R gtp[3] = {R(0), R(0), R(0)};
                if (lambda_R > R(0))
                    for (int i = 0; i < 3; ++i) gtp[i] = gt[i * HW + pid];
                R prefix_wm = R(0), suffix_wm = R(0), prefix_w = R(0), suffix_w = R(0);
                if (need_dist)
                {
                    prefix_wm = out_depth[HW + pid];
                    prefix_w = R(1) - T_final;
                }
                for (uint32_t jj = rend; jj > rx; --jj)
                {
                    const uint32_t j = jj - 1;
                    if (j - rx >= nc) continue;
                    const uint32_t val = vox_list[j];
                    const uint32_t v = val & 0x1FFFFFFFu, q = val >> 29;
                    if (!pix_in_bbox(px, py, bbox[2 * v], bbox[2 * v + 1]) || q != ray.pq) continue;
                    const R* c = centers + 3 * (size_t)v;
                    const R l = lengths[v];
                    R a, b;
                    ray_aabb<R>(c, l, ray, a, b);
                    if (a > b) continue;
                    const R* g = geos + 8 * (size_t)v;
                    R dI[8], lv[3], edI[3][8];
                    const R vol = vox_samples<R>(c, l, g, a, b, ray, n_samp, lv, edI, dI);
                    const R e = rexp<R>(-vol);
                    const R alpha = std::min(R(MAX_ALPHA), R(1) - e);
                    if (alpha < R(MIN_ALPHA)) continue;
                    T = T / (R(1) - alpha);
                    const R pt_w = alpha * T;
                    const R* col = rgbs + 3 * (size_t)v;
                    R dpt = rfma<R>(dpix[2], col[2], rfma<R>(dpix[1], col[1], dpix[0] * col[0]));
                    R dgeo[8] = {R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0)};
                    if (need_dist)
                    {
                        const R ad = depth_contracted<R>(a), bd = depth_contracted<R>(b);
                        const R now_m = R(0.5f) * (ad + bd);
                        const R now_wm = now_m * pt_w;
                        prefix_wm = prefix_wm - now_wm;
                        prefix_w = prefix_w - pt_w;
                        const R uni = R(0.6666666f) * pt_w * (bd - ad);
                        const R bi = R(2) * (now_m * (prefix_w - suffix_w) - (prefix_wm - suffix_wm));
                        dpt = dpt + w_dist * (uni + bi);
                        suffix_wm = suffix_wm + now_wm;
                        suffix_w = suffix_w + pt_w;
                    }
                    if (need_normal)
                    {
                        R n[3];
                        lin_normal<R>(g, n);
                        const R rl = safe_rnorm<R>(n[0], n[1], n[2]);
                        const R sn[3] = {rl * n[0], rl * n[1], rl * n[2]};
                        dpt = dpt + (dN[0] * sn[0] + dN[1] * sn[1] + dN[2] * sn[2]);
                        const R ds[3] = {pt_w * dN[0], pt_w * dN[1], pt_w * dN[2]};
                        const R dsn = ds[0] * n[0] + ds[1] * n[1] + ds[2] * n[2];
                        R dl[3];
                        for (int i = 0; i < 3; ++i) dl[i] = rl * (ds[i] - dsn * rl * sn[i]);
                        for (int k = 0; k < 8; ++k)
                            dgeo[k] = ((k & 4) ? dl[0] : -dl[0]) + ((k & 2) ? dl[1] : -dl[1]) +
                                      ((k & 1) ? dl[2] : -dl[2]);
                    }
                    const R diff = dpt - last;
                    const R dalpha = T * diff;
                    last = rfma<R>(alpha, diff, last);
                    const R dIs = dalpha * e;
                    for (int k = 0; k < 8; ++k) dgeo[k] = dgeo[k] + dIs * dI[k];
                    if (need_depth)
                    {
                        R dD[3];
                        const R dval = entry_depth<R>(n_samp, a, b, alpha, lv, dD);
                        last = last + dL_dD * dval;
                        const R f = dL_dD * T;
                        if (n_samp == 1)
                        {
                            for (int k = 0; k < 8; ++k) dgeo[k] = dgeo[k] + (f * dD[0]) * dI[k];
                        }
                        else
                        {
                            for (int k = 0; k < 8; ++k)
                            {
                                R s2 = (f * dD[0]) * edI[0][k];
                                for (int m = 1; m < n_samp; ++m) s2 = s2 + (f * dD[m]) * edI[m][k];
                                dgeo[k] = dgeo[k] + s2;
                            }
                        }
                    }
                    if (lambda_asc > R(0))
                    {
                        const R r = R(0.5f) * l, l_inv = R(1) / l;
                        R qa[3], qb[3];
                        for (int i = 0; i < 3; ++i)
                        {
                            qa[i] = (rfma<R>(a, ray.rd[i], ray.ro[i]) - (c[i] - r)) * l_inv;
                            qb[i] = (rfma<R>(b, ray.rd[i], ray.ro[i]) - (c[i] - r)) * l_inv;
                        }
                        R wa[8], wb[8];
                        tri_w<R>(qa, wa);
                        tri_w<R>(qb, wb);
                        R da = g[0] * wa[0], db = g[0] * wb[0];
                        for (int k = 1; k < 8; ++k)
                        {
                            da = rfma<R>(g[k], wa[k], da);
                            db = rfma<R>(g[k], wb[k], db);
                        }
                        const R rw = w_asc * pt_w * (R)(da > db ? 1 : 0);
                        for (int k = 0; k < 8; ++k) dgeo[k] = dgeo[k] + rw * (wa[k] - wb[k]);
                    }
                    R drgb[3] = {pt_w * dpix[0], pt_w * dpix[1], pt_w * dpix[2]};
                    if (lambda_R > R(0))
                        for (int i = 0; i < 3; ++i) drgb[i] = drgb[i] + w_R * pt_w * R(2) * (col[i] - gtp[i]);
                    double* gacc = acc + 12 * (size_t)v;
                    for (int k = 0; k < 8; ++k) atomic_add<R>(gacc + k, (double)dgeo[k]);
                    for (int i = 0; i < 3; ++i) atomic_add<R>(gacc + 8 + i, (double)drgb[i]);
                    atomic_add<R>(gacc + 11, (double)std::fabs(dalpha * alpha));
                }
            }
    }
}

// NC/src/sh_compute.cu:33-50
constexpr float SH_C0 = 0.28209479177387814f;
constexpr float SH_C1 = 0.4886025119029199f;
constexpr float SH_C2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                            -1.0925484305920792f, 0.5462742152960396f};
constexpr float SH_C3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                            0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                            -0.5900435899266435f};

// basis values b[0..15] for the direction (deg 0..3); NC/src/sh_compute.cu:55-176
inline int sh_basis(int deg, float x, float y, float z, float* b)
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

inline void sh_dir(const float* c, const float* cam, float& x, float& y, float& z)
{
    const float dx = c[0] - cam[0], dy = c[1] - cam[1], dz = c[2] - cam[2];
    const float rn = 1.f / std::sqrt(dx * dx + dy * dy + dz * dz);
    x = dx * rn; y = dy * rn; z = dz * rn;
}

}  // namespace

extern "C" {

float oracle_expf(float x) { return svr_ref_expf(x); }

void oracle_expf_array(int n, const float* x, float* y)
{
    for (int i = 0; i < n; ++i) y[i] = svr_ref_expf(x[i]);
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// NC/src/preprocess.cu:24-132 (kernel) and :171-174 (focal, tile grid)
void oracle_preprocess(int W, int H, float tanfovx, float tanfovy, float cx, float cy,
                       const float* w2c, const float* c2w, float near_, int N,
                       const float* centers, const float* lengths,
                       uint32_t* n_dup, uint32_t* bbox, uint32_t* qbits)
{
    const int gx = (W + BLOCK_X - 1) / BLOCK_X, gy = (H + BLOCK_Y - 1) / BLOCK_Y;
    const float focal_x = 0.5f * (float)W / tanfovx;
    const float focal_y = 0.5f * (float)H / tanfovy;
    const float ro[3] = {c2w[3], c2w[7], c2w[11]};
#pragma omp parallel for schedule(static)
    for (int v = 0; v < N; ++v)
    {
        n_dup[v] = 0; bbox[2 * v] = 0; bbox[2 * v + 1] = 0; qbits[v] = 0;
        const float* c = centers + 3 * (size_t)v;
        const float r = 0.5f * lengths[v];
        const float rel[3] = {c[0] - ro[0], c[1] - ro[1], c[2] - ro[2]};
        if (rel[0] * rel[0] + rel[1] * rel[1] + rel[2] * rel[2] < near_ * near_) continue;
        uint32_t bits = 0;
        float umin = 1e9f, vmin = 1e9f, umax = -1e9f, vmax = -1e9f;
        for (int i = 0; i < 8; ++i)
        {
            const float sx = (float)(((i & 4) >> 2) * 2 - 1);
            const float sy = (float)(((i & 2) >> 1) * 2 - 1);
            const float sz = (float)((i & 1) * 2 - 1);
            const float p[3] = {c[0] + r * sx, c[1] + r * sy, c[2] + r * sz};
            const float camx = w2c[0] * p[0] + w2c[1] * p[1] + w2c[2] * p[2] + w2c[3];
            const float camy = w2c[4] * p[0] + w2c[5] * p[1] + w2c[6] * p[2] + w2c[7];
            const float camz = w2c[8] * p[0] + w2c[9] * p[1] + w2c[10] * p[2] + w2c[11];
            if (camz < near_) continue;
            const float iz = 1.0f / camz;
            const float u = camx * iz, vv = camy * iz;
            const uint32_t q = ((p[0] < ro[0]) << 2) | ((p[1] < ro[1]) << 1) | (p[2] < ro[2]);
            umin = smin(umin, u); vmin = smin(vmin, vv);
            umax = smax(umax, u); vmax = smax(vmax, vv);
            bits |= 1u << q;
        }
        const float cxh = cx - 0.5f, cyh = cy - 0.5f;
        const float bminx = std::fmax(focal_x * umin + cxh, 0.0f);
        const float bminy = std::fmax(focal_y * vmin + cyh, 0.0f);
        const float bmaxx = std::fmin(focal_x * umax + cxh, (float)W);
        const float bmaxy = std::fmin(focal_y * vmax + cyh, (float)H);
        if (bminx > bmaxx || bminy > bmaxy) continue;
        const uint32_t bx = (((uint32_t)std::lrint(bminx)) << 16) | ((uint32_t)std::lrint(bminy));
        const uint32_t by = (((uint32_t)std::lrint(bmaxx)) << 16) | ((uint32_t)std::lrint(bmaxy));
        int txmin, tymin, txmax, tymax;
        tile_rect(bx, by, gx, gy, txmin, tymin, txmax, tymax);
        const int tiles = (1 + tymax - tymin) * (1 + txmax - txmin);
        n_dup[v] = (uint32_t)(tiles * __builtin_popcount(bits));
        bbox[2 * v] = bx; bbox[2 * v + 1] = by;
        qbits[v] = bits;
    }
}

// NC/src/forward.cu:523-577 (duplicateWithKeys), :649-663 (scan, R),
// :703-708 (stable sort of the 128-bit key), :582-604 + :733 (ranges).
// Outputs the sorted canonical keys (hi, lo), the sorted values and ranges.
void oracle_binning(int W, int H, int N, const int64_t* octpaths, const uint32_t* n_dup,
                    const uint32_t* bbox, const uint32_t* qbits, int64_t R,
                    uint64_t* keys_hi, uint64_t* keys_lo, uint32_t* vox_list, uint32_t* ranges)
{
    const int gx = (W + BLOCK_X - 1) / BLOCK_X, gy = (H + BLOCK_Y - 1) / BLOCK_Y;
    const int T_tiles = gx * gy;
    struct E { uint64_t hi, lo; uint32_t val; };
    std::vector<E> ent;
    ent.reserve((size_t)R);
    for (int v = 0; v < N; ++v)
    {
        if (n_dup[v] == 0) continue;
        int txmin, tymin, txmax, tymax;
        tile_rect(bbox[2 * v], bbox[2 * v + 1], gx, gy, txmin, tymin, txmax, tymax);
        const uint64_t path = (uint64_t)octpaths[v];
        for (uint32_t q = 0; q < 8; ++q)
        {
            if (!(qbits[v] & (1u << q))) continue;
            const uint64_t rank = path ^ order_table(q);
            for (int y = tymin; y <= tymax; ++y)
                for (int x = txmin; x <= txmax; ++x)
                {
                    const uint64_t tile = (uint64_t)((uint32_t)y * (uint32_t)gx + (uint32_t)x);
                    ent.push_back({tile >> 1, (tile << 63) | rank, (q << 29) | (uint32_t)v});
                }
        }
    }
    std::stable_sort(ent.begin(), ent.end(), [](const E& a, const E& b) {
        return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo);
    });
    std::memset(ranges, 0, sizeof(uint32_t) * 2 * (size_t)T_tiles);
    const size_t L = ent.size();
    for (size_t i = 0; i < L; ++i)
    {
        keys_hi[i] = ent[i].hi; keys_lo[i] = ent[i].lo; vox_list[i] = ent[i].val;
        const uint32_t cur = (uint32_t)((ent[i].hi << 1) | (ent[i].lo >> 63));
        if (i == 0) ranges[2 * cur] = 0;
        else
        {
            const uint32_t prev = (uint32_t)((ent[i - 1].hi << 1) | (ent[i - 1].lo >> 63));
            if (cur != prev) { ranges[2 * prev + 1] = (uint32_t)i; ranges[2 * cur] = (uint32_t)i; }
        }
        if (i == L - 1) ranges[2 * cur + 1] = (uint32_t)L;
    }
}

void oracle_render_fwd_f32(int W, int H, float tanfovx, float tanfovy, float cx, float cy,
                           const float* c2w, float bg, int n_samp,
                           const uint32_t* ranges, const uint32_t* vox_list, const uint32_t* bbox,
                           const float* centers, const float* lengths, const float* geos, const float* rgbs,
                           float* out_color, float* out_T, uint32_t* n_contrib, uint32_t* tile_last)
{
    render_fwd<float>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, n_samp, ranges, vox_list, bbox,
                      centers, lengths, geos, rgbs, out_color, out_T, n_contrib, tile_last);
}

void oracle_render_fwd_f64(int W, int H, double tanfovx, double tanfovy, double cx, double cy,
                           const double* c2w, double bg, int n_samp,
                           const uint32_t* ranges, const uint32_t* vox_list, const uint32_t* bbox,
                           const double* centers, const double* lengths, const double* geos, const double* rgbs,
                           double* out_color, double* out_T, uint32_t* n_contrib, uint32_t* tile_last)
{
    render_fwd<double>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, n_samp, ranges, vox_list, bbox,
                       centers, lengths, geos, rgbs, out_color, out_T, n_contrib, tile_last);
}

void oracle_render_bwd_f32(int W, int H, float tanfovx, float tanfovy, float cx, float cy,
                           const float* c2w, float bg, int n_samp,
                           const uint32_t* ranges, const uint32_t* vox_list, const uint32_t* bbox,
                           const float* centers, const float* lengths, const float* geos, const float* rgbs,
                           const float* out_T, const uint32_t* n_contrib, const uint32_t* tile_last,
                           const float* dL_dcolor, const float* dL_dT, int N, double* acc)
{
    render_bwd<float>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, n_samp, ranges, vox_list, bbox,
                      centers, lengths, geos, rgbs, out_T, n_contrib, tile_last, dL_dcolor, dL_dT, N, acc);
}

void oracle_render_bwd_f64(int W, int H, double tanfovx, double tanfovy, double cx, double cy,
                           const double* c2w, double bg, int n_samp,
                           const uint32_t* ranges, const uint32_t* vox_list, const uint32_t* bbox,
                           const double* centers, const double* lengths, const double* geos, const double* rgbs,
                           const double* out_T, const uint32_t* n_contrib, const uint32_t* tile_last,
                           const double* dL_dcolor, const double* dL_dT, int N, double* acc)
{
    render_bwd<double>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, n_samp, ranges, vox_list, bbox,
                       centers, lengths, geos, rgbs, out_T, n_contrib, tile_last, dL_dcolor, dL_dT, N, acc);
}

// NC/src/sh_compute.cu:55-106, 180-208.  idx == nullptr (K == 0) means all voxels.
void oracle_sh_fwd(int deg, int K, const int64_t* idx, int N, const float* centers, const float* cam,
                   const float* sh0, const float* shs, int M, float* rgbs)
{
    std::memset(rgbs, 0, sizeof(float) * 3 * (size_t)N);
    const int cnt = K ? K : N;
#pragma omp parallel for schedule(static)
    for (int t = 0; t < cnt; ++t)
    {
        const int64_t v = K ? idx[t] : t;
        float x, y, z, b[16];
        sh_dir(centers + 3 * v, cam, x, y, z);
        const int nb = sh_basis(deg, x, y, z, b);
        const float* s0 = sh0 + 3 * v;
        const float* s = shs + (size_t)v * (M - 1) * 3;
        for (int ch = 0; ch < 3; ++ch)
        {
            // NC/src/sh_compute.cu:66-99: result accumulated term by term
            float res = SH_C0 * s0[ch];
            if (nb > 1)
            {
                res = res - (SH_C1 * y) * s[0 * 3 + ch] + (SH_C1 * z) * s[1 * 3 + ch] - (SH_C1 * x) * s[2 * 3 + ch];
            }
            for (int k = 4; k < nb; ++k) res = res + b[k] * s[(k - 1) * 3 + ch];
            res = res + 0.5f;
            rgbs[3 * v + ch] = res > 0.f ? res : 0.f;
        }
    }
}

// NC/src/sh_compute.cu:110-176, 212-240 (no clamp mask, :118-121)
void oracle_sh_bwd(int deg, int K, const int64_t* idx, int N, const float* centers, const float* cam,
                   const float* dL_drgbs, int M, float* dL_dsh0, float* dL_dshs)
{
    std::memset(dL_dsh0, 0, sizeof(float) * 3 * (size_t)N);
    std::memset(dL_dshs, 0, sizeof(float) * 3 * (size_t)N * (M - 1));
    const int cnt = K ? K : N;
#pragma omp parallel for schedule(static)
    for (int t = 0; t < cnt; ++t)
    {
        const int64_t v = K ? idx[t] : t;
        float x, y, z, b[16];
        sh_dir(centers + 3 * v, cam, x, y, z);
        const int nb = sh_basis(deg, x, y, z, b);
        for (int ch = 0; ch < 3; ++ch)
        {
            const float g = dL_drgbs[3 * v + ch];
            dL_dsh0[3 * v + ch] = SH_C0 * g;
            for (int k = 1; k < nb; ++k) dL_dshs[((size_t)v * (M - 1) + (k - 1)) * 3 + ch] = b[k] * g;
        }
    }
}

// NC/src/adam_step.cu:21-64, 106-137 (unbiased); :67-102, 139-162 (biased)
void oracle_adam_step(int biased, int sparse, int64_t numel, float* p, const float* g, float* m, float* v,
                      double step, double lr, double beta1, double beta2, float eps)
{
    float step_size, rb2;
    if (!biased)
    {
        const double bc1 = 1.0 - std::pow(beta1, step);
        const double bc2 = 1.0 - std::pow(beta2, step);
        step_size = (float)(lr / bc1);
        rb2 = (float)(1.0 / std::sqrt(bc2));
    }
    const float b1 = (float)beta1, b2 = (float)beta2, lrf = (float)lr;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < numel; ++i)
    {
        const float gv = g[i];
        if (sparse && gv == 0.0f) continue;
        const float mv = std::fma(b1, m[i], std::fma(-b1, gv, gv));
        const float vv = std::fma(b2, v[i], std::fma(-b2, gv * gv, gv * gv));
        if (!biased)
        {
            const float denom = std::fma(std::sqrt(vv), rb2, eps);
            p[i] -= step_size * (mv / denom);
        }
        else
        {
            const float denom = std::sqrt(vv) + eps;
            p[i] -= lrf * (mv / denom);
        }
        m[i] = mv; v[i] = vv;
    }
}

// NC/src/utils.cu:106-127
void oracle_voxel_order_rank(int N, const int64_t* octpaths, int64_t* ranks /* 8*N */)
{
    for (int q = 0; q < 8; ++q)
        for (int v = 0; v < N; ++v) ranks[(size_t)q * N + v] = (int64_t)((uint64_t)octpaths[v] ^ order_table(q));
}

// geo_params_gather.cu:22-45 / 74-108: out[idx, k, d] = grid[key[idx, k], d]
// for idx in care (n_dim = 1 for the density form).
void oracle_gather(int64_t n_care, int n_dim, const int64_t* vox_key, const int64_t* care, const float* grid,
                   float* out)
{
    for (int64_t i = 0; i < n_care; ++i)
    {
        const int64_t idx = care[i];
        for (int k = 0; k < 8; ++k)
            for (int d = 0; d < n_dim; ++d)
                out[(idx * 8 + k) * n_dim + d] = grid[vox_key[idx * 8 + k] * n_dim + d];
    }
}

// geo_params_gather.cu:47-72 / 110-132: grad_grid[key[idx, k], d] += g[idx, k, d]
// (summed in double here).
void oracle_gather_bw(int64_t n_care, int n_dim, const int64_t* vox_key, const int64_t* care, const float* g,
                      double* grad_grid)
{
    for (int64_t i = 0; i < n_care; ++i)
    {
        const int64_t idx = care[i];
        for (int k = 0; k < 8; ++k)
            for (int d = 0; d < n_dim; ++d)
                grad_grid[vox_key[idx * 8 + k] * n_dim + d] += (double)g[(idx * 8 + k) * n_dim + d];
    }
}

// tv_compute.cu:23-80.  grad (double, in/out) holds the gradient before the
// call; with tv_sparse an item takes part when any corner gradient is
// non-zero in that incoming gradient.  Each corner adds w * sum over its 3
// cube neighbours of (v > v_neighbour ? 1 : -1).
void oracle_tv_bw(int64_t N, int C, const float* grid, const int64_t* vox_key, float weight,
                  const float* vox_size_inv, int no_tv_s, int tv_sparse, double* grad)
{
    std::vector<unsigned char> take((size_t)N * C, 1);
    if (tv_sparse)
        for (int64_t n = 0; n < N; ++n)
            for (int c = 0; c < C; ++c)
            {
                bool any = false;
                for (int k = 0; k < 8; ++k) any = any || (grad[vox_key[n * 8 + k] * C + c] != 0.0);
                take[(size_t)n * C + c] = any;
            }
    for (int64_t n = 0; n < N; ++n)
        for (int c = 0; c < C; ++c)
        {
            if (!take[(size_t)n * C + c]) continue;
            float v[8];
            for (int k = 0; k < 8; ++k) v[k] = grid[vox_key[n * 8 + k] * C + c];
            float w = weight;
            if (!no_tv_s) w *= 0.01f * vox_size_inv[n];
            for (int k = 0; k < 8; ++k)
            {
                int sgn = 0;
                sgn += (v[k] > v[k ^ 1]) ? 1 : -1;
                sgn += (v[k] > v[k ^ 2]) ? 1 : -1;
                sgn += (v[k] > v[k ^ 4]) ? 1 : -1;
                grad[vox_key[n * 8 + k] * C + c] += (double)(w * (float)sgn);
            }
        }
}

// utils.cu:23-43 (no contraction)
void oracle_is_in_cone(int64_t n, float tanfovx, float tanfovy, float near_plane, const float* w2c,
                       const float* pts, unsigned char* is_in)
{
    for (int64_t i = 0; i < n; ++i)
    {
        const float x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
        const float cx = w2c[0] * x + w2c[1] * y + w2c[2] * z + w2c[3];
        const float cy = w2c[4] * x + w2c[5] * y + w2c[6] * z + w2c[7];
        const float cz = w2c[8] * x + w2c[9] * y + w2c[10] * z + w2c[11];
        is_in[i] = (cz > near_plane) && (std::fabs(cx) <= tanfovx * cz) && (std::fabs(cy) <= tanfovy * cz);
    }
}

// utils.cu:45-102: rays (depth == nullptr) or points at depth, planar [3, H, W]
void oracle_pixel_rays(int W, int H, float cx, float cy, float tanfovx, float tanfovy, const float* c2w,
                       const float* depth, float* out)
{
    const int64_t n = (int64_t)W * H;
    for (int64_t pid = 0; pid < n; ++pid)
    {
        const float yid = (float)(pid / W), xid = (float)(pid % W);
        const float dx = (xid + 0.5f - cx) * 2.f * tanfovx / (float)W;
        const float dy = (yid + 0.5f - cy) * 2.f * tanfovy / (float)H;
        const float rx = c2w[0] * dx + c2w[1] * dy + c2w[2] * 1.f;
        const float ry = c2w[4] * dx + c2w[5] * dy + c2w[6] * 1.f;
        const float rz = c2w[8] * dx + c2w[9] * dy + c2w[10] * 1.f;
        if (depth)
        {
            out[pid] = c2w[3] + depth[pid] * rx;
            out[n + pid] = c2w[7] + depth[pid] * ry;
            out[2 * n + pid] = c2w[11] + depth[pid] * rz;
        }
        else
        {
            out[pid] = rx;
            out[n + pid] = ry;
            out[2 * n + pid] = rz;
        }
    }
}

#define ORACLE_EXT_FWD(NAME, R)                                                                              \
    void NAME(int W, int H, R tanfovx, R tanfovy, R cx, R cy, const R* c2w, R bg, int n_samp,                 \
              const uint32_t* ranges, const uint32_t* vox_list, const uint32_t* bbox, const R* centers,        \
              const R* lengths, const R* geos, const R* rgbs, int need_depth, int need_dist, int need_normal,  \
              int track_max_w, R* out_color, R* out_T, uint32_t* n_contrib, uint32_t* tile_last, R* out_depth, \
              R* out_normal, R* max_w, R* dist_loss)                                                           \
    {                                                                                                          \
        render_fwd_ext<R>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, n_samp, ranges, vox_list, bbox, centers,   \
                          lengths, geos, rgbs, need_depth, need_dist, need_normal, track_max_w, out_color,     \
                          out_T, n_contrib, tile_last, out_depth, out_normal, max_w, dist_loss);               \
    }
ORACLE_EXT_FWD(oracle_render_fwd_ext_f32, float)
ORACLE_EXT_FWD(oracle_render_fwd_ext_f64, double)

#define ORACLE_EXT_BWD(NAME, R)                                                                              \
    void NAME(int W, int H, R tanfovx, R tanfovy, R cx, R cy, const R* c2w, R bg, int n_samp,                 \
              const uint32_t* ranges, const uint32_t* vox_list, const uint32_t* bbox, const R* centers,        \
              const R* lengths, const R* geos, const R* rgbs, const R* out_T, const uint32_t* n_contrib,       \
              const uint32_t* tile_last, const R* dL_dcolor, const R* dL_dT, int need_depth, int need_normal,  \
              const R* dL_ddepth, const R* dL_dnormal, R lambda_R, const R* gt, R lambda_asc, R lambda_dist,   \
              const R* out_depth, int N, double* acc)                                                          \
    {                                                                                                          \
        render_bwd_ext<R>(W, H, tanfovx, tanfovy, cx, cy, c2w, bg, n_samp, ranges, vox_list, bbox, centers,   \
                          lengths, geos, rgbs, out_T, n_contrib, tile_last, dL_dcolor, dL_dT, need_depth,      \
                          need_normal, dL_ddepth, dL_dnormal, lambda_R, gt, lambda_asc, lambda_dist,           \
                          out_depth, N, acc);                                                                  \
    }
ORACLE_EXT_BWD(oracle_render_bwd_ext_f32, float)
ORACLE_EXT_BWD(oracle_render_bwd_ext_f64, double)

}  // extern "C"
