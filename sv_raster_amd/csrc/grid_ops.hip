// Grid-point parameters (gather to voxels and its adjoint, total-variation
// gradient) and the camera utilities of the reference module.
//
// Reference: backends/new_cuda/src/geo_params_gather.cu:22-224,
// tv_compute.cu:23-111, utils.cu:23-102 and 188-245, raster_state.cu:207-260.
//
// Mapping: the gathers run one lane per (voxel, corner) (or per (voxel,
// corner, channel)), so the 8 corner keys of a voxel and its 8 outputs are
// contiguous across neighbouring lanes; the adjoints scatter with float
// atomics into the grid points (shared by at most 8 voxels each).
#include "svr_device.h"
#include "svr_host.h"

namespace svr {

// geo_params_gather.cu:22-45
__global__ __launch_bounds__(256) void gather_geo_kernel(int64_t n_care, const int64_t* __restrict__ vox_key,
                                                         const int64_t* __restrict__ care_idx,
                                                         const float* __restrict__ grid_pts,
                                                         float* __restrict__ geo_params)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 8 * n_care) return;
    const int64_t idx = care_idx[t >> 3];
    const int k = (int)(t & 7);
    geo_params[8 * idx + k] = grid_pts[vox_key[8 * idx + k]];
}

// geo_params_gather.cu:47-72 (zero gradients add nothing and are skipped)
__global__ __launch_bounds__(256) void gather_geo_bw_kernel(int64_t n_care, const int64_t* __restrict__ vox_key,
                                                            const int64_t* __restrict__ care_idx,
                                                            const float* __restrict__ dL_dgeo,
                                                            float* __restrict__ dL_dgrid)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 8 * n_care) return;
    const int64_t idx = care_idx[t >> 3];
    const int k = (int)(t & 7);
    const float g = dL_dgeo[8 * idx + k];
    if (g != 0.f) unsafeAtomicAdd(dL_dgrid + vox_key[8 * idx + k], g);
}

// geo_params_gather.cu:74-132 for n_dim channels per grid point
__global__ __launch_bounds__(256) void gather_feat_kernel(int64_t n_care, int n_dim,
                                                          const int64_t* __restrict__ vox_key,
                                                          const int64_t* __restrict__ care_idx,
                                                          const float* __restrict__ grid_pts,
                                                          float* __restrict__ feat_params)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = 8 * (int64_t)n_dim;
    if (t >= per * n_care) return;
    const int64_t i = t / per;
    const int r = (int)(t - i * per);
    const int k = r / n_dim, d = r - k * n_dim;
    const int64_t idx = care_idx[i];
    feat_params[per * idx + r] = grid_pts[vox_key[8 * idx + k] * n_dim + d];
}

__global__ __launch_bounds__(256) void gather_feat_bw_kernel(int64_t n_care, int n_dim,
                                                             const int64_t* __restrict__ vox_key,
                                                             const int64_t* __restrict__ care_idx,
                                                             const float* __restrict__ dL_dfeat,
                                                             float* __restrict__ dL_dgrid)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = 8 * (int64_t)n_dim;
    if (t >= per * n_care) return;
    const int64_t i = t / per;
    const int r = (int)(t - i * per);
    const int k = r / n_dim, d = r - k * n_dim;
    const int64_t idx = care_idx[i];
    const float g = dL_dfeat[per * idx + r];
    if (g != 0.f) unsafeAtomicAdd(dL_dgrid + vox_key[8 * idx + k] * n_dim + d, g);
}

// tv_compute.cu:36-43: an item (voxel, channel) takes part in the sparse form
// when any of its 8 corner gradients is non-zero.  Evaluated here on the
// gradient as it stands before this update, for all items at once.
__global__ __launch_bounds__(256) void tv_mask_kernel(int64_t NC, int C, const int64_t* __restrict__ vox_key,
                                                      const float* __restrict__ grad, uint8_t* __restrict__ mask)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NC) return;
    const int64_t iN = t / C;
    const int iC = (int)(t - iN * C);
    bool any = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) any = any || (grad[vox_key[8 * iN + k] * C + iC] != 0.f);
    mask[t] = any ? 1 : 0;
}

// tv_compute.cu:45-80: gradient of the L1 differences along the 12 edges of
// each voxel's corner cube, each corner getting +-w per neighbour (-w on ties).
template <bool NO_TV_S>
__global__ __launch_bounds__(256) void tv_bw_kernel(int64_t NC, int C, const float* __restrict__ grid_pts,
                                                    const int64_t* __restrict__ vox_key, float weight,
                                                    const float* __restrict__ vox_size_inv,
                                                    const uint8_t* __restrict__ mask, float* __restrict__ grad)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NC) return;
    if (mask && !mask[t]) return;
    const int64_t iN = t / C;
    const int iC = (int)(t - iN * C);
    int64_t key[8];
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
        key[k] = vox_key[8 * iN + k];
        v[k] = grid_pts[key[k] * C + iC];
    }
    float w = weight;
    if (!NO_TV_S) w *= 0.01f * vox_size_inv[iN];
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
        int s = 0;
        s += (v[k] > v[k ^ 1]) ? 1 : -1;
        s += (v[k] > v[k ^ 2]) ? 1 : -1;
        s += (v[k] > v[k ^ 4]) ? 1 : -1;
        unsafeAtomicAdd(grad + key[k] * C + iC, w * (float)s);
    }
}

// utils.cu:23-43
__global__ __launch_bounds__(256) void is_in_cone_kernel(int64_t n, float tanfovx, float tanfovy, float near_plane,
                                                         const float* __restrict__ w2c, const float* __restrict__ pts,
                                                         uint8_t* __restrict__ is_in)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    const float cx = w2c[0] * x + w2c[1] * y + w2c[2] * z + w2c[3];
    const float cy = w2c[4] * x + w2c[5] * y + w2c[6] * z + w2c[7];
    const float cz = w2c[8] * x + w2c[9] * y + w2c[10] * z + w2c[11];
    is_in[i] = (cz > near_plane) && (fabsf(cx) <= tanfovx * cz) && (fabsf(cy) <= tanfovy * cz);
}

// utils.cu:45-102: per-pixel world ray direction (not normalised), and the
// point at depth d along it from the camera centre; planar [3, H, W].
template <bool PTS>
__global__ __launch_bounds__(256) void pixel_rays_kernel(int W, int H, float cx, float cy, float tanfovx,
                                                         float tanfovy, const float* __restrict__ c2w,
                                                         const float* __restrict__ depth, float* __restrict__ out)
{
    const int64_t n_pix = (int64_t)W * H;
    const int64_t pid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pid >= n_pix) return;
    const float yid = (float)(pid / W);
    const float xid = (float)(pid % W);
    const float dx = (xid + 0.5f - cx) * 2.f * tanfovx / (float)W;
    const float dy = (yid + 0.5f - cy) * 2.f * tanfovy / (float)H;
    const float rx = c2w[0] * dx + c2w[1] * dy + c2w[2] * 1.f;
    const float ry = c2w[4] * dx + c2w[5] * dy + c2w[6] * 1.f;
    const float rz = c2w[8] * dx + c2w[9] * dy + c2w[10] * 1.f;
    if (PTS)
    {
        const float d = depth[pid];
        out[pid] = c2w[3] + d * rx;
        out[n_pix + pid] = c2w[7] + d * ry;
        out[2 * n_pix + pid] = c2w[11] + d * rz;
    }
    else
    {
        out[pid] = rx;
        out[n_pix + pid] = ry;
        out[2 * n_pix + pid] = rz;
    }
}

// raster_state.cu:207-225
__global__ __launch_bounds__(256) void filter_geom_kernel(int64_t n, const int64_t* __restrict__ indices,
                                                          const uint32_t* __restrict__ n_dup,
                                                          const uint2* __restrict__ bbox,
                                                          const uint32_t* __restrict__ qbits,
                                                          const OctBox* __restrict__ oct,
                                                          uint32_t* __restrict__ n_dup_out,
                                                          uint2* __restrict__ bbox_out,
                                                          uint32_t* __restrict__ qbits_out,
                                                          OctBox* __restrict__ oct_out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t src = indices[i];
    n_dup_out[i] = n_dup[src];
    bbox_out[i] = bbox[src];
    qbits_out[i] = qbits[src];
    oct_out[i] = oct[src];
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace svr

using namespace svr;

extern "C" {

int svr_gather_geo_params(int64_t n_care, const int64_t* vox_key, const int64_t* care_idx, const float* grid_pts,
                          float* geo_params, svr_stream_t stream)
{
    SVR_REQUIRE(n_care >= 0 && n_care < (1ll << 40), SVR_ERR_INVALID, "care count out of range");
    if (n_care == 0) return SVR_OK;
    SVR_REQUIRE(vox_key && care_idx && grid_pts && geo_params, SVR_ERR_INVALID, "svr_gather_geo_params: null pointer");
    gather_geo_kernel<<<blocks_for(8 * n_care), 256, 0, stream>>>(n_care, vox_key, care_idx, grid_pts, geo_params);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

int svr_gather_geo_params_bw(int64_t n_care, const int64_t* vox_key, const int64_t* care_idx,
                             const float* dL_dgeo_params, float* dL_dgrid_pts, svr_stream_t stream)
{
    SVR_REQUIRE(n_care >= 0 && n_care < (1ll << 40), SVR_ERR_INVALID, "care count out of range");
    if (n_care == 0) return SVR_OK;
    SVR_REQUIRE(vox_key && care_idx && dL_dgeo_params && dL_dgrid_pts, SVR_ERR_INVALID,
                "svr_gather_geo_params_bw: null pointer");
    gather_geo_bw_kernel<<<blocks_for(8 * n_care), 256, 0, stream>>>(n_care, vox_key, care_idx, dL_dgeo_params,
                                                                      dL_dgrid_pts);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

int svr_gather_feat_params(int64_t n_care, int32_t n_dim, const int64_t* vox_key, const int64_t* care_idx,
                           const float* grid_pts, float* feat_params, svr_stream_t stream)
{
    SVR_REQUIRE(n_dim == 3, SVR_ERR_INVALID, "Only support n_dim=3 now.");
    SVR_REQUIRE(n_care >= 0 && n_care < (1ll << 38), SVR_ERR_INVALID, "care count out of range");
    if (n_care == 0) return SVR_OK;
    SVR_REQUIRE(vox_key && care_idx && grid_pts && feat_params, SVR_ERR_INVALID,
                "svr_gather_feat_params: null pointer");
    gather_feat_kernel<<<blocks_for(8 * n_dim * n_care), 256, 0, stream>>>(n_care, n_dim, vox_key, care_idx, grid_pts,
                                                                           feat_params);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

int svr_gather_feat_params_bw(int64_t n_care, int32_t n_dim, const int64_t* vox_key, const int64_t* care_idx,
                              const float* dL_dfeat_params, float* dL_dgrid_pts, svr_stream_t stream)
{
    SVR_REQUIRE(n_dim == 3, SVR_ERR_INVALID, "Only support n_dim=3");
    SVR_REQUIRE(n_care >= 0 && n_care < (1ll << 38), SVR_ERR_INVALID, "care count out of range");
    if (n_care == 0) return SVR_OK;
    SVR_REQUIRE(vox_key && care_idx && dL_dfeat_params && dL_dgrid_pts, SVR_ERR_INVALID,
                "svr_gather_feat_params_bw: null pointer");
    gather_feat_bw_kernel<<<blocks_for(8 * n_dim * n_care), 256, 0, stream>>>(n_care, n_dim, vox_key, care_idx,
                                                                              dL_dfeat_params, dL_dgrid_pts);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

size_t svr_total_variation_workspace_size(int64_t n_voxels, int32_t n_channels)
{
    const int64_t nc = (n_voxels > 0 ? n_voxels : 0) * (n_channels > 0 ? n_channels : 0);
    return (size_t)((nc + 255) / 256 * 256);
}

int svr_total_variation_bw(int64_t N, int32_t C, const float* grid_pts, const int64_t* vox_key, float weight,
                           const float* vox_size_inv, int32_t no_tv_s, int32_t tv_sparse, void* workspace,
                           float* grid_pts_grad, svr_stream_t stream)
{
    SVR_REQUIRE(N >= 0 && C >= 1 && N * (int64_t)C < (1ll << 40), SVR_ERR_INVALID, "voxel / channel count out of range");
    if (N == 0) return SVR_OK;
    SVR_REQUIRE(grid_pts && vox_key && grid_pts_grad && (no_tv_s || vox_size_inv), SVR_ERR_INVALID,
                "svr_total_variation_bw: null pointer");
    SVR_REQUIRE(!tv_sparse || workspace, SVR_ERR_INVALID, "svr_total_variation_bw: the sparse form needs a workspace");
    const int64_t NC = N * (int64_t)C;
    const uint8_t* mask = nullptr;
    if (tv_sparse)
    {
        tv_mask_kernel<<<blocks_for(NC), 256, 0, stream>>>(NC, C, vox_key, grid_pts_grad, (uint8_t*)workspace);
        SVR_CHECK_LAUNCH();
        mask = (const uint8_t*)workspace;
    }
    if (no_tv_s)
        tv_bw_kernel<true><<<blocks_for(NC), 256, 0, stream>>>(NC, C, grid_pts, vox_key, weight, vox_size_inv, mask,
                                                               grid_pts_grad);
    else
        tv_bw_kernel<false><<<blocks_for(NC), 256, 0, stream>>>(NC, C, grid_pts, vox_key, weight, vox_size_inv, mask,
                                                                grid_pts_grad);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

int svr_is_in_cone(int64_t n, float tanfovx, float tanfovy, float near_plane, const float* w2c, const float* pts,
                   uint8_t* is_in, svr_stream_t stream)
{
    SVR_REQUIRE(n >= 0, SVR_ERR_INVALID, "point count out of range");
    if (n == 0) return SVR_OK;
    SVR_REQUIRE(w2c && pts && is_in, SVR_ERR_INVALID, "svr_is_in_cone: null pointer");
    is_in_cone_kernel<<<blocks_for(n), 256, 0, stream>>>(n, tanfovx, tanfovy, near_plane, w2c, pts, is_in);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

int svr_compute_rd(int32_t W, int32_t H, float cx, float cy, float tanfovx, float tanfovy, const float* c2w,
                   float* rd, svr_stream_t stream)
{
    SVR_REQUIRE(W > 0 && H > 0, SVR_ERR_INVALID, "image size %dx%d must be positive", W, H);
    SVR_REQUIRE(c2w && rd, SVR_ERR_INVALID, "svr_compute_rd: null pointer");
    pixel_rays_kernel<false><<<blocks_for((int64_t)W * H), 256, 0, stream>>>(W, H, cx, cy, tanfovx, tanfovy, c2w,
                                                                              nullptr, rd);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

int svr_depth2pts(int32_t W, int32_t H, float cx, float cy, float tanfovx, float tanfovy, const float* c2w,
                  const float* depth, float* pts, svr_stream_t stream)
{
    SVR_REQUIRE(W > 0 && H > 0, SVR_ERR_INVALID, "image size %dx%d must be positive", W, H);
    SVR_REQUIRE(c2w && depth && pts, SVR_ERR_INVALID, "svr_depth2pts: null pointer");
    pixel_rays_kernel<true><<<blocks_for((int64_t)W * H), 256, 0, stream>>>(W, H, cx, cy, tanfovx, tanfovy, c2w,
                                                                             depth, pts);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

int svr_filter_geom_state(int64_t n_orig, int64_t n_keep, const int64_t* indices, const void* geom_state,
                          void* new_geom_state, svr_stream_t stream)
{
    SVR_REQUIRE(n_orig >= 0 && n_keep >= 0, SVR_ERR_INVALID, "voxel count out of range");
    if (n_keep == 0) return SVR_OK;
    SVR_REQUIRE(indices && geom_state && new_geom_state, SVR_ERR_INVALID, "svr_filter_geom_state: null pointer");
    GeomState src = GeomState::carve((void*)geom_state, n_orig, nullptr);
    GeomState dst = GeomState::carve(new_geom_state, n_keep, nullptr);
    filter_geom_kernel<<<blocks_for(n_keep), 256, 0, stream>>>(n_keep, indices, src.n_dup, src.bbox, src.qbits,
                                                               src.oct, dst.n_dup, dst.bbox, dst.qbits, dst.oct);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

}  // extern "C"
