/*
 * svraster_hip.h — C ABI of the MI355X sparse-voxel rasterizer (libsvraster_hip.so).
 *
 * The entry points replace the pybind11 functions of the reference backend
 * module `new_svraster_cuda._C` (backends/new_cuda/binding.cpp:24-57) for the
 * rasterize path.  Conventions:
 *   - every pointer is a device pointer unless its name ends in `_host`;
 *   - matrices (w2c, c2w) are 16 floats, row-major 4x4, on the device;
 *   - images are planar float [C, H, W];
 *   - every call is asynchronous on `stream` unless documented otherwise;
 *   - the library never allocates device memory: callers size the opaque state
 *     buffers with the *_size() queries and pass them in;
 *   - the return value is an svr_status; on failure svr_last_error() holds a
 *     message for the calling thread.
 */
#ifndef SVRASTER_HIP_H_
#define SVRASTER_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* svr_stream_t; /* == hipStream_t */

typedef enum svr_status {
    SVR_OK = 0,
    SVR_ERR_INVALID = 1,     /* bad argument (maps to the reference's Exception / AT_ERROR) */
    SVR_ERR_HIP = 2,         /* a HIP runtime call or launch failed */
    SVR_ERR_UNSUPPORTED = 3, /* option not implemented by this build */
    SVR_ERR_CAPACITY = 4     /* a preallocated buffer is too small */
} svr_status;

/* Message of the last failed call on this thread ("" if none). */
const char* svr_last_error(void);

/* Read-only constants exported by the reference module (binding.cpp:51-57; config.h:14-34). */
typedef struct svr_constants {
    int32_t max_num_levels;       /* MAX_NUM_LEVELS = 21 */
    float step_sz_scale;          /* STEP_SZ_SCALE = 100 */
    int32_t max_n_samp;           /* MAX_N_SAMP = 3 */
    int32_t max_sort_key_bits;    /* MAX_SORT_KEY_BITS = 128 */
    int32_t max_packed_image_dim; /* MAX_PACKED_IMAGE_DIM = 65535 */
    int64_t max_render_tiles;     /* MAX_RENDER_TILES */
    int32_t tile_x, tile_y;       /* BLOCK_X, BLOCK_Y = 16 */
} svr_constants;
int svr_get_constants(svr_constants* out_host);

/* ---- state-buffer sizes (bytes) ---------------------------------------- */
/* Geometry state of `rasterize_preprocess` (raster_state.h:43-59). */
size_t svr_geom_state_size(int64_t n_voxels);
/* Per-octree rank-order table (no reference counterpart: replaces the per-frame
 * 63-bit order-rank part of the 128-bit sort key, forward.cu:666-730). */
size_t svr_order_table_size(int64_t n_voxels);
size_t svr_order_table_scratch_size(int64_t n_voxels);
/* Binning state of `rasterize_voxels` (raster_state.h:70-82) for R entries. */
size_t svr_binning_state_size(int64_t num_rendered);
/* Image state of a W x H frame with R entries: ranges, tile_last, n_contrib
 * (raster_state.h:61-68) and the forward's per-pixel hit masks that the
 * backward walks (64 bits per pixel and 64-entry chunk of its tile). */
size_t svr_image_state_size(int32_t width, int32_t height, int64_t num_rendered);

/* ---- forward path ------------------------------------------------------- */
/* Replaces `_C.rasterize_preprocess` (preprocess.cu:136-199): per voxel the
 * near-sphere cull, the 8-corner projection, the packed pixel bbox, the
 * camera-quadrant bitset and n_duplicates = tiles x popcount(quadrants).
 * Writes out_n_duplicates[N] (int32) and the geometry state. */
int svr_rasterize_preprocess(int32_t width, int32_t height, float tanfovx, float tanfovy, float cx, float cy,
                             const float* w2c, const float* c2w, float near_plane, int64_t n_voxels,
                             const float* vox_centers, const float* vox_lengths, int32_t* out_n_duplicates,
                             void* geom_state, svr_stream_t stream);

/* The number of rendered entries R, the sum of n_duplicates (forward.cu:
 * 649-663, whose inclusive scan places each voxel's entries; here the
 * forward places them in rank order itself).  Synchronises `stream` to
 * return R to the host, as the reference does before sizing its binning
 * buffer. */
int svr_num_rendered(int64_t n_voxels, void* geom_state, svr_stream_t stream, int64_t* num_rendered_host);

/* The in-frustum voxel list and R with one synchronisation: replaces the two
 * host round trips of rasterize_voxels between preprocess and the raster,
 * `torch.where(n_duplicates > 0)[0]` (renderer.py:81) and num_rendered
 * (forward.cu:649-663).  Writes the ascending ids of the voxels with
 * n_duplicates > 0 into idx_out[0 .. *num_visible_host) (int64, room for
 * n_voxels ids) and returns R and that count on the host.  With an order
 * table (svr_build_order_table of the same octree; may be NULL) it also
 * queues the part of the forward's binning that does not need R, to run on
 * the device while the host goes on: a svr_rasterize_forward(_ext) of this
 * geometry state then passes SVR_FWD_PAIRS_READY. */
int svr_visible_voxels(int64_t n_voxels, void* geom_state, const void* order_table, int64_t* idx_out,
                       svr_stream_t stream, int64_t* num_rendered_host, int64_t* num_visible_host);

/* flags of svr_rasterize_binning and svr_rasterize_forward(_ext) */
enum
{
    SVR_FWD_PAIRS_READY = 1, /* svr_visible_voxels already prepared the binning with this order table */
    SVR_FWD_BINNED = 2       /* svr_rasterize_binning already filled the binning and image states */
};

/* Builds the rank-order table of an octree: the position of every
 * (voxel, quadrant) pair in the order of the reference's 63-bit order rank
 * octpath ^ repeat(q) (auxiliary.h:57-60), ties by (voxel, quadrant).  Valid
 * until octree_paths changes.  `scratch` has svr_order_table_scratch_size()
 * bytes.  Synchronises `stream` once (reads the finest used level). */
int svr_build_order_table(int64_t n_voxels, const int64_t* octree_paths, void* order_table, void* scratch,
                          svr_stream_t stream);

/* The binning half of the forward (forward.cu:665-753: key duplication, sort,
 * tile ranges) on its own, so that it can be queued before the per-voxel
 * colours exist (it needs only the geometry): a svr_rasterize_forward(_ext)
 * on the same states then passes SVR_FWD_BINNED and only composites.  flags:
 * SVR_FWD_PAIRS_READY as for the forward. */
int svr_rasterize_binning(int32_t width, int32_t height, int64_t n_voxels, const void* geom_state,
                          const void* order_table, int64_t num_rendered, void* binning_state, void* image_state,
                          int32_t flags, svr_stream_t stream);

/* Replaces `_C.rasterize_voxels` (forward.cu:792-879) for n_samp in [1,3]
 * without depth / normal / distortion / max_w: key duplication, sort, tile
 * ranges and the front-to-back compositor.  Outputs out_color [3,H,W] and
 * out_T [1,H,W]; fills the binning and image states used by the backward. */
int svr_rasterize_forward(int32_t n_samp, int32_t width, int32_t height, float tanfovx, float tanfovy, float cx,
                          float cy, const float* c2w, float bg_color, int64_t n_voxels, const float* vox_centers,
                          const float* vox_lengths, const float* geos, const float* rgbs, const void* geom_state,
                          const void* order_table, int64_t num_rendered, void* binning_state, void* image_state,
                          float* out_color, float* out_T, int32_t flags, svr_stream_t stream);

/* Workspace of svr_rasterize_backward for R entries: one 12-float partial
 * gradient row per list entry. */
size_t svr_backward_workspace_size(int64_t num_rendered);
/* Replaces `_C.rasterize_voxels_backward` (backward.cu:654-772): reverse
 * traversal with T recovered from the stored final T and n_contrib.  Gradients
 * are summed per list entry in `workspace`, then per voxel over its entries.
 * Writes every row of dL_dgeos [N,8], dL_drgbs [N,3], subdiv_p_bw [N,1]
 * (zero for voxels with no entry).  geos, dL_dgeos and workspace must be
 * 16-byte aligned.  dL_dout_T may be NULL (no loss on T: the same as zeros;
 * the batch backward below takes NULL the same way). */
int svr_rasterize_backward(int32_t n_samp, int32_t width, int32_t height, float tanfovx, float tanfovy, float cx,
                           float cy, const float* c2w, float bg_color, int64_t n_voxels, const float* vox_centers,
                           const float* vox_lengths, const float* geos, const float* rgbs, const void* geom_state,
                           int64_t num_rendered, const void* binning_state, const void* image_state,
                           const float* out_T, const float* dL_dout_color, const float* dL_dout_T, void* workspace,
                           float* dL_dgeos, float* dL_drgbs, float* subdiv_p_bw, svr_stream_t stream);

/* No counterpart in the reference: svr_rasterize_backward for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and outputs
 * without the workspace: the per-entry gradient rows live in a fixed device
 * array of the library, sized for SVR_NOALLOC_MAX_ENTRIES entries (192 MiB,
 * held by every GPU the library runs on until the process ends).  Calls no
 * allocator; num_rendered above the maximum returns SVR_ERR_CAPACITY.  All
 * calls on one GPU share that array: calls that overlap on one GPU (other
 * streams or threads) race and must not be issued. */
#define SVR_NOALLOC_MAX_ENTRIES (1ll << 22)
int svr_rasterize_backward_noalloc(int32_t n_samp, int32_t width, int32_t height, float tanfovx, float tanfovy,
                                   float cx, float cy, const float* c2w, float bg_color, int64_t n_voxels,
                                   const float* vox_centers, const float* vox_lengths, const float* geos,
                                   const float* rgbs, const void* geom_state, int64_t num_rendered,
                                   const void* binning_state, const void* image_state, const float* out_T,
                                   const float* dL_dout_color, const float* dL_dout_T, float* dL_dgeos,
                                   float* dL_drgbs, float* subdiv_p_bw, svr_stream_t stream);

/* Replaces `_C.rasterize_voxels` (forward.cu:792-879) with its optional
 * outputs (forward.cu:262-394): need_depth writes the mean depth (channel 0)
 * and the median depth (channel 2) of out_depth [3,H,W], need_distortion the
 * distortion sum sum_i w_i (contracted(a_i) + contracted(b_i)) / 2 (channel 1);
 * channels not asked for are written as zero (out_depth may be null when
 * neither is set).  need_normal writes out_normal [3,H,W]; track_max_w zeroes
 * max_w [N] and stores each voxel's largest blending weight.  Colour, T and
 * the states are those of svr_rasterize_forward. */
int svr_rasterize_forward_ext(int32_t n_samp, int32_t width, int32_t height, float tanfovx, float tanfovy, float cx,
                              float cy, const float* c2w, float bg_color, int32_t need_depth, int32_t need_distortion,
                              int32_t need_normal, int32_t track_max_w, int64_t n_voxels, const float* vox_centers,
                              const float* vox_lengths, const float* geos, const float* rgbs, const void* geom_state,
                              const void* order_table, int64_t num_rendered, void* binning_state, void* image_state,
                              float* out_color, float* out_T, float* out_depth, float* out_normal, float* max_w,
                              int32_t flags, svr_stream_t stream);

/* Replaces `_C.rasterize_voxels_backward` (backward.cu:654-772) with the
 * optional terms (backward.cu:160-510): need_depth back-propagates
 * dL_dout_depth channel 0 (mean depth), need_normal dL_dout_normal [3,H,W];
 * lambda_dist > 0 adds the distortion gradient (out_depth: the forward's
 * [3,H,W] output with need_distortion), lambda_R_concen > 0 the colour
 * concentration toward gt_color [3,H,W], lambda_ascending > 0 the
 * density-ascending regularizer; each regularizer is weighted by
 * lambda / (width * height).  Same workspace, alignment and outputs as
 * svr_rasterize_backward. */
int svr_rasterize_backward_ext(int32_t n_samp, int32_t width, int32_t height, float tanfovx, float tanfovy, float cx,
                               float cy, const float* c2w, float bg_color, int64_t n_voxels, const float* vox_centers,
                               const float* vox_lengths, const float* geos, const float* rgbs,
                               const void* geom_state, int64_t num_rendered, const void* binning_state,
                               const void* image_state, const float* out_T, const float* dL_dout_color,
                               const float* dL_dout_depth, const float* dL_dout_normal, const float* dL_dout_T,
                               float lambda_R_concen, const float* gt_color, float lambda_ascending,
                               float lambda_dist, int32_t need_depth, int32_t need_normal, const float* out_depth,
                               void* workspace, float* dL_dgeos, float* dL_drgbs, float* subdiv_p_bw,
                               svr_stream_t stream);

/* Replaces `_C.unpack_ImageState` (raster_state.cu:150-205): copies ranges
 * [T,2], tile_last [T] and n_contrib [H,W] (int32, device) out of the image state. */
int svr_unpack_image_state(int32_t width, int32_t height, const void* image_state, int32_t* ranges,
                           int32_t* tile_last, int32_t* n_contrib, svr_stream_t stream);

/* Parity export: the sorted keys in the reference's 128-bit form
 * hi = tile >> 1, lo = (tile << 63) | (octpath ^ repeat(q)) (auxiliary.h:67-76),
 * and the sorted values (q << 29) | vox_id. */
int svr_export_sort_keys(int32_t width, int32_t height, int64_t num_rendered, const void* binning_state,
                         const void* order_table, const int64_t* octree_paths, int64_t n_voxels, uint64_t* keys_hi,
                         uint64_t* keys_lo, uint32_t* values, svr_stream_t stream);

/* Parity export of the geometry state: n_duplicates [N] (u32), packed bboxes
 * [N,2] (x << 16 | y for min and max) and camera-quadrant bitsets [N]
 * (raster_state.h:43-59; the reference only reads them on the device). */
int svr_unpack_geom_state(int64_t n_voxels, const void* geom_state, uint32_t* n_dup, uint32_t* bbox,
                          uint32_t* qbits, svr_stream_t stream);

/* ---- colour ----------------------------------------------------------- */
/* Replaces `_C.sh_compute` (sh_compute.cu:244-270): rgbs [N,3] = SH(deg) of
 * the direction cam_pos -> centre, +0.5, clamped at 0; rows outside idx are
 * zero.  n_idx == 0 evaluates all voxels.  M = 1 + shs.shape[1]. */
int svr_sh_compute(int32_t deg, int64_t n_idx, const int64_t* idx, int64_t n_voxels, const float* vox_centers,
                   const float* cam_pos, const float* sh0, const float* shs, int32_t M, float* rgbs,
                   svr_stream_t stream);
/* Replaces `_C.sh_compute_bw` (sh_compute.cu:272-299), no clamp mask. */
int svr_sh_compute_bw(int32_t deg, int32_t M, int64_t n_idx, const int64_t* idx, int64_t n_voxels,
                      const float* vox_centers, const float* cam_pos, const float* dL_drgbs, float* dL_dsh0,
                      float* dL_dshs, svr_stream_t stream);

/* svr_sh_compute / svr_sh_compute_bw over the in-frustum voxels of a
 * preprocess geometry state (n_dup > 0: the list svr_visible_voxels returns
 * for it), without the index list: the same rgbs / gradients as the calls with
 * that list (sh_compute.cu:244-299 with idx = renderer.py:81's in-frustum
 * ids), minus the zero fill and marking pass they need.  M must be 16 and
 * sh0, shs, rgbs (dL_dsh0, dL_dshs) 16-byte aligned; SVR_ERR_INVALID
 * otherwise. */
int svr_sh_compute_visible(int32_t deg, int64_t n_voxels, const void* geom_state, const float* vox_centers,
                           const float* cam_pos, const float* sh0, const float* shs, int32_t M, float* rgbs,
                           svr_stream_t stream);
int svr_sh_compute_bw_visible(int32_t deg, int32_t M, int64_t n_voxels, const void* geom_state,
                              const float* vox_centers, const float* cam_pos, const float* dL_drgbs, float* dL_dsh0,
                              float* dL_dshs, svr_stream_t stream);

/* ---- optimizer -------------------------------------------------------- */
/* Replaces `_C.unbiased_adam_step` (adam_step.cu:106-137); sparse skips grad == 0. */
int svr_unbiased_adam_step(int32_t sparse, int64_t numel, float* param, const float* grad, float* exp_avg,
                           float* exp_avg_sq, double step, double lr, double beta1, double beta2, float eps,
                           svr_stream_t stream);
/* The bias terms of step `step` exactly as svr_unbiased_adam_step computes
 * them on the host (adam_step.cu:119-124): step_size = lr / (1 - beta1^step),
 * rb2 = 1 / sqrt(1 - beta2^step), each rounded to float once. */
int svr_adam_bias_terms(double step, double lr, double beta1, double beta2, float* step_size_host,
                        float* rb2_host);
/* Capture-safe form of svr_unbiased_adam_step for hipGraph replay (no
 * reference counterpart): increments the step count *step_dev (int32, device)
 * and takes the bias terms of the new count from `bias_table` (device:
 * int32 first step, int32 count, then `count` (step_size, rb2) float pairs
 * filled with svr_adam_bias_terms), so the update is bit-equal to the
 * host-step form; a count outside the table falls back to pow() in double on
 * the device.  bias_table may be NULL (always the fallback).  skip_dev
 * (uint32, device, may be NULL): when nonzero the call changes nothing, the
 * step count included (a frame over its entry capacity).  coef_dev is 3
 * floats of device scratch. */
int svr_unbiased_adam_step_device(int32_t sparse, int64_t numel, float* param, const float* grad, float* exp_avg,
                                  float* exp_avg_sq, int32_t* step_dev, const void* bias_table,
                                  const uint32_t* skip_dev, float* coef_dev, double lr, double beta1, double beta2,
                                  float eps, svr_stream_t stream);
/* SparseAdam.step over several tensors in one launch (sparse_adam.py:43-72
 * calls the per-tensor step once per parameter): tensor i of `count` takes
 * the update of svr_unbiased_adam_step (biased = 0; step count steps[i]) or
 * svr_biased_adam_step (biased = 1) with its own lr, betas and eps, bit-equal
 * to those calls.  The arrays are host arrays of `count` entries; the tensors
 * are device pointers. */
int svr_adam_step_multi(int32_t biased, int32_t sparse, int32_t count, const int64_t* numels, float* const* params,
                        const float* const* grads, float* const* exp_avgs, float* const* exp_avg_sqs,
                        const double* steps, const double* lrs, const double* beta1s, const double* beta2s,
                        const float* eps, svr_stream_t stream);
/* Replaces `_C.biased_adam_step` (adam_step.cu:139-162). */
int svr_biased_adam_step(int32_t sparse, int64_t numel, float* param, const float* grad, float* exp_avg,
                         float* exp_avg_sq, float lr, float beta1, float beta2, float eps, svr_stream_t stream);

/* ---- grid-point parameters ------------------------------------------ */
/* Replaces `_C.gather_triinterp_geo_params` (geo_params_gather.cu:22-45,
 * 134-153): geo_params[idx, k] = grid_pts[vox_key[idx, k]] for the n_care voxels
 * idx of care_idx (other rows untouched).  vox_key [N,8] int64. */
int svr_gather_geo_params(int64_t n_care, const int64_t* vox_key, const int64_t* care_idx, const float* grid_pts,
                          float* geo_params, svr_stream_t stream);
/* Replaces `_C.gather_triinterp_geo_params_bw` (geo_params_gather.cu:47-72,
 * 155-173): adds dL_dgeo_params[idx, k] into dL_dgrid_pts[vox_key[idx, k]]
 * (float atomics; the caller zeroes dL_dgrid_pts). */
int svr_gather_geo_params_bw(int64_t n_care, const int64_t* vox_key, const int64_t* care_idx,
                             const float* dL_dgeo_params, float* dL_dgrid_pts, svr_stream_t stream);
/* Replaces `_C.gather_triinterp_feat_params(_bw)` (geo_params_gather.cu:
 * 74-132, 175-224) for n_dim = 3 channels per grid point: feat [N,8,3]. */
int svr_gather_feat_params(int64_t n_care, int32_t n_dim, const int64_t* vox_key, const int64_t* care_idx,
                           const float* grid_pts, float* feat_params, svr_stream_t stream);
int svr_gather_feat_params_bw(int64_t n_care, int32_t n_dim, const int64_t* vox_key, const int64_t* care_idx,
                              const float* dL_dfeat_params, float* dL_dgrid_pts, svr_stream_t stream);
/* Replaces `_C.total_variation_bw` (tv_compute.cu:23-111): adds the gradient
 * of the weighted L1 differences along each voxel's 12 corner-cube edges into
 * grid_pts_grad [M, C] in place (weight x 0.01 / voxel size unless no_tv_s).
 * tv_sparse: only (voxel, channel) items with a non-zero corner gradient take
 * part, decided on grid_pts_grad as it is before this call (workspace:
 * svr_total_variation_workspace_size bytes). */
size_t svr_total_variation_workspace_size(int64_t n_voxels, int32_t n_channels);
int svr_total_variation_bw(int64_t n_voxels, int32_t n_channels, const float* grid_pts, const int64_t* vox_key,
                           float weight, const float* vox_size_inv, int32_t no_tv_s, int32_t tv_sparse,
                           void* workspace, float* grid_pts_grad, svr_stream_t stream);

/* ---- camera utilities (utils.cu:23-102, 188-245) ---------------------- */
/* Replaces `_C.is_in_cone`: is_in [n] (bool bytes) = the point is in front of
 * `near` and inside the camera's field-of-view cone. */
int svr_is_in_cone(int64_t n, float tanfovx, float tanfovy, float near_plane, const float* w2c, const float* pts,
                   uint8_t* is_in, svr_stream_t stream);
/* Replaces `_C.compute_rd`: per-pixel world ray directions (unnormalised), [3,H,W]. */
int svr_compute_rd(int32_t width, int32_t height, float cx, float cy, float tanfovx, float tanfovy,
                   const float* c2w, float* rd, svr_stream_t stream);
/* Replaces `_C.depth2pts`: camera centre + depth x ray direction, [3,H,W]. */
int svr_depth2pts(int32_t width, int32_t height, float cx, float cy, float tanfovx, float tanfovy,
                  const float* c2w, const float* depth, float* pts, svr_stream_t stream);
/* Replaces `_C.filter_geomState` (raster_state.cu:207-260): a geometry state
 * of the n_keep voxels `indices` (n_duplicates, bboxes, quadrant bits). */
int svr_filter_geom_state(int64_t n_orig, int64_t n_keep, const int64_t* indices, const void* geom_state,
                          void* new_geom_state, svr_stream_t stream);

/* ---- octree utilities (utils.cu:106-189) ---------------------------- */
int svr_voxel_order_rank(int64_t n_voxels, const int64_t* octree_paths, int64_t* order_ranks /* [8,N] */,
                         svr_stream_t stream);
int svr_ijk_2_octpath(int64_t n, const int64_t* ijk, const int8_t* octlevel, int64_t* octpath,
                      svr_stream_t stream);
int svr_octpath_2_ijk(int64_t n, const int64_t* octpath, const int8_t* octlevel, int64_t* ijk,
                      svr_stream_t stream);

/* ---- multi-view batch ---------------------------------------------------
 * No reference counterpart: the reference renders one camera per call
 * (renderer.py:39-127).  These entry points rasterize V cameras of one octree
 * in one pass per stage, for training steps whose loss sums over views.  Each
 * view's results equal what the single-view entry points give for that
 * camera; gradients of shared parameters are the sums over the views.
 * Per (view, voxel) arrays are view-major: index u * N + v. */

/* One camera of a batch.  The caller keeps an array of these in host memory
 * (for validation and sizes) and the same bytes in device memory (read by the
 * kernels).  tile_base = sum of the tile counts (ceil(W / tile width) *
 * ceil(H / tile height)) of the previous views; pix_base = sum of their pixel
 * counts. */
typedef struct svr_view {
    int32_t width, height;
    float tanfovx, tanfovy, cx, cy, near_plane, bg_color;
    float w2c[12]; /* first 3 rows of the row-major world-to-camera matrix */
    float c2w[12]; /* first 3 rows of the row-major camera-to-world matrix */
    int32_t tile_base;
    int32_t tile_shape; /* SVR_TILE_*: binning tile of the batch, the same in every view */
    int32_t bwd_window; /* backward LDS depth in 64-entry chunks: 2, 4, 6 or 8 (0: 4); the same in every view */
    int32_t bwd_sync;   /* backward wave-sum threshold in pixels: 4, 8 or 16 (0: 8); the same in every view */
    int64_t pix_base;
} svr_view;

/* Binning tiles of a batch (the single-view entry points use the reference's
 * 16x16).  Images and gradients do not depend on the tile; n_contrib and the
 * tile ranges are relative to the batch's tiles.
 *
 * Backward depth of a batch: the adjoint stages bwd_window chunks of 64 list
 * entries (payload and gradient accumulators) in LDS at a time, and an entry
 * used by at least bwd_sync pixels of a wave has its gradient terms summed
 * across the wave at once instead of per pixel.  Pure performance choices:
 * gradients do not depend on them beyond float summation order.  n_samp 2
 * and 3 always use the defaults. */
enum { SVR_TILE_16X16 = 0, SVR_TILE_8X8 = 1, SVR_TILE_32X8 = 2 };

/* Sizes: geometry state of V views (its count runs over V * N entries, so
 * svr_num_rendered(V * N, ...) returns the batch's total entry count R), the
 * batch image state for T_total tiles, P_total pixels and R entries (the
 * forward's hit masks, as in svr_image_state_size), and the backward
 * workspace (per-entry partial rows plus per-(view, voxel) gradient rows). */
size_t svr_batch_geom_state_size(int32_t n_views, int64_t n_voxels);
size_t svr_batch_image_state_size(int64_t total_tiles, int64_t total_pixels, int64_t num_rendered);
size_t svr_batch_backward_workspace_size(int32_t n_views, int64_t n_voxels, int64_t num_rendered);

/* Preprocess of every (view, voxel) pair: out_n_duplicates [V, N] (int32) and
 * the geometry state, as svr_rasterize_preprocess per view. */
int svr_batch_preprocess(int32_t n_views, const svr_view* views_host, const svr_view* views_dev, int64_t n_voxels,
                         const float* vox_centers, const float* vox_lengths, int32_t* out_n_duplicates,
                         void* geom_state, svr_stream_t stream);

/* SH colours of every (view, voxel) pair with entries (rgbs [V, N, 3], zero
 * rows elsewhere), reading each voxel's coefficients once for all views.
 * Same arithmetic per pair as svr_sh_compute.  M = 16 (degree-3 layout). */
int svr_batch_sh_compute(int32_t deg, int32_t n_views, const svr_view* views_dev, int64_t n_voxels,
                         const float* vox_centers, const void* geom_state, const float* sh0, const float* shs,
                         float* rgbs, svr_stream_t stream);
/* dL_dsh0 [N, 3] and dL_dshs [N, 15, 3] summed over the views (in view
 * order) from dL_drgbs [V, N, 3], as svr_sh_compute_bw per view. */
int svr_batch_sh_compute_bw(int32_t deg, int32_t n_views, const svr_view* views_dev, int64_t n_voxels,
                            const float* vox_centers, const void* geom_state, const float* dL_drgbs,
                            float* dL_dsh0, float* dL_dshs, svr_stream_t stream);

/* Key duplication, one sort over all views' entries, tile ranges, and the
 * forward compositor as one persistent launch over a queue of all views'
 * tiles (longest lists first).  out_color holds each view's planar [3, H, W]
 * block at 3 * pix_base, out_T its [H, W] block at pix_base. */
int svr_batch_rasterize_forward(int32_t n_samp, int32_t n_views, const svr_view* views_host,
                                const svr_view* views_dev, int64_t n_voxels, const float* vox_centers,
                                const float* vox_lengths, const float* geos, const float* rgbs,
                                const void* geom_state, const void* order_table, int64_t num_rendered,
                                void* binning_state, void* image_state, float* out_color, float* out_T,
                                svr_stream_t stream);

/* Reverse-order adjoint over all views (persistent launch), then per voxel:
 * dL_dgeos [N, 8] and subdiv_p_bw [N] summed over the views, dL_drgbs
 * [V, N, 3] per view.  workspace: svr_batch_backward_workspace_size bytes. */
int svr_batch_rasterize_backward(int32_t n_samp, int32_t n_views, const svr_view* views_host,
                                 const svr_view* views_dev, int64_t n_voxels, const float* vox_centers,
                                 const float* vox_lengths, const float* geos, const float* rgbs,
                                 const void* geom_state, int64_t num_rendered, const void* binning_state,
                                 const void* image_state, const float* out_T, const float* dL_dout_color,
                                 const float* dL_dout_T, void* workspace, float* dL_dgeos, float* dL_drgbs,
                                 float* subdiv_p_bw, svr_stream_t stream);

/* Capture-safe forms for hipGraph replay of a batch training step: no host
 * synchronisation.  svr_scan_entries runs the sum of svr_num_rendered and
 * leaves the entry count on the device.  The capacity forms take buffers
 * sized for `capacity` entries (binning state, backward workspace) and read
 * the count on the device.  overflow_dev is 2 uint32 words on the device:
 * when the count exceeds the capacity, overflow_dev[0] is set to 1 and the
 * frame renders no entries (background colour, T = 1, zero gradients);
 * overflow_dev[1] is rewritten every frame (1: this frame was over the
 * capacity, else 0; svr_unbiased_adam_step_device takes it as skip_dev).
 * Word 0 is only ever set, never cleared: the caller zeroes it, checks it
 * after any number of replays and re-sizes. */
int svr_scan_entries(int64_t n_counts, void* geom_state, svr_stream_t stream);
int svr_batch_rasterize_forward_capacity(int32_t n_samp, int32_t n_views, const svr_view* views_host,
                                         const svr_view* views_dev, int64_t n_voxels, const float* vox_centers,
                                         const float* vox_lengths, const float* geos, const float* rgbs,
                                         const void* geom_state, const void* order_table, int64_t capacity,
                                         void* binning_state, void* image_state, float* out_color, float* out_T,
                                         uint32_t* overflow_dev, svr_stream_t stream);
int svr_batch_rasterize_backward_capacity(int32_t n_samp, int32_t n_views, const svr_view* views_host,
                                          const svr_view* views_dev, int64_t n_voxels, const float* vox_centers,
                                          const float* vox_lengths, const float* geos, const float* rgbs,
                                          const void* geom_state, int64_t capacity, const void* binning_state,
                                          const void* image_state, const float* out_T, const float* dL_dout_color,
                                          const float* dL_dout_T, void* workspace, float* dL_dgeos,
                                          float* dL_drgbs, float* subdiv_p_bw, svr_stream_t stream);

/* Copies ranges [T_total, 2], tile_last [T_total] and n_contrib [P_total] out
 * of a batch image state. */
int svr_batch_unpack_image_state(int64_t total_tiles, int64_t total_pixels, const void* image_state,
                                 int32_t* ranges, int32_t* tile_last, int32_t* n_contrib, svr_stream_t stream);

/* Per-voxel statistics over many cameras, replacing the per-camera loops of
 * SRC/sparse_voxel_gears/adaptive.py:190-221 (compute_training_stat) and
 * NC/new_svraster_cuda/renderer.py:408-438 (mark_max_samp_rate, mark_near).
 *
 * svr_batch_rasterize_max_w: the batch forward as a weights-only pass (no
 * colour; rgbs of svr_batch_rasterize_forward not needed), writing out_T and
 * pair_max_w [V, N], each (view, voxel) pair's largest blending weight (the
 * per-camera max_w of rasterize_voxels with track_max_w; zeroed here).
 *
 * svr_batch_training_stat: per voxel over the views in order, from
 * pair_max_w: max_w = max_u w_u; min_samp_interval = min over the views with
 * w_u > 0 of ((c - position_u) . lookat_u) * pix_size[u] (1e30 if none);
 * view_cnt = number of such views (as float).  pix_size [V] (device):
 * 2 * tanfovx / width per view, rounded to float once.
 *
 * svr_batch_mark_samp: per voxel over the views with pair_n_dup > 0 (the
 * batch preprocess, whose near plane is each view's near_plane):
 * max_samp_rate = max of length / (zdist * pix_size[u]) over views with
 * zdist > near_plane (0 if none); is_near = 1 when some such view has
 * zdist <= near_plane + length.  Either output may be null. */
int svr_batch_rasterize_max_w(int32_t n_samp, int32_t n_views, const svr_view* views_host, const svr_view* views_dev,
                              int64_t n_voxels, const float* vox_centers, const float* vox_lengths,
                              const float* geos, const void* geom_state, const void* order_table,
                              int64_t num_rendered, void* binning_state, void* image_state, float* out_T,
                              float* pair_max_w, svr_stream_t stream);
int svr_batch_training_stat(int32_t n_views, const svr_view* views_dev, int64_t n_voxels, const float* vox_centers,
                            const float* pix_size, const float* pair_max_w, float* max_w, float* min_samp_interval,
                            float* view_cnt, svr_stream_t stream);
int svr_batch_mark_samp(int32_t n_views, const svr_view* views_dev, int64_t n_voxels, const float* vox_centers,
                        const float* vox_lengths, const float* pix_size, const int32_t* pair_n_dup,
                        float* max_samp_rate, uint8_t* is_near, svr_stream_t stream);

/* Live stage timing (no reference counterpart; measurement support).  While
 * enabled, each entry point records a HIP event pair on its stream around the
 * kernels of its stage: 0 preprocess, 1 scan, 2 binning (duplicate + sort +
 * ranges), 3 forward compositor kernel, 4 backward compositor kernel, 5 SH
 * forward, 6 SH backward, 7 Adam, 8 per-voxel gradient reduction.  Enabling
 * resets the records. */
int svr_profile_enable(int32_t on);
/* Synchronises on the recorded events of `stage`, returns their summed
 * duration and count, and clears them. */
int svr_profile_collect(int32_t stage, double* total_ms_host, int64_t* count_host);

/* Waits for `stream` and reports any pending device error (the reference's
 * CHECK_CUDA(debug), auxiliary.h:394-401). */
int svr_debug_sync(svr_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* SVRASTER_HIP_H_ */
