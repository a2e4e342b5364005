// Host-side helpers for the C-ABI: status codes, the per-thread error string,
// launch checks and the state-buffer layouts.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/svraster_hip.h"
#include "svr_device.h"

namespace svr {

void set_error(const char* fmt, ...);

#define SVR_CHECK_HIP(expr)                                                       \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            ::svr::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                             __FILE__, __LINE__);                                 \
            return SVR_ERR_HIP;                                                   \
        }                                                                         \
    } while (0)

#define SVR_CHECK_LAUNCH()                                                        \
    do {                                                                          \
        hipError_t _e = hipGetLastError();                                        \
        if (_e != hipSuccess) {                                                   \
            ::svr::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e), \
                             __FILE__, __LINE__);                                 \
            return SVR_ERR_HIP;                                                   \
        }                                                                         \
    } while (0)

#define SVR_REQUIRE(cond, code, ...)                                              \
    do {                                                                          \
        if (!(cond)) {                                                            \
            ::svr::set_error(__VA_ARGS__);                                        \
            return (code);                                                        \
        }                                                                         \
    } while (0)

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Carves typed sub-arrays out of one opaque byte buffer (256-byte aligned).
struct Carver
{
    char* base;
    size_t off;
    explicit Carver(void* p) : base((char*)p), off(0) {}
    template <typename T>
    T* take(size_t count)
    {
        off = align_up(off, 256);
        T* p = base ? (T*)(base + off) : nullptr;
        off += sizeof(T) * count;
        return p;
    }
    size_t size() const { return align_up(off, 256); }
};

// Voxels with more list entries than this are expanded / reduced by a whole
// workgroup (they are queued in GeomState::big_list by the duplication pass).
constexpr uint32_t BIG_DUP = 32;
// Voxels with more entries than this (the outside shell: up to ~10^5 entries
// on a 1080p frame) are spread over every workgroup of the launch instead, so
// that no single workgroup becomes its tail; they are queued from the end of
// big_list (count in counters[3]).
constexpr uint32_t HUGE_DUP = 4096;
// Workgroups of grad_reduce_big_kernel; a huge item is split into up to
// HUGE_SPLIT_ROWS / (number of huge items) row ranges, whose partial sums go
// to a scratch of HUGE_SPLIT_ROWS 12-float rows and are added in a fixed order.
constexpr uint32_t HUGE_SPLIT_ROWS = 1024;

// Queues item `id` with nd entries (nd > BIG_DUP) for the large-voxel passes.
__device__ inline void queue_large(uint32_t* __restrict__ big_list, uint32_t* __restrict__ counts, int64_t n_items,
                                   uint32_t nd, uint32_t id)
{
    if (nd > HUGE_DUP)
        big_list[n_items - 1 - (int64_t)atomicAdd(counts + 1, 1u)] = id;
    else
        big_list[atomicAdd(counts, 1u)] = id;
}

constexpr int SCAN_BLOCK = 1024;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

// The pair scan (entry offsets in rank-order position order) uses wider tiles.
constexpr int PSCAN_ITEMS = 16;
constexpr int PSCAN_TILE = SCAN_BLOCK * PSCAN_ITEMS;

// Geometry state (one per preprocess call): per-voxel results + scan scratch.
// Items are voxels (one view) or (view, voxel) pairs u * N + v (a batch).
//
// Entry emission order: the entries of the (item, quadrant) pairs are laid
// out pair after pair in the rank order of the pairs (the order table's
// position), each pair's run covering its tile rectangle row by row.  Every
// pair has at most one entry per tile, so a stable sort of the entries on the
// tile bits alone leaves each tile's entries in rank order: the sorted list
// equals the sort on the full (tile << pos_bits | pos) key.
struct GeomState
{
    uint32_t* n_dup;       // [n] entries of the item (tiles x quadrants)
    uint32_t* run0;        // [n] first entry of the item's run when it has one quadrant
    uint2* bbox;           // [n] packed pixel bbox (x<<16 | y) min / max
    OctBox* oct;           // [n] projected octagon and pixel box (forward candidate bounds)
    uint32_t* qbits;       // [n] camera-quadrant bitset
    uint32_t* block_sums;  // [ceil(n / SCAN_TILE) + 1]
    uint32_t* vis_sums;    // [ceil(n / SCAN_TILE) + 1] in-frustum items per scan block
    uint32_t* counters;    // [8]: 0-1 R (total entries, 64-bit), 2 large-voxel queue length,
                           // 3 huge-voxel queue length (from the end of big_list),
                           // 4-5 entries to sort (R, or 0 when R exceeds a fixed capacity),
                           // 6-7 in-frustum items (n_dup > 0)
    uint32_t* big_list;    // [n] voxels with many entries, expanded by whole workgroups
    uint32_t* pos_off;     // [8n] per view, indexed by rank-order position: the pair's
                           // entry count, then (scanned in place) its first entry
    uint32_t* qoff;        // [8n] first entry of pair (item p, quadrant q) at 8p + q,
                           // written for the items with several quadrants only
    uint32_t* pos_sums;    // [ceil(8n / PSCAN_TILE) + 2]
    static GeomState carve(void* p, int64_t n, size_t* bytes)
    {
        Carver c(p);
        GeomState g;
        g.n_dup = c.take<uint32_t>(n);
        g.run0 = c.take<uint32_t>(n);
        g.bbox = c.take<uint2>(n);
        g.oct = c.take<OctBox>(n);
        g.qbits = c.take<uint32_t>(n);
        g.block_sums = c.take<uint32_t>((n + SCAN_TILE - 1) / SCAN_TILE + 1);
        g.vis_sums = c.take<uint32_t>((n + SCAN_TILE - 1) / SCAN_TILE + 1);
        g.counters = c.take<uint32_t>(8);
        g.big_list = c.take<uint32_t>(n);
        g.pos_off = c.take<uint32_t>(8 * n);
        g.qoff = c.take<uint32_t>(8 * n);
        g.pos_sums = c.take<uint32_t>((8 * n + PSCAN_TILE - 1) / PSCAN_TILE + 2);
        if (bytes) *bytes = c.size();
        return g;
    }
};

// Quadrant of the qi-th set bit of bits.
__host__ __device__ inline uint32_t nth_set_bit(uint32_t bits, uint32_t qi)
{
    for (uint32_t s = 0; s < qi; ++s) bits &= bits - 1u;
    return (uint32_t)__builtin_ctz(bits);
}

// First entry of the run of (item p, quadrant q); bits = the item's quadrants.
__device__ inline uint32_t run_offset(const GeomState& g, size_t p, uint32_t bits, uint32_t q)
{
    return (bits & (bits - 1u)) ? g.qoff[8 * p + q] : g.run0[p];
}

// Records the run offset of (item p, quadrant q).
__device__ inline void set_run_offset(const GeomState& g, size_t p, uint32_t bits, uint32_t q, uint32_t off)
{
    if (bits & (bits - 1u))
        g.qoff[8 * p + q] = off;
    else
        g.run0[p] = off;
}

// Per-octree order table: the rank-order position of every (voxel, quadrant) pair.
struct OrderTable
{
    uint32_t* pos;    // [8N] pair (v*8+q) -> position in the global rank order
    uint32_t* pair;   // [8N] position -> pair
    uint32_t* meta;   // [4]: 0 = pos_bits
    static OrderTable carve(void* p, int64_t N, size_t* bytes)
    {
        Carver c(p);
        OrderTable t;
        t.pos = c.take<uint32_t>(8 * N);
        t.pair = c.take<uint32_t>(8 * N);
        t.meta = c.take<uint32_t>(4);
        if (bytes) *bytes = c.size();
        return t;
    }
};

constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;
constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;

// Keys per thread of a sort pass: 8 up to SORT_SMALL_MAX keys (more, shorter
// tiles: the tile latency and the look-back chain dominate small sorts), 16
// beyond (fewer tiles per workgroup in the persistent loop).
constexpr int SORT_ITEMS_SMALL = 8;
constexpr int64_t SORT_SMALL_MAX = 1ll << 22;
inline int sort_items(int64_t n) { return n <= SORT_SMALL_MAX ? SORT_ITEMS_SMALL : SORT_ITEMS; }
// Tiles of a sort pass over n keys.
inline int64_t sort_blocks(int64_t n)
{
    const int64_t tile = (int64_t)SORT_THREADS * sort_items(n);
    return (n + tile - 1) / tile;
}

// Scratch of the onesweep radix sort (radix_sort_ex) of n keys in at most
// max_passes passes, in 32-bit words: a header (word 0: look-back error
// flag), SORT_HIST_COPIES partial digit histograms of every pass (workgroups
// add into copy blockIdx % SORT_HIST_COPIES, so that no counter takes every
// workgroup's atomic), the look-back status of every (pass, group of
// SORT_LB_GROUP tiles, digit), one 64-bit word each (flag << 32 | count), and
// of every (pass, tile, digit), one 32-bit word each (published bit | count).
// Cleared by radix_sort_ex itself.
constexpr int SORT_HEADER_WORDS = 64;
constexpr int SORT_MAX_PASSES = 8;
constexpr int SORT_HIST_COPIES = 8;
constexpr int SORT_LB_GROUP = 16;
inline int64_t sort_lb_groups(int64_t n) { return (sort_blocks(n > 0 ? n : 1) + SORT_LB_GROUP - 1) / SORT_LB_GROUP; }
inline int64_t sort_scratch_words(int64_t n, int max_passes)
{
    return SORT_HEADER_WORDS + (int64_t)max_passes * RADIX *
                                   (SORT_HIST_COPIES + sort_blocks(n > 0 ? n : 1) + 2 * sort_lb_groups(n));
}
// Passes of the per-frame sorts: tile ids of at most 32 bits in 8-bit digits.
constexpr int SORT_FRAME_PASSES = 4;

// Per-frame binning state: sort keys (double buffer), the sorted entry list.
struct BinningState
{
    uint64_t* keys;     // [R] sorted (tile << pos_bits | pos)
    uint64_t* keys_alt; // [R]
    uint32_t* vox_list; // [R] (q << 29) | vox_id in sorted order
    uint32_t* orig;     // [R] emission-order position of each sorted entry
    uint32_t* orig_alt; // [R]
    uint32_t* counts;   // [sort_scratch_words(R, SORT_FRAME_PASSES)] radix-sort scratch
    static BinningState carve(void* p, int64_t R, size_t* bytes)
    {
        Carver c(p);
        BinningState b;
        b.keys = c.take<uint64_t>(R);
        b.keys_alt = c.take<uint64_t>(R);
        b.vox_list = c.take<uint32_t>(R);
        b.orig = c.take<uint32_t>(R);
        b.orig_alt = c.take<uint32_t>(R);
        b.counts = c.take<uint32_t>(sort_scratch_words(R, SORT_FRAME_PASSES));
        if (bytes) *bytes = c.size();
        return b;
    }
};

// Hit-mask slots of a frame with R entries over T tiles: chunk c (entries
// [range.x + 64c, range.x + 64c + 64)) of tile t uses slot
// floor(range.x / 64) + t + c; these are distinct across tiles and all below
// R / 64 + T + 1.
inline int64_t hit_mask_slots(int64_t R, int64_t T) { return R > 0 ? R / 64 + T + 1 : 0; }

// Per-frame image state (backends/new_cuda/src/raster_state.h:61-68), plus
// the forward's hit masks: per slot, wave and lane (pixel) of a 16x16 tile,
// bit l is set when entry l of the chunk contributed to the pixel (the ray
// meets the box and alpha >= MIN_ALPHA, before early termination).  The
// backward walks exactly those bits instead of re-testing every candidate.
// The masks sit after the reference's fields, so R = 0 carves the prefix.
struct ImageState
{
    uint2* ranges;       // [T] [start, end) of each tile's entries
    uint32_t* tile_last; // [T] range.x + max n_contrib over the tile
    uint32_t* n_contrib; // [H*W]
    uint32_t* order;     // [T] tile queue: longest lists first (workgroup b renders tile order[b])
    uint32_t* order_alt; // [T]
    uint64_t* qkeys;     // [T] queue sort keys
    uint64_t* qkeys_alt; // [T]
    uint32_t* qcounts;   // radix counts of the queue sort
    uint32_t* ctr;       // [8]: 0 tiles in exact_list (cleared with the tile queue)
    uint32_t* exact_list;// [T] tiles with a ray whose 1/rd is not finite (the exact ray-box launches)
    uint64_t* hit_masks; // [hit_mask_slots(R, T)][4 waves][64 lanes]
    static ImageState carve(void* p, int64_t P, int64_t T, size_t* bytes, int64_t R = 0)
    {
        Carver c(p);
        ImageState s;
        s.ranges = c.take<uint2>(T);
        s.tile_last = c.take<uint32_t>(T);
        s.n_contrib = c.take<uint32_t>(P);
        s.order = c.take<uint32_t>(T);
        s.order_alt = c.take<uint32_t>(T);
        s.qkeys = c.take<uint64_t>(T);
        s.qkeys_alt = c.take<uint64_t>(T);
        s.qcounts = c.take<uint32_t>(sort_scratch_words(T, 1));
        s.ctr = c.take<uint32_t>(8);
        s.exact_list = c.take<uint32_t>(T);
        s.hit_masks = c.take<uint64_t>(hit_mask_slots(R, T) * 256);
        if (bytes) *bytes = c.size();
        return s;
    }
};

__host__ __device__ inline int tiles_x(int W) { return (W + 15) / 16; }
__host__ __device__ inline int tiles_y(int H) { return (H + 15) / 16; }

// Stable LSD radix sort of 64-bit keys (optionally with 32-bit values) over
// bits [lo_bit, lo_bit + nbits), in radix_passes(nbits) passes of at most
// RADIX_BITS bits: one histogram launch for every pass, then one onesweep
// launch per pass (decoupled look-back).  The sorted data ends in the buffers
// returned through out_keys / out_vals; counts needs
// sort_scratch_words(n, radix_passes(nbits)) words.  With n_dev, the kernels
// sort min(n, *n_dev) keys (a 64-bit count stored as two 32-bit words on the
// device); n then only sizes the launch.  identity_vals: the values are the
// keys' input positions (0, 1, 2, ...); the first pass generates them, so
// `vals` need not be written beforehand (it is still the ping-pong buffer).
inline int radix_passes(int nbits) { return (nbits + RADIX_BITS - 1) / RADIX_BITS; }
// also_zero / also_zero_n: another buffer (uint32 words) cleared by the same
// launch that clears the sort's scratch.
int radix_sort_ex(uint64_t* keys, uint64_t* keys_alt, uint32_t* vals, uint32_t* vals_alt, int64_t n, int nbits,
                  uint32_t* counts, hipStream_t stream, uint64_t** out_keys, uint32_t** out_vals,
                  const uint32_t* n_dev = nullptr, int lo_bit = 0, bool identity_vals = false,
                  void* also_zero = nullptr, int64_t also_zero_n = 0);

// Entry offsets of the (item, quadrant) pairs in rank order (GeomState's
// emission order): pos_off cleared, filled with each active pair's tile count
// at view u's block [8 N u, 8 N (u + 1)) indexed by rank position, then
// exclusive-scanned in place (only the active positions are written).  The
// two words at g.counters + 2 are cleared by the same launch as pos_off.
// queue_large_items (one view): the same launch queues the items with more
// than BIG_DUP entries in big_list for the duplication.
int pair_offsets_launch(int V, int64_t N, const GeomState& g, const uint32_t* order_pos, hipStream_t stream,
                        bool queue_large_items);
int pair_scan_launch(uint32_t* c, int64_t n, uint32_t* sums, hipStream_t stream);

// Key duplication, sort, tile ranges and the sorted entry list of one frame.
// pairs_ready: pair_prepare already ran on this geometry state (svr_visible_voxels
// with an order table, SVR_FWD_PAIRS_READY).
int binning_forward(int W, int H, int64_t N, const GeomState& g, const OrderTable& ot, int64_t R,
                    BinningState& b, ImageState& img, hipStream_t stream, bool pairs_ready = false);

// Clears the large-voxel queue counters and builds the rank-ordered pair
// offsets of one view (the part of binning that does not need R).
int pair_prepare(int64_t N, const GeomState& g, const OrderTable& ot, hipStream_t stream);

int fill_u32(void* dst, uint32_t value, int64_t count, hipStream_t stream);
int fill_u32_pair(void* dst0, int64_t n0, void* dst1, int64_t n1, uint32_t value, hipStream_t stream);

// A timing-disabled event of the calling thread for the current device,
// created on first use and kept until the process ends.
hipError_t thread_event(hipEvent_t* ev);

// 64 words of pinned host memory of the calling thread (device-to-host
// staging of counts), allocated on first use and kept until the process ends.
hipError_t thread_pinned_words(uint32_t** words);

// Tile queue: order[0..T) = tile ids sorted by length_bucket of their list,
// longest first.  Up to 32768 tiles one workgroup builds it (histogram and
// scatter in LDS); larger batches sort the keys (built in the _alt buffers,
// one stable 8-bit radix pass lands them in qkeys / order).  clear8, when
// given, has 8 words zeroed on the way (the frame's counters).
int tile_queue_launch(int64_t T, const uint2* ranges, uint64_t* qkeys, uint64_t* qkeys_alt, uint32_t* order,
                      uint32_t* order_alt, uint32_t* qcounts, hipStream_t stream, uint32_t* clear8 = nullptr);

// Tile ranges and the sorted entry list from sorted (tile << pos_bits | pos)
// keys; ranges must be zeroed first.
int ranges_launch(const uint64_t* keys, int64_t R, int pos_bits, const uint32_t* order_pair, uint2* ranges,
                  uint32_t* vox_list, hipStream_t stream, const uint32_t* R_dev = nullptr);

// counters[4..5] = counters[0..1] (the entry count) when it is at most cap,
// else 0, and then *overflow = 1 (the flag is never cleared here).
int clamp_count_launch(uint32_t* counters, uint64_t cap, uint32_t* overflow, hipStream_t stream);

// Per-voxel (or per (view, voxel) pair) sums of the per-entry partial
// gradient rows over the item's quadrant runs (GeomState::qoff), quadrants
// ascending and each run in emission order.  A run ending beyond `cap` was
// never emitted and counts as zero.
// Scratch rows of the huge-item split in the single-view backward workspace:
// after the R partial rows (svr_backward_workspace_size).
inline float* huge_rows(float* partial, int64_t R)
{
    return partial + align_up(sizeof(float) * 12 * (size_t)R, 256) / sizeof(float);
}

int grad_reduce_launch(int64_t n_pairs, const GeomState& g, const float* partial, float* dgeos, float* drgbs,
                       float* dsub, float* huge_rows, hipStream_t stream, uint32_t cap = 0xFFFFFFFFu);
// The items with more than BIG_DUP entries only (the queue of the duplication
// pass): their sums go to dgeos / drgbs / dsub at the item's index.
int grad_reduce_big_launch(int64_t n_pairs, const GeomState& g, const float* partial, float* dgeos, float* drgbs,
                           float* dsub, float* huge_rows, hipStream_t stream, uint32_t cap = 0xFFFFFFFFu);

// Argument checks, state carving and (forward) binning shared by the plain and
// the extended compositor entry points.  backward_setup returns 1 after
// zeroing the gradients when there is nothing to composite (N == 0 or R == 0).
int forward_setup(int32_t n_samp, int32_t W, int32_t H, const float* c2w, int64_t N, const float* centers,
                  const float* lengths, const float* geos, const float* rgbs, const void* geom_state,
                  const void* order_table, int64_t R, void* binning_state, void* image_state, float* out_color,
                  float* out_T, GeomState& g, BinningState& b, ImageState& img, hipStream_t stream,
                  int32_t flags = 0);
int backward_setup(int32_t n_samp, int32_t W, int32_t H, const float* c2w, int64_t N, const float* geos,
                   const void* geom_state, int64_t R, const void* binning_state, const void* image_state,
                   const float* out_T, const float* dL_dcolor, const float* dL_dT, void* workspace, float* dgeos,
                   float* drgbs, float* dsub, GeomState& g, BinningState& b, ImageState& img, hipStream_t stream);

// Live stage timing: when enabled, prof_begin/prof_end record a HIP event pair
// on the launch stream around the kernels of one stage.
enum ProfStage { PROF_PREPROCESS = 0, PROF_SCAN = 1, PROF_BINNING = 2, PROF_RENDER_FWD = 3, PROF_RENDER_BWD = 4,
                 PROF_SH_FWD = 5, PROF_SH_BWD = 6, PROF_ADAM = 7, PROF_GRAD_REDUCE = 8,
                 PROF_NUM_STAGES = 9 };
void prof_begin(int stage, hipStream_t stream);
void prof_end(int stage, hipStream_t stream);

}  // namespace svr
