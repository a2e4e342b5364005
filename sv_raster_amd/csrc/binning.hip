// Binning: per-octree rank-order table, key duplication, LSD radix sort,
// tile ranges and the sorted entry list.
//
// Reference: backends/new_cuda/src/forward.cu:507-604 (duplicateWithKeys,
// identifyTileRanges), :666-753 (128-bit CUB sort + ranges),
// auxiliary.h:33-127 (order rank, key and value encodings).
//
// Design: the reference sorts 76..77-bit keys (tile << 63 | order_rank).  The
// order rank octpath ^ repeat(q) depends only on the octree, so its order is
// computed once per octree as a table pos[v*8+q] (the rank position of the
// pair, ties by (v, q) exactly as the reference's stable sort orders them).
// Per frame the key is (tile << pos_bits) | pos, which sorts into exactly the
// reference order with 30..40 significant bits instead of 76.
#include <algorithm>
#include <mutex>
#include <utility>
#include <vector>

#include "voxel_ops.h"
#include "svr_host.h"

namespace svr {


// Lanes of `active` whose digit d (of DBITS bits) equals this lane's.
template <int DBITS>
__device__ inline uint64_t match_digit(uint32_t d, uint64_t active)
{
    uint64_t m = active;
#pragma unroll
    for (int b = 0; b < DBITS; ++b)
    {
        const uint32_t bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        m &= bit ? bal : ~bal;
    }
    return m;
}

// Adds one count per lane to h[d] with one LDS atomic per run of equal
// digits on consecutive lanes (the keys arrive with long runs of one tile,
// which as single-lane atomics would all queue on one LDS address).  The
// valid lanes must be a prefix of the wave.
__device__ inline void add_digit_runs(uint32_t* h, uint32_t d, bool valid, int lane)
{
    const uint32_t prev = __shfl_up(d, 1, 64);
    const bool head = valid && (lane == 0 || prev != d);
    const uint64_t heads = __ballot(head);
    const uint64_t vmask = __ballot(valid);
    if (head)
    {
        const uint64_t above = heads & ~((2ull << lane) - 1ull);
        const int end = above ? __ffsll((unsigned long long)above) - 1 : __popcll(vmask);
        atomicAdd(&h[d], (uint32_t)(end - lane));
    }
}

// Number of keys to sort: n, or the smaller count a previous kernel left at
// n_dev (a 64-bit count as two 32-bit words) when n_dev is given.
__device__ inline int64_t sort_count(int64_t n, const uint32_t* __restrict__ n_dev)
{
    if (!n_dev) return n;
    const int64_t d = (int64_t)(((uint64_t)n_dev[1] << 32) | n_dev[0]);
    return d < n ? d : n;
}

// ---- onesweep LSD radix sort ----------------------------------------------
//
// One launch counts the digits of every pass (one read of the keys); then
// each pass is one launch in which a workgroup takes tiles of
// 256 x sort_items(n) keys, ranks the tile's keys by digit (stable),
// publishes the tile's digit counts, finds the counts of the tiles before it
// by a two-level decoupled look-back and scatters the keys.  Tiles form groups of SORT_LB_GROUP:
// a tile sums the counts of the tiles before it in its group (all published
// at about the same time, one round of loads), and the last tile of a group
// publishes the group's total (FLAG_AGG), then its running total over all
// groups (FLAG_PREFIX) once it has looked back over the groups before it;
// every tile looks back over the groups the same way, newest first, until a
// group with a running total.  A chained look-back over single tiles would
// make the last of a few hundred tiles wait some tens of dependent rounds of
// cross-XCD loads; over groups it is a few.
//
// The pass grid is no larger than the number of its workgroups the device
// holds at once, and workgroup b takes tiles b, b + grid, b + 2 grid, ... in
// that order.  A tile waits only on tiles with smaller ids (its group's
// earlier tiles, and the last tiles of earlier groups), so the unfinished
// tile with the smallest id never waits and runs to its end: every wait
// ends.  (No tile counter is needed for that, and a counter every workgroup
// adds to would be a serial chain of device-scope atomics on one address.)
// Each wait is also bounded (LOOKBACK_SPIN_LIMIT polls, then the error word
// of the scratch is set and the wait gives up) so that a broken invariant
// cannot stall the device.

constexpr uint32_t FLAG_AGG = 1u, FLAG_PREFIX = 2u;
constexpr uint32_t LOOKBACK_SPIN_LIMIT = 1u << 22;
constexpr int LOOKBACK_WINDOW = 4;
constexpr int SORT_HIST_GRID = 2048;
constexpr int SORT_RANK_CHAINS = 1;

struct SortScratch
{
    uint32_t* err;      // [1]
    uint32_t* hist;     // [passes][SORT_HIST_COPIES][RADIX]
    uint32_t* status;   // [passes][nb][RADIX] a tile's digit count | TILE_PUBLISHED
    uint64_t* gstatus;  // [passes][ngroups][RADIX] flag << 32 | count
    int64_t words;      // words in use (cleared before the sort)
    static SortScratch carve(uint32_t* p, int64_t n, int passes)
    {
        SortScratch s;
        const int64_t nb = sort_blocks(n > 0 ? n : 1);
        const int64_t hist_words = (int64_t)passes * SORT_HIST_COPIES * RADIX;
        s.err = p;
        s.hist = p + SORT_HEADER_WORDS;
        s.gstatus = reinterpret_cast<uint64_t*>(p + SORT_HEADER_WORDS + hist_words);
        s.status = reinterpret_cast<uint32_t*>(s.gstatus + (int64_t)passes * sort_lb_groups(n) * RADIX);
        s.words = SORT_HEADER_WORDS + hist_words + (int64_t)passes * (nb + 2 * sort_lb_groups(n)) * RADIX;
        return s;
    }
};

// Digit counts of every pass in one read of the keys: pass p covers bits
// [lo_bit + p * step, lo_bit + min((p + 1) * step, nbits)).  Workgroup b
// counts tiles b, b + grid, ... in LDS and adds them into histogram copy
// b % SORT_HIST_COPIES.
__global__ __launch_bounds__(SORT_THREADS) void onesweep_hist_kernel(const uint64_t* __restrict__ keys, int64_t n_cap,
                                                                     const uint32_t* __restrict__ n_dev, int lo_bit,
                                                                     int step, int nbits, int passes,
                                                                     uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[SORT_MAX_PASSES][RADIX];
    const int64_t n = sort_count(n_cap, n_dev);
    const int t = threadIdx.x;
    for (int p = 0; p < passes; ++p) h[p][t] = 0u;
    __syncthreads();
    for (int64_t b0 = (int64_t)blockIdx.x * SORT_TILE; b0 < n; b0 += (int64_t)gridDim.x * SORT_TILE)
    {
        // every load of the tile in flight at once, then the counting
        uint64_t k[SORT_ITEMS];
#pragma unroll
        for (int r = 0; r < SORT_ITEMS; ++r)
        {
            const int64_t i = b0 + r * SORT_THREADS + t;
            k[r] = i < n ? keys[i] : 0ull;
        }
        for (int p = 0; p < passes; ++p)
        {
            const int sh = lo_bit + p * step;
            const uint32_t mask = (1u << min(step, nbits - p * step)) - 1u;
#pragma unroll
            for (int r = 0; r < SORT_ITEMS; ++r)
                add_digit_runs(h[p], (uint32_t)(k[r] >> sh) & mask, b0 + r * SORT_THREADS + t < n, t & 63);
        }
    }
    __syncthreads();
    uint32_t* out = hist + (blockIdx.x % SORT_HIST_COPIES) * RADIX;
    for (int p = 0; p < passes; ++p)
        if (h[p][t]) atomicAdd(&out[p * SORT_HIST_COPIES * RADIX + t], h[p][t]);
}

// A status word carries its flag and its count together, and nothing else is
// read on the strength of it, so relaxed device-scope atomics suffice: they
// are coherent across the XCDs' L2s without the L2 write-back / invalidate
// that release / acquire ordering would add to every poll.
__device__ inline uint64_t status_load(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void status_store(uint64_t* p, uint32_t flag, uint32_t value)
{
    __hip_atomic_store(p, ((uint64_t)flag << 32) | value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A tile's status word: its digit count (at most 256 x ITEMS) with the top bit
// set once published (zero before).
constexpr uint32_t TILE_PUBLISHED = 0x80000000u;

__device__ inline uint32_t tile_load(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void tile_store(uint32_t* p, uint32_t count)
{
    __hip_atomic_store(p, count | TILE_PUBLISHED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of digit t's counts of tiles [first, first + r) (r < SORT_LB_GROUP),
// every load in flight at once; words not published yet are polled again.
__device__ inline uint32_t group_sum(const uint32_t* __restrict__ status, int64_t first, int r, int t,
                                     uint32_t* __restrict__ err)
{
    uint32_t sum = 0;
    uint32_t pending = (1u << r) - 1u;
    uint32_t spins = 0;
    while (pending)
    {
        uint32_t sw[SORT_LB_GROUP - 1];
#pragma unroll
        for (int q = 0; q < SORT_LB_GROUP - 1; ++q)
            sw[q] = (pending >> q) & 1u ? tile_load(status + (first + q) * RADIX + t) : 0u;
#pragma unroll
        for (int q = 0; q < SORT_LB_GROUP - 1; ++q)
            if (((pending >> q) & 1u) && (sw[q] & TILE_PUBLISHED))
            {
                sum += sw[q] & ~TILE_PUBLISHED;
                pending &= ~(1u << q);
            }
        if (pending)
        {
            if (++spins > LOOKBACK_SPIN_LIMIT)
            {
                atomicOr(err, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return sum;
}

// Exclusive count of digit t over the entries before `tile` (tile > 0) of a
// status array whose entry 0 publishes FLAG_PREFIX: newest first,
// LOOKBACK_WINDOW predecessors per round (their loads in flight together); a
// round consumes words until a prefix (done), an entry that has not
// published yet (poll again from it) or the end of the window.
__device__ inline uint32_t look_back(const uint64_t* __restrict__ status, int64_t tile, int t,
                                     uint32_t* __restrict__ err)
{
    uint32_t excl = 0;
    int64_t j = tile - 1;
    uint32_t spins = 0;
    bool done = false;
    while (!done)
    {
        uint64_t sw[LOOKBACK_WINDOW];
#pragma unroll
        for (int q = 0; q < LOOKBACK_WINDOW; ++q)
            sw[q] = j - q >= 0 ? status_load(status + (j - q) * RADIX + t) : 0ull;
        int used = 0;
        bool stalled = false;
#pragma unroll
        for (int q = 0; q < LOOKBACK_WINDOW; ++q)
        {
            if (done || stalled || j - q < 0) continue;
            const uint32_t flag = (uint32_t)(sw[q] >> 32);
            if (flag == 0u)
            {
                stalled = true;
                continue;
            }
            excl += (uint32_t)sw[q];
            ++used;
            done = flag == FLAG_PREFIX || j - q == 0;
        }
        j -= used;
        if (!done && stalled)
        {
            if (++spins > LOOKBACK_SPIN_LIMIT)
            {
                atomicOr(err, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return excl;
}

// One pass: for each of the workgroup's tiles, rank the tile's keys by digit
// (stable; wave w ranks its 64 x ITEMS keys in order, then the waves in
// order), find the digit's global offset (the pass's exclusive digit prefix +
// the counts of earlier tiles by look-back), reorder the tile's keys by digit
// in LDS and write each digit's run with consecutive threads on consecutive
// addresses.
template <bool HAS_VALS, int DBITS, int ITEMS>
__global__ __launch_bounds__(SORT_THREADS) void onesweep_pass_kernel(
    const uint64_t* __restrict__ keys_in, uint64_t* __restrict__ keys_out, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ vals_out, int64_t n_cap, const uint32_t* __restrict__ n_dev, int shift, int ident,
    const uint32_t* __restrict__ hist, uint32_t* __restrict__ status, uint64_t* __restrict__ gstatus,
    uint32_t* __restrict__ err)
{
    const int64_t n = sort_count(n_cap, n_dev);
    constexpr uint32_t dmask = (1u << DBITS) - 1u;
    constexpr int TILE = SORT_THREADS * ITEMS;
    constexpr int NW = SORT_THREADS / 64;
    constexpr int NC = SORT_RANK_CHAINS;
    // per wave, SORT_RANK_CHAINS running digit counts: the wave ranks
    // consecutive slices of its keys as independent chains of LDS
    // read-modify-writes (interleaved, so that their latencies overlap)
    __shared__ uint32_t wcnt[NW][NC][RADIX];
    __shared__ uint32_t gdelta[RADIX];  // global start of this tile's run of digit d minus lbase[d]
    __shared__ uint32_t lbase[RADIX];   // start of digit d in the tile's reordered keys
    __shared__ uint32_t wsum[NW];
    __shared__ uint8_t ldig[TILE];  // digit of each reordered key
    __shared__ uint64_t lkeys[TILE];  // reordered keys, then (as 32-bit words) values
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    // the pass's exclusive digit starts (digit = thread), once per workgroup:
    // the digit's count over the histogram copies, scanned across the digits
    uint32_t gstart = 0;
    {
        uint32_t tot = 0;
#pragma unroll
        for (int c = 0; c < SORT_HIST_COPIES; ++c) tot += hist[c * RADIX + t];
        uint32_t sc = tot;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const uint32_t y = __shfl_up(sc, o, 64);
            if (lane >= o) sc += y;
        }
        if (lane == 63) wsum[w] = sc;
        __syncthreads();
        uint32_t pre = 0;
        for (int q = 0; q < w; ++q) pre += wsum[q];
        gstart = pre + sc - tot;
    }
    const int64_t ntiles = (n + TILE - 1) / TILE;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
    {
        const int64_t blk0 = tile * TILE;
        __syncthreads();  // the previous tile's LDS reads are done
#pragma unroll
        for (int i = 0; i < NW; ++i)
#pragma unroll
            for (int h = 0; h < NC; ++h) wcnt[i][h][t] = 0;
        const int64_t base = blk0 + (int64_t)w * (ITEMS * 64);
        uint64_t k[ITEMS];
        uint32_t v[ITEMS];
        uint32_t rk[ITEMS];
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
        {
            const int64_t idx = base + r * 64 + lane;
            const bool valid = idx < n;
            k[r] = valid ? keys_in[idx] : 0ull;
            if (HAS_VALS) v[r] = valid ? (ident ? (uint32_t)idx : vals_in[idx]) : 0u;
        }
        __syncthreads();
        static_assert(ITEMS % NC == 0, "ranking chains of equal length");
        constexpr int SLICE = ITEMS / NC;
#pragma unroll
        for (int r2 = 0; r2 < ITEMS; ++r2)
        {
            const int h = r2 % NC;                // chain
            const int r = (r2 / NC) + h * SLICE;  // item
            const int64_t idx = base + r * 64 + lane;
            const bool valid = idx < n;
            const uint64_t active = __ballot(valid);
            const uint32_t d = (uint32_t)(k[r] >> shift) & dmask;
            const uint64_t m = match_digit<DBITS>(d, active);
            const uint32_t before = (uint32_t)__popcll(m & lanemask_lt(lane));
            uint32_t cnt = 0;
            if (valid) cnt = wcnt[w][h][d];
            rk[r] = cnt + before;
            if (valid && lane == 63 - __clzll(m)) wcnt[w][h][d] = cnt + (uint32_t)__popcll(m);
        }
        // a chain's keys follow all keys of the chains before it
#pragma unroll
        for (int r = SLICE; r < ITEMS; ++r)
        {
            const uint32_t d = (uint32_t)(k[r] >> shift) & dmask;
#pragma unroll
            for (int h = 0; h < r / SLICE; ++h) rk[r] += wcnt[w][h][d];
        }
        __syncthreads();
        // per digit (thread t = digit): exclusive offsets of the waves, the
        // tile's count and its start in the tile's reordered keys
        uint32_t dtot = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i)
        {
            uint32_t c = 0;
#pragma unroll
            for (int h = 0; h < NC; ++h) c += wcnt[i][h][t];
            wcnt[i][0][t] = dtot;
            dtot += c;
        }
        {
            uint32_t l = dtot;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1)
            {
                const uint32_t z = __shfl_up(l, o, 64);
                if (lane >= o) l += z;
            }
            if (lane == 63) wsum[w] = l;
            __syncthreads();
            uint32_t pre_l = 0;
            for (int q = 0; q < w; ++q) pre_l += wsum[q];
            lbase[t] = pre_l + l - dtot;
        }
        __syncthreads();
        // the keys in digit order in LDS (so that neither they nor their
        // ranks stay in registers through the look-back)
        uint32_t lp[ITEMS];
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
        {
            const int64_t idx = base + r * 64 + lane;
            const uint32_t d = (uint32_t)(k[r] >> shift) & dmask;
            lp[r] = lbase[d] + wcnt[w][0][d] + rk[r];
            if (idx < n)
            {
                lkeys[lp[r]] = k[r];
                ldig[lp[r]] = (uint8_t)d;
            }
        }
        // digits above dmask never occur: their counts stay zero and take no
        // part in the look-back (a 6-bit pass polls a quarter of the words)
        const bool live = (uint32_t)t <= dmask;
        uint32_t excl = 0;
        if (live)
        {
            tile_store(status + tile * RADIX + t, dtot);
            const int64_t g = tile / SORT_LB_GROUP;
            const int r = (int)(tile - g * SORT_LB_GROUP);
            const uint32_t in_group = r > 0 ? group_sum(status, tile - r, r, t, err) : 0u;
            uint32_t before = 0;
            if (r == SORT_LB_GROUP - 1)
            {
                const uint32_t total = in_group + dtot;
                status_store(gstatus + g * RADIX + t, g == 0 ? FLAG_PREFIX : FLAG_AGG, total);
                if (g > 0)
                {
                    before = look_back(gstatus, g, t, err);
                    status_store(gstatus + g * RADIX + t, FLAG_PREFIX, before + total);
                }
            }
            else if (g > 0)
                before = look_back(gstatus, g, t, err);
            excl = before + in_group;
        }
        // destination of the tile's run of digit t, as an offset from its
        // reordered position
        gdelta[t] = gstart + excl - lbase[t];
        __syncthreads();
        const int nk = (int)min((int64_t)TILE, n - blk0);
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
        {
            const int i = t + r * SORT_THREADS;
            if (i < nk) keys_out[gdelta[ldig[i]] + (uint32_t)i] = lkeys[i];
        }
        if (HAS_VALS)
        {
            uint32_t* lv = reinterpret_cast<uint32_t*>(lkeys);
            __syncthreads();
#pragma unroll
            for (int r = 0; r < ITEMS; ++r)
            {
                const int64_t idx = base + r * 64 + lane;
                if (idx < n) lv[lp[r]] = v[r];
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < ITEMS; ++r)
            {
                const int i = t + r * SORT_THREADS;
                if (i < nk) vals_out[gdelta[ldig[i]] + (uint32_t)i] = lv[i];
            }
        }
    }
}

// Workgroups of `kernel` (SORT_THREADS threads each) the current device
// holds at once, queried once per (device, kernel).
static int64_t resident_blocks(const void* kernel)
{
    static std::mutex mu;
    static std::vector<std::pair<std::pair<int, const void*>, int64_t>> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lock(mu);
    for (const auto& e : cache)
        if (e.first.first == dev && e.first.second == kernel) return e.second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, SORT_THREADS, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    cache.push_back({{dev, kernel}, (int64_t)per_cu * cus});
    return (int64_t)per_cu * cus;
}

template <bool HAS_VALS, int ITEMS>
static const void* pass_kernel_items(int bits)
{
    switch (bits)
    {
        case 1: return (const void*)onesweep_pass_kernel<HAS_VALS, 1, ITEMS>;
        case 2: return (const void*)onesweep_pass_kernel<HAS_VALS, 2, ITEMS>;
        case 3: return (const void*)onesweep_pass_kernel<HAS_VALS, 3, ITEMS>;
        case 4: return (const void*)onesweep_pass_kernel<HAS_VALS, 4, ITEMS>;
        case 5: return (const void*)onesweep_pass_kernel<HAS_VALS, 5, ITEMS>;
        case 6: return (const void*)onesweep_pass_kernel<HAS_VALS, 6, ITEMS>;
        case 7: return (const void*)onesweep_pass_kernel<HAS_VALS, 7, ITEMS>;
        default: return (const void*)onesweep_pass_kernel<HAS_VALS, 8, ITEMS>;
    }
}

template <bool HAS_VALS>
static const void* pass_kernel(int bits, int items)
{
    return items == SORT_ITEMS_SMALL ? pass_kernel_items<HAS_VALS, SORT_ITEMS_SMALL>(bits)
                                     : pass_kernel_items<HAS_VALS, SORT_ITEMS>(bits);
}

// Stable LSD radix sort over bits [lo_bit, lo_bit + nbits) in digits of at
// most 8 bits.  The sorted data ends in the buffers returned through
// out_keys / out_vals.
int radix_sort_ex(uint64_t* keys, uint64_t* keys_alt, uint32_t* vals, uint32_t* vals_alt, int64_t n, int nbits,
                  uint32_t* counts, hipStream_t stream, uint64_t** out_keys, uint32_t** out_vals,
                  const uint32_t* n_dev, int lo_bit, bool identity_vals, void* also_zero, int64_t also_zero_n)
{
    *out_keys = keys;
    if (out_vals) *out_vals = vals;
    if (n <= 0 || nbits <= 0) return also_zero ? fill_u32(also_zero, 0u, also_zero_n, stream) : SVR_OK;
    SVR_REQUIRE(lo_bit >= 0 && lo_bit + nbits <= 64, SVR_ERR_INVALID, "radix sort bits [%d, %d) out of range", lo_bit,
                lo_bit + nbits);
    const int64_t nb = sort_blocks(n);
    SVR_REQUIRE(n <= (int64_t)UINT32_MAX, SVR_ERR_INVALID, "radix sort of %lld keys too large", (long long)n);
    const int passes = radix_passes(nbits);
    SVR_REQUIRE(passes <= SORT_MAX_PASSES, SVR_ERR_INVALID, "radix sort of %d bits needs too many passes", nbits);
    const int step = (nbits + passes - 1) / passes;
    SortScratch sc = SortScratch::carve(counts, n, passes);
    if (fill_u32_pair(counts, sc.words, also_zero, also_zero ? also_zero_n : 0, 0u, stream) != SVR_OK)
        return SVR_ERR_HIP;
    onesweep_hist_kernel<<<(unsigned)std::min<int64_t>(nb, SORT_HIST_GRID), SORT_THREADS, 0, stream>>>(
        keys, n, n_dev, lo_bit, step, nbits, passes, sc.hist);
    SVR_CHECK_LAUNCH();
    uint64_t* kin = keys; uint64_t* kout = keys_alt;
    uint32_t* vin = vals; uint32_t* vout = vals_alt;
    for (int p = 0; p < passes; ++p)
    {
        const int done = p * step;
        const int shift = lo_bit + done;
        const int bits = (nbits - done) < step ? (nbits - done) : step;
        uint32_t* status = sc.status + (int64_t)p * nb * RADIX;
        uint64_t* gstatus = sc.gstatus + (int64_t)p * sort_lb_groups(n) * RADIX;
        const uint32_t* hist = sc.hist + (int64_t)p * SORT_HIST_COPIES * RADIX;
        const void* kernel = vals ? pass_kernel<true>(bits, sort_items(n)) : pass_kernel<false>(bits, sort_items(n));
        const int64_t resident = resident_blocks(kernel);
        SVR_REQUIRE(resident > 0, SVR_ERR_HIP, "radix sort occupancy query failed");
        const unsigned grid = (unsigned)std::min(nb, resident);
        const uint64_t* kin_c = kin;
        const uint32_t* vin_c = vin;
        int ident = (identity_vals && p == 0) ? 1 : 0;
        void* args[] = {(void*)&kin_c, (void*)&kout,  (void*)&vin_c, (void*)&vout,   (void*)&n,
                        (void*)&n_dev, (void*)&shift, (void*)&ident, (void*)&hist,   (void*)&status,
                        (void*)&gstatus, (void*)&sc.err};
        SVR_CHECK_HIP(hipLaunchKernel(kernel, dim3(grid), dim3(SORT_THREADS), args, 0, stream));
        uint64_t* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
    }
    *out_keys = kin;
    if (out_vals) *out_vals = vin;
    return SVR_OK;
}

// ---- order table -----------------------------------------------------

__global__ void or_reduce_kernel(const int64_t* __restrict__ paths, int64_t n, uint32_t* __restrict__ out2)
{
    uint64_t acc = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) acc |= (uint64_t)paths[i];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
    {
        const uint32_t lo = __shfl_xor((uint32_t)acc, o, 64);
        const uint32_t hi = __shfl_xor((uint32_t)(acc >> 32), o, 64);
        acc |= ((uint64_t)hi << 32) | lo;
    }
    if ((threadIdx.x & 63) == 0)
    {
        atomicOr(out2 + 0, (uint32_t)acc);
        atomicOr(out2 + 1, (uint32_t)(acc >> 32));
    }
}

__global__ void order_keys_kernel(const int64_t* __restrict__ paths, int64_t n8, int lshift, int with_q,
                                  uint64_t* __restrict__ keys)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    const uint32_t q = (uint32_t)(i & 7);
    const uint64_t rank = (uint64_t)paths[i >> 3] ^ order_table(q);
    keys[i] = with_q ? (((rank >> lshift) << 3) | q) : rank;
}

__global__ void order_scatter_kernel(const uint32_t* __restrict__ sorted_pairs, int64_t n8, uint32_t* __restrict__ pos,
                                     uint32_t* __restrict__ pair)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n8) return;
    const uint32_t pr = sorted_pairs[p];
    pos[pr] = (uint32_t)p;
    pair[p] = pr;
}

struct OrderScratch
{
    uint64_t *keys, *keys_alt;
    uint32_t *vals, *vals_alt, *counts, *misc;
    static OrderScratch carve(void* p, int64_t N, size_t* bytes)
    {
        Carver c(p);
        OrderScratch s;
        const int64_t n8 = 8 * N;
        s.keys = c.take<uint64_t>(n8);
        s.keys_alt = c.take<uint64_t>(n8);
        s.vals = c.take<uint32_t>(n8);
        s.vals_alt = c.take<uint32_t>(n8);
        s.counts = c.take<uint32_t>(sort_scratch_words(n8, SORT_MAX_PASSES));
        s.misc = c.take<uint32_t>(4);
        if (bytes) *bytes = c.size();
        return s;
    }
};

// ---- per-frame duplication, ranges and key export --------------------

// Each active (item, quadrant) pair's tile count at its rank position, in
// view blockIdx.y's block of pos_off (pos_off cleared beforehand).
// With big_list (one view), items with more than BIG_DUP entries are also
// queued there (counted in counters[2], huge ones from the end in
// counters[3]) for the duplication's workgroup-per-item part.
__global__ __launch_bounds__(256) void pair_count_kernel(int64_t N, const uint32_t* __restrict__ n_dup,
                                                         const uint32_t* __restrict__ qbits,
                                                         const uint32_t* __restrict__ order_pos,
                                                         uint32_t* __restrict__ pos_off,
                                                         uint32_t* __restrict__ big_list,
                                                         uint32_t* __restrict__ big_counts)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const size_t p = (size_t)blockIdx.y * (size_t)N + (size_t)v;
    const uint32_t nd = n_dup[p];
    if (nd == 0) return;
    if (big_list && nd > BIG_DUP) queue_large(big_list, big_counts, N, nd, (uint32_t)v);
    const uint32_t bits = qbits[p];
    const uint32_t tiles = nd / (uint32_t)__popc(bits);
    uint32_t* blk = pos_off + 8 * (size_t)blockIdx.y * (size_t)N;
    for (uint32_t b = bits; b; b &= b - 1u)
    {
        const uint32_t q = (uint32_t)__builtin_ctz(b);
        blk[order_pos[8 * v + q]] = tiles;
    }
}

int pair_offsets_launch(int V, int64_t N, const GeomState& g, const uint32_t* order_pos, hipStream_t stream,
                        bool queue_large_items)
{
    const int64_t n8 = 8 * (int64_t)V * N;
    if (n8 == 0) return fill_u32(g.counters + 2, 0u, 2, stream);
    if (fill_u32_pair(g.pos_off, n8, g.counters + 2, 2, 0u, stream) != SVR_OK) return SVR_ERR_HIP;
    pair_count_kernel<<<dim3((unsigned)((N + 255) / 256), (unsigned)V), 256, 0, stream>>>(
        N, g.n_dup, g.qbits, order_pos, g.pos_off, queue_large_items ? g.big_list : nullptr, g.counters + 2);
    SVR_CHECK_LAUNCH();
    return pair_scan_launch(g.pos_off, n8, g.pos_sums, stream);
}

// forward.cu:523-577, emitting (tile << pos_bits | pos) keys: each quadrant's
// run at its rank-ordered offset (recorded in qoff for the later passes).
// Item v (items with more than BIG_DUP entries only get their run offsets
// here; pair_count_kernel queued them for duplicate_big).
__device__ inline void duplicate_small(const GeomState& g, int64_t N, int gx, int gy, int pos_bits, int64_t v,
                                       const uint32_t* __restrict__ order_pos, uint64_t* __restrict__ keys)
{
    if (v >= N) return;
    const uint32_t nd = g.n_dup[v];
    if (nd == 0) return;
    const bool large = nd > BIG_DUP;
    int txmin, tymin, txmax, tymax;
    rect_of(g.bbox[v], gx, gy, txmin, tymin, txmax, tymax);
    const uint32_t bits = g.qbits[v];
    for (uint32_t b = bits; b; b &= b - 1u)
    {
        const uint32_t q = (uint32_t)__builtin_ctz(b);
        const uint64_t pos = order_pos[8 * v + q];
        uint32_t off = g.pos_off[pos];
        set_run_offset(g, (size_t)v, bits, q, off);
        if (large) continue;
        for (int y = tymin; y <= tymax; ++y)
            for (int x = txmin; x <= txmax; ++x)
            {
                keys[off] = ((uint64_t)((uint32_t)y * (uint32_t)gx + (uint32_t)x) << pos_bits) | pos;
                ++off;
            }
    }
}

// Large voxels: one workgroup each; then the huge ones with every workgroup
// taking a strided share of each.  Entry e of a voxel: quadrant (the qi-th set
// bit of its bits) major, then the tile rows of its rectangle, written into
// that quadrant's run (its offset read from pos_off, not from the run offsets
// duplicate_small records in the same launch).  Workgroup bid of nblk.
__device__ inline void duplicate_big(const GeomState& g, int64_t N, int gx, int gy, int pos_bits, uint32_t bid,
                                     uint32_t nblk, const uint32_t* __restrict__ order_pos,
                                     uint64_t* __restrict__ keys)
{
    // pass 0: one workgroup per large voxel; pass 1: each huge voxel split into
    // K entry ranges taken by K different workgroups (item = chunk * huge + voxel)
    const uint32_t count = g.counters[2], huge = g.counters[3];
    const uint32_t K = huge ? max(1u, nblk / huge) : 1u;
    for (uint32_t pass = 0; pass < 2; ++pass)
    {
        for (uint32_t it = bid; it < (pass == 0 ? count : huge * K); it += nblk)
        {
            const uint32_t i = pass == 0 ? it : it % huge;
            const uint32_t v = pass == 0 ? g.big_list[i] : g.big_list[N - 1 - i];
            const uint32_t nd = g.n_dup[v];
            const uint32_t c = pass == 0 ? 0u : it / huge;
            const uint32_t e_lo = pass == 0 ? 0u : (uint32_t)(((uint64_t)nd * c) / K);
            const uint32_t e_hi = pass == 0 ? nd : (uint32_t)(((uint64_t)nd * (c + 1)) / K);
            int txmin, tymin, txmax, tymax;
            rect_of(g.bbox[v], gx, gy, txmin, tymin, txmax, tymax);
            const uint32_t w = (uint32_t)(txmax - txmin + 1);
            const uint32_t tiles = w * (uint32_t)(tymax - tymin + 1);
            const uint32_t bits = g.qbits[v];
            uint32_t qi = 0;
            for (uint32_t b = bits; b; b &= b - 1u, ++qi)
            {
                // the part of [e_lo, e_hi) in this quadrant's run
                const uint32_t q = (uint32_t)__builtin_ctz(b);
                const uint32_t r_lo = max(e_lo, qi * tiles), r_hi = min(e_hi, (qi + 1) * tiles);
                if (r_lo >= r_hi) continue;
                const uint64_t pos = order_pos[8 * (size_t)v + q];
                const uint32_t off = g.pos_off[pos];
                for (uint32_t e = r_lo + threadIdx.x; e < r_hi; e += blockDim.x)
                {
                    const uint32_t rem = e - qi * tiles;
                    const uint32_t y = (uint32_t)tymin + rem / w, x = (uint32_t)txmin + rem % w;
                    keys[off + rem] = ((uint64_t)(y * (uint32_t)gx + x) << pos_bits) | pos;
                }
            }
        }
    }
}

// forward.cu:582-604 plus the value decode: vox_list[i] = (q << 29) | vox_id.
__global__ __launch_bounds__(256) void ranges_kernel(const uint64_t* __restrict__ keys, int64_t R_cap,
                                                     const uint32_t* __restrict__ R_dev, int pos_bits,
                                                     const uint32_t* __restrict__ order_pair,
                                                     uint2* __restrict__ ranges, uint32_t* __restrict__ vox_list)
{
    const int64_t R = sort_count(R_cap, R_dev);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    const uint64_t k = keys[i];
    const uint32_t tile = (uint32_t)(k >> pos_bits);
    const uint32_t pair = order_pair[k & ((1ull << pos_bits) - 1ull)];
    vox_list[i] = ((pair & 7u) << 29) | (pair >> 3);
    uint32_t* rg = (uint32_t*)ranges;
    if (i == 0)
        rg[2 * tile] = 0;
    else
    {
        const uint32_t prev = (uint32_t)(keys[i - 1] >> pos_bits);
        if (prev != tile)
        {
            rg[2 * prev + 1] = (uint32_t)i;
            rg[2 * tile] = (uint32_t)i;
        }
    }
    if (i == R - 1) rg[2 * tile + 1] = (uint32_t)R;
}

__global__ void export_keys_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vox_list,
                                   int64_t R, int pos_bits, const int64_t* __restrict__ paths,
                                   uint64_t* __restrict__ hi, uint64_t* __restrict__ lo, uint32_t* __restrict__ vals)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    const uint64_t tile = keys[i] >> pos_bits;
    const uint32_t val = vox_list[i];
    const uint64_t rank = (uint64_t)paths[val & VOX_ID_MASK] ^ order_table(val >> 29);
    hi[i] = tile >> 1;
    lo[i] = (tile << 63) | rank;
    vals[i] = val;
}

// The duplication in one launch: the first DUP_BIG_BLOCKS workgroups take the
// items pair_count_kernel queued as large (they run longest, so they start
// first), the rest one item per thread.
constexpr uint32_t DUP_BIG_BLOCKS = 1024;
__global__ __launch_bounds__(256) void duplicate_kernel(GeomState g, int64_t N, int gx, int gy, int pos_bits,
                                                        const uint32_t* __restrict__ order_pos,
                                                        uint64_t* __restrict__ keys)
{
    if (blockIdx.x < DUP_BIG_BLOCKS)
        duplicate_big(g, N, gx, gy, pos_bits, blockIdx.x, DUP_BIG_BLOCKS, order_pos, keys);
    else
        duplicate_small(g, N, gx, gy, pos_bits, (int64_t)(blockIdx.x - DUP_BIG_BLOCKS) * blockDim.x + threadIdx.x,
                        order_pos, keys);
}

int ranges_launch(const uint64_t* keys, int64_t R, int pos_bits, const uint32_t* order_pair, uint2* ranges,
                  uint32_t* vox_list, hipStream_t stream, const uint32_t* R_dev)
{
    ranges_kernel<<<(unsigned)((R + 255) / 256), 256, 0, stream>>>(keys, R, R_dev, pos_bits, order_pair, ranges,
                                                                   vox_list);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

__global__ void clamp_count_kernel(uint32_t* __restrict__ counters, uint64_t cap, uint32_t* __restrict__ overflow)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint64_t R = ((uint64_t)counters[1] << 32) | counters[0];
    const uint64_t eff = R <= cap ? R : 0ull;
    counters[4] = (uint32_t)eff;
    counters[5] = (uint32_t)(eff >> 32);
    if (R > cap) overflow[0] = 1u;  // sticky
    overflow[1] = R > cap ? 1u : 0u;  // this frame
}

int clamp_count_launch(uint32_t* counters, uint64_t cap, uint32_t* overflow, hipStream_t stream)
{
    clamp_count_kernel<<<1, 64, 0, stream>>>(counters, cap, overflow);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

// Tile queue keys: 255 - length_bucket (an ascending sort puts the longest
// lists first), values the tile ids.
__global__ __launch_bounds__(256) void tile_queue_keys_kernel(int64_t T, const uint2* __restrict__ ranges,
                                                              uint64_t* __restrict__ qkeys)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= T) return;
    const uint2 r = ranges[g];
    qkeys[g] = 255u - length_bucket(r.y - r.x);
}

// The whole queue in one workgroup (frames of at most QUEUE_ONE_BLOCK_MAX
// tiles): bucket histogram in LDS, bucket offsets longest first, then every
// tile takes the next slot of its bucket.  The order inside a bucket follows
// the LDS atomics (it only affects scheduling, never a result).
constexpr int64_t QUEUE_ONE_BLOCK_MAX = 32768;
__global__ __launch_bounds__(1024) void tile_queue_block_kernel(int64_t T, const uint2* __restrict__ ranges,
                                                                uint32_t* __restrict__ order,
                                                                uint32_t* __restrict__ clear8)
{
    __shared__ uint32_t hist[256];
    __shared__ uint32_t base[256];
    const int t = threadIdx.x;
    if (t < 256) hist[t] = 0u;
    if (clear8 && t < 8) clear8[t] = 0u;
    __syncthreads();
    for (int64_t g = t; g < T; g += blockDim.x)
    {
        const uint2 r = ranges[g];
        atomicAdd(&hist[length_bucket(r.y - r.x)], 1u);
    }
    __syncthreads();
    if (t < 64)
    {
        // exclusive offsets over buckets 255, 254, ..., 0 (lane t owns buckets 255 - 4t - k)
        uint32_t c[4], s = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            c[k] = hist[255 - 4 * t - k];
            s += c[k];
        }
        uint32_t incl = s;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const uint32_t y = __shfl_up(incl, o, 64);
            if (t >= o) incl += y;
        }
        uint32_t run = incl - s;
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            base[255 - 4 * t - k] = run;
            run += c[k];
        }
    }
    __syncthreads();
    for (int64_t g = t; g < T; g += blockDim.x)
    {
        const uint2 r = ranges[g];
        const uint32_t slot = atomicAdd(&base[length_bucket(r.y - r.x)], 1u);
        order[slot] = (uint32_t)g;
    }
}

int tile_queue_launch(int64_t T, const uint2* ranges, uint64_t* qkeys, uint64_t* qkeys_alt, uint32_t* order,
                      uint32_t* order_alt, uint32_t* qcounts, hipStream_t stream, uint32_t* clear8)
{
    if (T <= QUEUE_ONE_BLOCK_MAX)
    {
        tile_queue_block_kernel<<<1, 1024, 0, stream>>>(T, ranges, order, clear8);
        SVR_CHECK_LAUNCH();
        return SVR_OK;
    }
    if (clear8 && fill_u32(clear8, 0u, 8, stream) != SVR_OK) return SVR_ERR_HIP;
    tile_queue_keys_kernel<<<(unsigned)((T + 255) / 256), 256, 0, stream>>>(T, ranges, qkeys_alt);
    SVR_CHECK_LAUNCH();
    uint64_t* qk = nullptr;
    uint32_t* qo = nullptr;
    const int rc = radix_sort_ex(qkeys_alt, qkeys, order_alt, order, T, 8, qcounts, stream, &qk, &qo, nullptr, 0, true);
    if (rc != SVR_OK) return rc;
    SVR_REQUIRE(qo == order, SVR_ERR_HIP, "tile queue sort ended in the wrong buffer");
    return SVR_OK;
}

int pair_prepare(int64_t N, const GeomState& g, const OrderTable& ot, hipStream_t stream)
{
    return pair_offsets_launch(1, N, g, ot.pos, stream, true);
}

int binning_forward(int W, int H, int64_t N, const GeomState& g, const OrderTable& ot, int64_t R,
                    BinningState& b, ImageState& img, hipStream_t stream, bool pairs_ready)
{
    const int gx = tiles_x(W), gy = tiles_y(H);
    const int64_t T = (int64_t)gx * gy;
    prof_begin(PROF_BINNING, stream);
    if (R == 0)
    {
        SVR_REQUIRE(fill_u32(img.ranges, 0u, 2 * T, stream) == SVR_OK, SVR_ERR_HIP, "fill ranges failed");
        const int rc0 = tile_queue_launch(T, img.ranges, img.qkeys, img.qkeys_alt, img.order, img.order_alt,
                                          img.qcounts, stream, img.ctr);
        prof_end(PROF_BINNING, stream);
        return rc0;
    }
    const int pos_bits = (int)required_bits((uint64_t)8 * N);
    const int tile_bits = (int)required_bits((uint64_t)T);
    SVR_REQUIRE(radix_passes(tile_bits) <= SORT_FRAME_PASSES, SVR_ERR_INVALID, "%lld tiles are too many to sort",
                (long long)T);
    int rc = pairs_ready ? SVR_OK : pair_prepare(N, g, ot, stream);
    if (rc != SVR_OK) return rc;
    // entries are emitted in rank order, so the sort only orders the tile
    // bits; it ping-pongs between the two buffers, one digit per pass: emit
    // into the buffer that makes the sorted result land in b.keys / b.orig
    const bool odd = radix_passes(tile_bits) & 1;
    uint64_t* k0 = odd ? b.keys_alt : b.keys;
    uint32_t* o0 = odd ? b.orig_alt : b.orig;
    duplicate_kernel<<<(unsigned)(DUP_BIG_BLOCKS + (N + 255) / 256), 256, 0, stream>>>(g, N, gx, gy, pos_bits,
                                                                                      ot.pos, k0);
    SVR_CHECK_LAUNCH();
    uint64_t* sorted = nullptr;
    uint32_t* sorted_orig = nullptr;
    // the tile ranges (empty tiles stay (0, 0)) are cleared with the sort's scratch
    rc = radix_sort_ex(k0, odd ? b.keys : b.keys_alt, o0, odd ? b.orig : b.orig_alt, R, tile_bits, b.counts, stream,
                       &sorted, &sorted_orig, nullptr, pos_bits, true, img.ranges, 2 * T);
    if (rc != SVR_OK) return rc;
    SVR_REQUIRE(sorted == b.keys && sorted_orig == b.orig, SVR_ERR_HIP, "sort ended in the wrong buffer");
    rc = ranges_launch(b.keys, R, pos_bits, ot.pair, img.ranges, b.vox_list, stream);
    if (rc != SVR_OK) return rc;
    rc = tile_queue_launch(T, img.ranges, img.qkeys, img.qkeys_alt, img.order, img.order_alt, img.qcounts, stream,
                           img.ctr);
    if (rc != SVR_OK) return rc;
    prof_end(PROF_BINNING, stream);
    return SVR_OK;
}

}  // namespace svr

using namespace svr;

extern "C" {

size_t svr_order_table_scratch_size(int64_t N)
{
    size_t bytes = 0;
    OrderScratch::carve(nullptr, N < 0 ? 0 : N, &bytes);
    return bytes;
}

int svr_build_order_table(int64_t N, const int64_t* paths, void* order_table_buf, void* scratch,
                          svr_stream_t stream)
{
    SVR_REQUIRE(N >= 0 && N <= (int64_t)VOX_ID_MASK + 1, SVR_ERR_INVALID, "voxel count out of range");
    if (N == 0) return SVR_OK;
    SVR_REQUIRE(paths && order_table_buf && scratch, SVR_ERR_INVALID, "svr_build_order_table: null pointer");
    OrderTable ot = OrderTable::carve(order_table_buf, N, nullptr);
    OrderScratch s = OrderScratch::carve(scratch, N, nullptr);
    const int64_t n8 = 8 * N;
    if (fill_u32(s.misc, 0u, 4, stream) != SVR_OK) return SVR_ERR_HIP;
    or_reduce_kernel<<<256, 256, 0, stream>>>(paths, N, s.misc);
    SVR_CHECK_LAUNCH();
    uint32_t orv[2] = {0, 0};
    SVR_CHECK_HIP(hipMemcpyAsync(orv, s.misc, sizeof(orv), hipMemcpyDeviceToHost, stream));
    SVR_CHECK_HIP(hipStreamSynchronize(stream));
    const uint64_t all = ((uint64_t)orv[1] << 32) | orv[0];
    // finest level any octpath uses: below it every rank bit is q repeated
    const int finest = all ? (MAX_NUM_LEVELS - (__builtin_ctzll(all) / 3)) : 0;
    const int with_q = finest < MAX_NUM_LEVELS;
    const int lshift = 3 * (MAX_NUM_LEVELS - finest);
    const int key_bits = with_q ? 3 * finest + 3 : ORDER_RANK_BITS;
    order_keys_kernel<<<(unsigned)((n8 + 255) / 256), 256, 0, stream>>>(paths, n8, lshift, with_q, s.keys);
    SVR_CHECK_LAUNCH();
    uint64_t* ko = nullptr;
    uint32_t* vo = nullptr;
    int rc = radix_sort_ex(s.keys, s.keys_alt, s.vals, s.vals_alt, n8, key_bits, s.counts, stream, &ko, &vo, nullptr,
                           0, true);
    if (rc != SVR_OK) return rc;
    uint32_t sort_err = 0;
    SVR_CHECK_HIP(hipMemcpyAsync(&sort_err, s.counts, sizeof(sort_err), hipMemcpyDeviceToHost, stream));
    SVR_CHECK_HIP(hipStreamSynchronize(stream));
    SVR_REQUIRE(sort_err == 0u, SVR_ERR_HIP, "order sort look-back timed out");
    order_scatter_kernel<<<(unsigned)((n8 + 255) / 256), 256, 0, stream>>>(vo, n8, ot.pos, ot.pair);
    SVR_CHECK_LAUNCH();
    const uint32_t meta[4] = {required_bits((uint64_t)n8), (uint32_t)key_bits, (uint32_t)finest, 0u};
    SVR_CHECK_HIP(hipMemcpyAsync(ot.meta, meta, sizeof(meta), hipMemcpyHostToDevice, stream));
    SVR_CHECK_HIP(hipStreamSynchronize(stream));
    return SVR_OK;
}

int svr_unpack_geom_state(int64_t N, const void* geom_state, uint32_t* n_dup, uint32_t* bbox, uint32_t* qbits,
                          svr_stream_t stream)
{
    if (N == 0) return SVR_OK;
    GeomState g = GeomState::carve((void*)geom_state, N, nullptr);
    SVR_CHECK_HIP(hipMemcpyAsync(n_dup, g.n_dup, sizeof(uint32_t) * N, hipMemcpyDeviceToDevice, stream));
    SVR_CHECK_HIP(hipMemcpyAsync(bbox, g.bbox, sizeof(uint2) * N, hipMemcpyDeviceToDevice, stream));
    SVR_CHECK_HIP(hipMemcpyAsync(qbits, g.qbits, sizeof(uint32_t) * N, hipMemcpyDeviceToDevice, stream));
    return SVR_OK;
}

int svr_export_sort_keys(int32_t W, int32_t H, int64_t R, const void* binning_state, const void* order_table_buf,
                         const int64_t* paths, int64_t N, uint64_t* hi, uint64_t* lo, uint32_t* vals,
                         svr_stream_t stream)
{
    (void)W; (void)H; (void)order_table_buf;
    if (R == 0) return SVR_OK;
    BinningState b = BinningState::carve((void*)binning_state, R, nullptr);
    const int pos_bits = (int)required_bits((uint64_t)8 * N);
    export_keys_kernel<<<(unsigned)((R + 255) / 256), 256, 0, stream>>>(b.keys, b.vox_list, R, pos_bits, paths, hi,
                                                                        lo, vals);
    SVR_CHECK_LAUNCH();
    return SVR_OK;
}

}  // extern "C"
