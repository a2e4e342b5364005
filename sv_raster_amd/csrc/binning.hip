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
#include "voxel_ops.h"
#include "svr_host.h"

namespace svr {


__device__ inline uint64_t match_digit(uint32_t d, uint64_t active)
{
    uint64_t m = active;
#pragma unroll
    for (int b = 0; b < RADIX_BITS; ++b)
    {
        const uint32_t bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        m &= bit ? bal : ~bal;
    }
    return m;
}

// Number of keys to sort: n, or the smaller count a previous kernel left at
// n_dev (a 64-bit count as two 32-bit words) when n_dev is given.
__device__ inline int64_t sort_count(int64_t n, const uint32_t* __restrict__ n_dev)
{
    if (!n_dev) return n;
    const int64_t d = (int64_t)(((uint64_t)n_dev[1] << 32) | n_dev[0]);
    return d < n ? d : n;
}

__global__ __launch_bounds__(SORT_THREADS) void radix_upsweep(const uint64_t* __restrict__ keys, int64_t n_cap,
                                                              const uint32_t* __restrict__ n_dev, int shift,
                                                              uint32_t dmask, uint32_t* __restrict__ counts,
                                                              int64_t nb)
{
    const int64_t n = sort_count(n_cap, n_dev);
    __shared__ uint32_t hist[SORT_THREADS / 64][RADIX];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
    for (int i = 0; i < SORT_THREADS / 64; ++i) hist[i][t] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * SORT_TILE + (int64_t)w * (SORT_ITEMS * 64);
    for (int r = 0; r < SORT_ITEMS; ++r)
    {
        const int64_t idx = base + r * 64 + lane;
        const bool valid = idx < n;
        const uint64_t active = __ballot(valid);
        if (active == 0) break;
        const uint32_t d = valid ? (uint32_t)(keys[idx] >> shift) & dmask : 0u;
        const uint64_t m = match_digit(d, active);
        if (valid && lane == 63 - __clzll(m)) hist[w][d] += (uint32_t)__popcll(m);
    }
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < SORT_THREADS / 64; ++i) c += hist[i][t];
    counts[(int64_t)t * nb + blockIdx.x] = c;
}

// Exclusive scan of each digit row of counts[RADIX][nb]; row totals -> counts[RADIX*nb + d].
__global__ __launch_bounds__(256) void radix_scan_rows(uint32_t* __restrict__ counts, int64_t nb)
{
    __shared__ uint32_t lds[4];
    __shared__ uint32_t carry_s;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t* row = counts + (int64_t)blockIdx.x * nb;
    if (t == 0) carry_s = 0;
    __syncthreads();
    for (int64_t off = 0; off < nb; off += 256)
    {
        const int64_t i = off + t;
        const uint32_t x = i < nb ? row[i] : 0u;
        uint32_t s = x;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const uint32_t y = __shfl_up(s, o, 64);
            if (lane >= o) s += y;
        }
        if (lane == 63) lds[w] = s;
        __syncthreads();
        uint32_t pre = 0;
        for (int k = 0; k < w; ++k) pre += lds[k];
        const uint32_t tot = lds[0] + lds[1] + lds[2] + lds[3];
        const uint32_t carry = carry_s;
        if (i < nb) row[i] = carry + pre + s - x;
        __syncthreads();
        if (t == 0) carry_s = carry + tot;
        __syncthreads();
    }
    if (t == 0) counts[(int64_t)RADIX * nb + blockIdx.x] = carry_s;
}

// Ranks each key of the block (stable within its digit), reorders the block's
// keys in LDS by digit, then writes each digit's run to its global position
// with consecutive threads on consecutive addresses (coalesced stores instead
// of one scattered 12-byte store per key).
template <bool HAS_VALS>
__global__ __launch_bounds__(SORT_THREADS) void radix_downsweep(
    const uint64_t* __restrict__ keys_in, uint64_t* __restrict__ keys_out, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ vals_out, int64_t n_cap, const uint32_t* __restrict__ n_dev, int shift, uint32_t dmask,
    const uint32_t* __restrict__ counts, int64_t nb)
{
    const int64_t n = sort_count(n_cap, n_dev);
    constexpr int NW = SORT_THREADS / 64;
    __shared__ uint32_t wcnt[NW][RADIX];
    __shared__ uint32_t gbase[RADIX];  // global start of this block's run of digit d
    __shared__ uint32_t lbase[RADIX];  // start of digit d in the block's reordered keys
    __shared__ uint32_t wsum[NW];
    __shared__ uint64_t lkeys[SORT_TILE];  // reordered keys, then (as 32-bit words) values
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;

    // digit base = exclusive prefix over digit totals + this block's row offset
    {
        const uint32_t tot = counts[(int64_t)RADIX * nb + t];
        uint32_t s = tot;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const uint32_t y = __shfl_up(s, o, 64);
            if (lane >= o) s += y;
        }
        if (lane == 63) wsum[w] = s;
        __syncthreads();
        uint32_t pre = 0;
        for (int k = 0; k < w; ++k) pre += wsum[k];
        gbase[t] = pre + s - tot + counts[(int64_t)t * nb + blockIdx.x];
#pragma unroll
        for (int i = 0; i < NW; ++i) wcnt[i][t] = 0;
    }
    __syncthreads();

    const int64_t blk0 = (int64_t)blockIdx.x * SORT_TILE;
    const int64_t base = blk0 + (int64_t)w * (SORT_ITEMS * 64);
    uint64_t k[SORT_ITEMS];
    uint32_t v[SORT_ITEMS];
    uint32_t rk[SORT_ITEMS];
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; ++r)
    {
        const int64_t idx = base + r * 64 + lane;
        const bool valid = idx < n;
        k[r] = valid ? keys_in[idx] : 0ull;
        if (HAS_VALS) v[r] = valid ? vals_in[idx] : 0u;
        const uint64_t active = __ballot(valid);
        const uint32_t d = (uint32_t)(k[r] >> shift) & dmask;
        const uint64_t m = match_digit(d, active);
        const uint32_t before = (uint32_t)__popcll(m & lanemask_lt(lane));
        uint32_t cnt = 0;
        if (valid) cnt = wcnt[w][d];
        rk[r] = cnt + before;
        if (valid && lane == 63 - __clzll(m)) wcnt[w][d] = cnt + (uint32_t)__popcll(m);
    }
    __syncthreads();
    // per digit: exclusive offsets of the waves, and the block's digit total
    uint32_t dtot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i)
    {
        const uint32_t c = wcnt[i][t];
        wcnt[i][t] = dtot;
        dtot += c;
    }
    // exclusive scan of the digit totals over the block (digit = thread)
    {
        uint32_t s = dtot;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const uint32_t y = __shfl_up(s, o, 64);
            if (lane >= o) s += y;
        }
        __syncthreads();
        if (lane == 63) wsum[w] = s;
        __syncthreads();
        uint32_t pre = 0;
        for (int q = 0; q < w; ++q) pre += wsum[q];
        lbase[t] = pre + s - dtot;
    }
    __syncthreads();
    uint32_t lp[SORT_ITEMS];
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; ++r)
    {
        const int64_t idx = base + r * 64 + lane;
        const uint32_t d = (uint32_t)(k[r] >> shift) & dmask;
        lp[r] = lbase[d] + wcnt[w][d] + rk[r];
        if (idx < n) lkeys[lp[r]] = k[r];
    }
    __syncthreads();
    const int nk = (int)min((int64_t)SORT_TILE, n - blk0);
    // the digit of reordered position i, kept for the values pass
    uint32_t dd[SORT_ITEMS];
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; ++r)
    {
        const int i = t + r * SORT_THREADS;
        dd[r] = 0;
        if (i < nk)
        {
            const uint64_t key = lkeys[i];
            const uint32_t d = (uint32_t)(key >> shift) & dmask;
            dd[r] = d;
            keys_out[gbase[d] + (uint32_t)i - lbase[d]] = key;
        }
    }
    if (HAS_VALS)
    {
        uint32_t* lv = reinterpret_cast<uint32_t*>(lkeys);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < SORT_ITEMS; ++r)
        {
            const int64_t idx = base + r * 64 + lane;
            if (idx < n) lv[lp[r]] = v[r];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < SORT_ITEMS; ++r)
        {
            const int i = t + r * SORT_THREADS;
            if (i < nk) vals_out[gbase[dd[r]] + (uint32_t)i - lbase[dd[r]]] = lv[i];
        }
    }
}

// Stable LSD radix sort over bits [lo_bit, lo_bit + nbits) in digits of at
// most 8 bits.  The sorted data ends in the buffers returned through
// out_keys / out_vals.
int radix_sort_ex(uint64_t* keys, uint64_t* keys_alt, uint32_t* vals, uint32_t* vals_alt, int64_t n, int nbits,
                  uint32_t* counts, hipStream_t stream, uint64_t** out_keys, uint32_t** out_vals,
                  const uint32_t* n_dev, int lo_bit)
{
    *out_keys = keys;
    if (out_vals) *out_vals = vals;
    if (n <= 0 || nbits <= 0) return SVR_OK;
    SVR_REQUIRE(lo_bit >= 0 && lo_bit + nbits <= 64, SVR_ERR_INVALID, "radix sort bits [%d, %d) out of range", lo_bit,
                lo_bit + nbits);
    const int64_t nb = sort_blocks(n);
    SVR_REQUIRE(nb < (1ll << 31), SVR_ERR_INVALID, "radix sort of %lld keys too large", (long long)n);
    uint64_t* kin = keys; uint64_t* kout = keys_alt;
    uint32_t* vin = vals; uint32_t* vout = vals_alt;
    // radix_passes(nbits) passes with the bits spread evenly over them
    const int passes = radix_passes(nbits);
    const int step = (nbits + passes - 1) / passes;
    for (int done = 0; done < nbits; done += step)
    {
        const int shift = lo_bit + done;
        const int bits = (nbits - done) < step ? (nbits - done) : step;
        const uint32_t dmask = (1u << bits) - 1u;
        radix_upsweep<<<(unsigned)nb, SORT_THREADS, 0, stream>>>(kin, n, n_dev, shift, dmask, counts, nb);
        SVR_CHECK_LAUNCH();
        radix_scan_rows<<<RADIX, 256, 0, stream>>>(counts, nb);
        SVR_CHECK_LAUNCH();
        if (vals)
            radix_downsweep<true><<<(unsigned)nb, SORT_THREADS, 0, stream>>>(kin, kout, vin, vout, n, n_dev, shift,
                                                                             dmask, counts, nb);
        else
            radix_downsweep<false><<<(unsigned)nb, SORT_THREADS, 0, stream>>>(kin, kout, nullptr, nullptr, n, n_dev,
                                                                              shift, dmask, counts, nb);
        SVR_CHECK_LAUNCH();
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
                                  uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    const uint32_t q = (uint32_t)(i & 7);
    const uint64_t rank = (uint64_t)paths[i >> 3] ^ order_table(q);
    keys[i] = with_q ? (((rank >> lshift) << 3) | q) : rank;
    vals[i] = (uint32_t)i;
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
        s.counts = c.take<uint32_t>(RADIX * sort_blocks(n8 > 0 ? n8 : 1) + RADIX);
        s.misc = c.take<uint32_t>(4);
        if (bytes) *bytes = c.size();
        return s;
    }
};

// ---- per-frame duplication, ranges and key export --------------------

// Each active (item, quadrant) pair's tile count at its rank position, in
// view blockIdx.y's block of pos_off (pos_off cleared beforehand).
__global__ __launch_bounds__(256) void pair_count_kernel(int64_t N, const uint32_t* __restrict__ n_dup,
                                                         const uint32_t* __restrict__ qbits,
                                                         const uint32_t* __restrict__ order_pos,
                                                         uint32_t* __restrict__ pos_off)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const size_t p = (size_t)blockIdx.y * (size_t)N + (size_t)v;
    const uint32_t nd = n_dup[p];
    if (nd == 0) return;
    const uint32_t bits = qbits[p];
    const uint32_t tiles = nd / (uint32_t)__popc(bits);
    uint32_t* blk = pos_off + 8 * (size_t)blockIdx.y * (size_t)N;
    for (uint32_t b = bits; b; b &= b - 1u)
    {
        const uint32_t q = (uint32_t)__builtin_ctz(b);
        blk[order_pos[8 * v + q]] = tiles;
    }
}

int pair_offsets_launch(int V, int64_t N, const GeomState& g, const uint32_t* order_pos, hipStream_t stream)
{
    const int64_t n8 = 8 * (int64_t)V * N;
    if (n8 == 0) return SVR_OK;
    if (fill_u32(g.pos_off, 0u, n8, stream) != SVR_OK) return SVR_ERR_HIP;
    pair_count_kernel<<<dim3((unsigned)((N + 255) / 256), (unsigned)V), 256, 0, stream>>>(N, g.n_dup, g.qbits,
                                                                                         order_pos, g.pos_off);
    SVR_CHECK_LAUNCH();
    return pair_scan_launch(g.pos_off, n8, g.pos_sums, stream);
}

// forward.cu:523-577, emitting (tile << pos_bits | pos) keys: each quadrant's
// run at its rank-ordered offset (recorded in qoff for the later passes).
__global__ __launch_bounds__(256) void duplicate_kernel(GeomState g, int64_t N, int gx, int gy, int pos_bits,
                                                        const uint32_t* __restrict__ order_pos,
                                                        uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
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
                vals[off] = off;
                ++off;
            }
    }
    if (large) queue_large(g.big_list, g.counters + 2, N, nd, (uint32_t)v);
}

// Large voxels: one workgroup each; then the huge ones with every workgroup
// taking a strided share of each.  Entry e of a voxel: quadrant (the qi-th set
// bit of its bits) major, then the tile rows of its rectangle, written into
// that quadrant's run.
__global__ __launch_bounds__(256) void duplicate_big_kernel(GeomState g, int64_t N, int gx, int gy, int pos_bits,
                                                            const uint32_t* __restrict__ order_pos,
                                                            uint64_t* __restrict__ keys,
                                                            uint32_t* __restrict__ vals)
{
    // pass 0: one workgroup per large voxel; pass 1: each huge voxel split into
    // K entry ranges taken by K different workgroups (item = chunk * huge + voxel)
    const uint32_t count = g.counters[2], huge = g.counters[3];
    const uint32_t K = huge ? max(1u, gridDim.x / huge) : 1u;
    for (uint32_t pass = 0; pass < 2; ++pass)
    {
        for (uint32_t it = blockIdx.x; it < (pass == 0 ? count : huge * K); it += gridDim.x)
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
                const uint32_t off = run_offset(g, v, bits, q);
                const uint64_t pos = order_pos[8 * (size_t)v + q];
                for (uint32_t e = r_lo + threadIdx.x; e < r_hi; e += blockDim.x)
                {
                    const uint32_t rem = e - qi * tiles;
                    const uint32_t y = (uint32_t)tymin + rem / w, x = (uint32_t)txmin + rem % w;
                    keys[off + rem] = ((uint64_t)(y * (uint32_t)gx + x) << pos_bits) | pos;
                    vals[off + rem] = off + rem;
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
                                                              uint64_t* __restrict__ qkeys,
                                                              uint32_t* __restrict__ order)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= T) return;
    const uint2 r = ranges[g];
    qkeys[g] = 255u - length_bucket(r.y - r.x);
    order[g] = (uint32_t)g;
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
    tile_queue_keys_kernel<<<(unsigned)((T + 255) / 256), 256, 0, stream>>>(T, ranges, qkeys_alt, order_alt);
    SVR_CHECK_LAUNCH();
    uint64_t* qk = nullptr;
    uint32_t* qo = nullptr;
    const int rc = radix_sort_ex(qkeys_alt, qkeys, order_alt, order, T, 8, qcounts, stream, &qk, &qo);
    if (rc != SVR_OK) return rc;
    SVR_REQUIRE(qo == order, SVR_ERR_HIP, "tile queue sort ended in the wrong buffer");
    return SVR_OK;
}

int pair_prepare(int64_t N, const GeomState& g, const OrderTable& ot, hipStream_t stream)
{
    SVR_REQUIRE(fill_u32(g.counters + 2, 0u, 2, stream) == SVR_OK, SVR_ERR_HIP, "fill counters failed");
    return pair_offsets_launch(1, N, g, ot.pos, stream);
}

int binning_forward(int W, int H, int64_t N, const GeomState& g, const OrderTable& ot, int64_t R,
                    BinningState& b, ImageState& img, hipStream_t stream, bool pairs_ready)
{
    const int gx = tiles_x(W), gy = tiles_y(H);
    const int64_t T = (int64_t)gx * gy;
    prof_begin(PROF_BINNING, stream);
    SVR_REQUIRE(fill_u32(img.ranges, 0u, 2 * T, stream) == SVR_OK, SVR_ERR_HIP, "fill ranges failed");
    if (R == 0)
    {
        const int rc0 = tile_queue_launch(T, img.ranges, img.qkeys, img.qkeys_alt, img.order, img.order_alt,
                                          img.qcounts, stream, img.ctr);
        prof_end(PROF_BINNING, stream);
        return rc0;
    }
    const int pos_bits = (int)required_bits((uint64_t)8 * N);
    const int tile_bits = (int)required_bits((uint64_t)T);
    int rc = pairs_ready ? SVR_OK : pair_prepare(N, g, ot, stream);
    if (rc != SVR_OK) return rc;
    // entries are emitted in rank order, so the sort only orders the tile
    // bits; it ping-pongs between the two buffers, one digit per pass: emit
    // into the buffer that makes the sorted result land in b.keys / b.orig
    const bool odd = radix_passes(tile_bits) & 1;
    uint64_t* k0 = odd ? b.keys_alt : b.keys;
    uint32_t* o0 = odd ? b.orig_alt : b.orig;
    duplicate_kernel<<<(unsigned)((N + 255) / 256), 256, 0, stream>>>(g, N, gx, gy, pos_bits, ot.pos, k0, o0);
    SVR_CHECK_LAUNCH();
    duplicate_big_kernel<<<1024, 256, 0, stream>>>(g, N, gx, gy, pos_bits, ot.pos, k0, o0);
    SVR_CHECK_LAUNCH();
    uint64_t* sorted = nullptr;
    uint32_t* sorted_orig = nullptr;
    rc = radix_sort_ex(k0, odd ? b.keys : b.keys_alt, o0, odd ? b.orig : b.orig_alt, R, tile_bits, b.counts, stream,
                       &sorted, &sorted_orig, nullptr, pos_bits);
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
    order_keys_kernel<<<(unsigned)((n8 + 255) / 256), 256, 0, stream>>>(paths, n8, lshift, with_q, s.keys, s.vals);
    SVR_CHECK_LAUNCH();
    uint64_t* ko = nullptr;
    uint32_t* vo = nullptr;
    int rc = radix_sort_ex(s.keys, s.keys_alt, s.vals, s.vals_alt, n8, key_bits, s.counts, stream, &ko, &vo);
    if (rc != SVR_OK) return rc;
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
