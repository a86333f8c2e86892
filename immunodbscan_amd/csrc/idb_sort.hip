// Stable LSD radix sort of (u64 key, u32 value) pairs for the entry order of
// runs up to IDB_SORT_SMALL_MAX entries.  At these sizes a general device
// sort spends its time on launches (a block sort, then merge passes of three
// kernels each); this one takes two launches per digit pass:
//   count:   each block histograms the digit over its tile of RS_TILE pairs
//            into column b of a 256 x nb count matrix;
//   scatter: each block derives its own offsets from the matrix (per digit:
//            pairs of all blocks, scanned over digits, plus pairs of lower
//            blocks), ranks its pairs among the equal digits before them in
//            the tile, and scatters.  Wave w owns a contiguous run of RS_IPT
//            slots of 64 pairs; lanes with the same digit in one slot find
//            each other with one ballot per digit bit, and a per-wave counter
//            per digit (LDS) carries the count across the wave's slots.
// A pair's rank counts only pairs before it in input order (lower block,
// lower wave, lower slot, lower lane), so equal keys keep their order.
// Passes alternate between the two buffer pairs, as few as the key's bits
// allow (8 bits or fewer per pass); the caller learns which pair holds the
// result.
#include "idb_internal.h"

namespace {

constexpr uint32_t RS_THREADS = 256;
constexpr uint32_t RS_WAVES = RS_THREADS / 64;
constexpr uint32_t RS_IPT = 8;  // slots of 64 pairs per wave
constexpr uint32_t RS_TILE = RS_THREADS * RS_IPT;

__global__ __launch_bounds__(RS_THREADS) void k_rs_count(const uint64_t *ks, uint32_t n, uint32_t shift,
                                                         uint32_t dmask, uint32_t *gcnt)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t b = blockIdx.x;
#pragma unroll
    for (uint32_t k = 0; k < RS_IPT; ++k) {
        const uint32_t i = b * RS_TILE + k * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(ks[i] >> shift) & dmask], 1u);
    }
    __syncthreads();
    gcnt[(size_t)threadIdx.x * gridDim.x + b] = h[threadIdx.x];
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t *ks, const uint32_t *vs, uint64_t *kd,
                                                           uint32_t *vd, uint32_t n, uint32_t shift, uint32_t dbits,
                                                           const uint32_t *gcnt)
{
    __shared__ uint32_t wcnt[RS_WAVES][256];  // per wave: count, then exclusive prefix over waves
    __shared__ uint32_t sbase[256];           // digit start + pairs of lower blocks
    __shared__ uint32_t wsum[RS_WAVES];
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t dmask = (1u << dbits) - 1u;
    const uint64_t lanes_below = (1ull << lane) - 1ull;
    for (uint32_t x = 0; x < RS_WAVES; ++x) wcnt[x][threadIdx.x] = 0u;
    // this block's offsets: per digit, pairs of all blocks and of lower blocks
    uint32_t tot = 0, pre = 0;
    {
        // eight loads in flight at a time
        const uint32_t *g = gcnt + (size_t)threadIdx.x * nb;
        uint32_t x = 0;
        for (; x + 8 <= nb; x += 8) {
            uint32_t c[8];
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) c[q] = g[x + q];
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) {
                tot += c[q];
                pre += x + q < b ? c[q] : 0u;
            }
        }
        for (; x < nb; ++x) {
            const uint32_t c = g[x];
            tot += c;
            pre += x < b ? c : 0u;
        }
    }
    uint32_t inc = tot;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    {
        uint32_t off = 0;
        for (uint32_t x = 0; x < w; ++x) off += wsum[x];
        sbase[threadIdx.x] = off + (inc - tot) + pre;
    }
    uint64_t key[RS_IPT];
    uint32_t val[RS_IPT], dr[RS_IPT];  // dr = digit | rank within the wave << 8
    const uint32_t base = b * RS_TILE + w * 64 * RS_IPT + lane;
#pragma unroll
    for (uint32_t k = 0; k < RS_IPT; ++k) {
        const uint32_t i = base + k * 64;
        const bool ok = i < n;
        key[k] = ok ? ks[i] : 0ull;
        val[k] = ok ? vs[i] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < RS_IPT; ++k) {
        const bool ok = base + k * 64 < n;
        const uint32_t d = (uint32_t)(key[k] >> shift) & dmask;
        uint64_t m = __ballot(ok);
        for (uint32_t t = 0; t < dbits; ++t) {
            const uint64_t bt = __ballot((d >> t) & 1u);
            m &= ((d >> t) & 1u) ? bt : ~bt;
        }
        uint32_t r = 0;
        if (ok) r = wcnt[w][d] + (uint32_t)__builtin_popcountll(m & lanes_below);
        // the lowest lane of each digit group moves the wave's counter on,
        // after every lane of the group has read it
        if (ok && (m & lanes_below) == 0ull) wcnt[w][d] += (uint32_t)__builtin_popcountll(m);
        dr[k] = d | (r << 8);
    }
    __syncthreads();
    {
        const uint32_t d = threadIdx.x;
        uint32_t s = 0;
        for (uint32_t x = 0; x < RS_WAVES; ++x) {
            const uint32_t c = wcnt[x][d];
            wcnt[x][d] = s;
            s += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < RS_IPT; ++k) {
        if (base + k * 64 < n) {
            const uint32_t d = dr[k] & 0xFFu;
            const uint32_t pos = sbase[d] + wcnt[w][d] + (dr[k] >> 8);
            if (pos < n) {  // always, when the counts match the tiles
                kd[pos] = key[k];
                vd[pos] = val[k];
            }
        }
    }
}

}  // namespace

uint32_t idb_sort_small_blocks(uint32_t n) { return (n + RS_TILE - 1) / RS_TILE; }

hipError_t idb_sort_small(uint64_t *k0, uint64_t *k1, uint32_t *v0, uint32_t *v1, uint32_t n, uint32_t end_bit,
                          uint32_t *gcnt, hipStream_t s, bool *in_first)
{
    *in_first = false;
    if (n == 0) return hipSuccess;
    if (n > IDB_SORT_SMALL_MAX || end_bit == 0 || end_bit > 64) return hipErrorInvalidValue;
    const uint32_t passes = (end_bit + 7) / 8;
    *in_first = passes % 2 == 0;
    const uint32_t wd = (end_bit + passes - 1) / passes;
    if ((passes - 1) * wd >= end_bit) return hipErrorInvalidValue;  // every pass must hold a digit bit
    const uint32_t nb = idb_sort_small_blocks(n);
    for (uint32_t p = 0; p < passes; ++p) {
        const uint64_t *ks = (p & 1) ? k1 : k0;
        const uint32_t *vs = (p & 1) ? v1 : v0;
        uint64_t *kd = (p & 1) ? k0 : k1;
        uint32_t *vd = (p & 1) ? v0 : v1;
        const uint32_t shift = p * wd;
        const uint32_t dbits = end_bit - shift < wd ? end_bit - shift : wd;
        hipLaunchKernelGGL(k_rs_count, dim3(nb), dim3(RS_THREADS), 0, s, ks, n, shift, (1u << dbits) - 1u, gcnt);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(RS_THREADS), 0, s, ks, vs, kd, vd, n, shift, dbits, gcnt);
    }
    return hipGetLastError();
}
