// idb_scan.hip — single-launch exclusive scans and flag selection for the
// sizing and renumbering steps of a run (entry offsets, bucket starts, query
// blocks, tile offsets, the uniques list, cluster ranks).
//
// One launch per scan, by decoupled look-back (D. Merrill, M. Garland,
// "Single-pass Parallel Prefix Scan with Decoupled Look-back", NVIDIA
// technical report NVR-2016-002): each block scans one tile of
// SCAN_TILE values, publishes the tile's aggregate, adds the aggregates of
// the tiles before it (or the first inclusive prefix it meets) and
// publishes its inclusive prefix.
//
// * Tiles are taken from a ticket counter, so a block only ever waits on
//   tiles that blocks already running have taken: every wait ends.  The
//   block that takes the last ticket resets the counter for the next launch.
// * Every launch has a generation (the caller's launch number, never 0).  A
//   tile's flag word carries it, so the status words of earlier launches
//   never need clearing; the workspace is cleared when it is allocated.
// * Status per tile: aggregate, inclusive prefix, and the flag word
//   (generation << 2 | 1 for the aggregate, 2 for the prefix).  All status
//   accesses are agent-scope atomics, which on gfx950 bypass the XCD-private
//   L2 state (sc1); a value is stored, its store waited for (vmcnt), and only
//   then the flag, and a reader loads a value only after its flag.  No
//   release or acquire fence is used: on gfx950 those write back or
//   invalidate the whole L2 of the XCD, which cost more than the scan.
// * A wait that does not end within about a second (it cannot happen while
//   every block runs) sets the error word instead of hanging; the run then
//   fails.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "idb_internal.h"

namespace {

constexpr uint32_t SCAN_THREADS = 256, SCAN_IPT = 16, SCAN_TILE = SCAN_THREADS * SCAN_IPT;
constexpr uint64_t FLAG_AGG = 1, FLAG_PREFIX = 2;
constexpr uint32_t SPIN_LIMIT = 1u << 22;

struct ScanState {
    uint32_t ticket;
    uint32_t gen;
    uint32_t err;
    uint32_t pad;
};

__device__ __forceinline__ uint64_t ld_status(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// a value and then its flag word: the flag goes out once the value's store
// has completed
__device__ __forceinline__ void publish(uint64_t *value, uint64_t v, uint64_t *flag, uint64_t f)
{
    __hip_atomic_store(value, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __hip_atomic_store(flag, f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One tile: values v (SEL: 1 for a set flag) of indices
// [tile * SCAN_TILE, ...), thread t holding SCAN_IPT consecutive values.
// Writes the exclusive prefix of every value (SEL: the index of every set
// flag at its prefix) and, from the block of the last tile, the total.
template <typename T, bool SEL>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan(const T *in, const uint8_t *flags, T *out, uint32_t *sel,
                                                       uint32_t *count, uint64_t n, ScanState *st, uint64_t *status,
                                                       uint64_t gen)
{
    __shared__ uint32_t s_tile;
    __shared__ T s_wave[SCAN_THREADS / 64];
    __shared__ T s_prefix;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid == 0) {
        const uint32_t t = __hip_atomic_fetch_add(&st->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t == gridDim.x - 1) __hip_atomic_store(&st->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_tile = t;
    }
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t base = (uint64_t)tile * SCAN_TILE + (uint64_t)tid * SCAN_IPT;

    // a thread's SCAN_IPT consecutive values in 16-byte loads when they are
    // all in range (the arrays are 256-byte aligned), else one by one
    T v[SCAN_IPT];
    const bool full = base + SCAN_IPT <= n &&
                      ((uintptr_t)(SEL ? (const void *)flags : (const void *)in) & 15u) == 0u;
    if (full) {
        if (SEL) {
            const uint4 f = *(const uint4 *)(flags + base);
            const uint32_t fw[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
            for (uint32_t i = 0; i < SCAN_IPT; ++i) v[i] = ((fw[i >> 2] >> (8 * (i & 3))) & 0xFFu) ? T(1) : T(0);
        } else {
            constexpr uint32_t PER4 = 16 / sizeof(T);
            const uint4 *src = (const uint4 *)(in + base);
#pragma unroll
            for (uint32_t q = 0; q < SCAN_IPT / PER4; ++q) {
                const uint4 x = src[q];
                if (sizeof(T) == 4) {
                    v[4 * q] = (T)x.x;
                    v[4 * q + 1] = (T)x.y;
                    v[4 * q + 2] = (T)x.z;
                    v[4 * q + 3] = (T)x.w;
                } else {
                    v[2 * q] = (T)(((uint64_t)x.y << 32) | x.x);
                    v[2 * q + 1] = (T)(((uint64_t)x.w << 32) | x.z);
                }
            }
        }
    } else {
#pragma unroll
        for (uint32_t i = 0; i < SCAN_IPT; ++i) {
            const uint64_t k = base + i;
            if (SEL) v[i] = (k < n && flags[k]) ? T(1) : T(0);
            else v[i] = k < n ? in[k] : T(0);
        }
    }
    T sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < SCAN_IPT; ++i) sum += v[i];
    // inclusive scan of the thread sums over the wave
    T incl = sum;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const T u = __shfl_up(incl, o, 64);
        if (lane >= o) incl += u;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    T wave_pre = 0, agg = 0;
#pragma unroll
    for (uint32_t w = 0; w < SCAN_THREADS / 64; ++w) {
        if (w < wave) wave_pre += s_wave[w];
        agg += s_wave[w];
    }
    if (wave == 0) {
        // The first wave publishes the aggregate and looks back over the
        // 64 tiles before a window's end at once, one tile per lane: the
        // window's sum runs up to the nearest tile with an inclusive prefix.
        uint64_t *my = status + 3ull * tile;
        if (lane == 0) {
            if (tile == 0) publish(my + 1, (uint64_t)agg, my + 2, (gen << 2) | FLAG_PREFIX);
            else publish(my, (uint64_t)agg, my + 2, (gen << 2) | FLAG_AGG);
        }
        T prefix = 0;
        int64_t hi = (int64_t)tile - 1;
        bool failed = false;
        while (hi >= 0) {
            const int64_t j = hi - (int64_t)lane;
            uint64_t f = (gen << 2) | FLAG_PREFIX;  // before tile 0: a prefix of 0
            T val = 0;
            if (j >= 0) {
                uint32_t spins = 0;
                for (;;) {
                    f = ld_status(status + 3ull * j + 2);
                    if ((f >> 2) == gen) break;
                    if (++spins > SPIN_LIMIT) {
                        failed = true;
                        f = (gen << 2) | FLAG_PREFIX;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                if (!failed) val = (T)ld_status(status + 3ull * j + ((f & 3u) == FLAG_PREFIX ? 1 : 0));
            }
            const uint64_t pm = __ballot((f & 3u) == FLAG_PREFIX);
            // lanes up to the first with a prefix (all 64 when none has one)
            const uint32_t last = pm ? (uint32_t)__builtin_ctzll(pm) : 63u;
            T part = lane <= last ? val : T(0);
#pragma unroll
            for (uint32_t o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
            prefix += part;
            if (pm) break;
            hi -= 64;
        }
        if (__ballot(failed) && lane == 0) __hip_atomic_store(&st->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (lane == 0) {
            if (tile != 0) publish(my + 1, (uint64_t)(prefix + agg), my + 2, (gen << 2) | FLAG_PREFIX);
            s_prefix = prefix;
            if (SEL && tile == gridDim.x - 1) *count = (uint32_t)(prefix + agg);
        }
    }
    __syncthreads();
    T run = s_prefix + wave_pre + (incl - sum);
    if (!SEL && full && ((uintptr_t)out & 15u) == 0u) {
        T o[SCAN_IPT];
#pragma unroll
        for (uint32_t i = 0; i < SCAN_IPT; ++i) {
            o[i] = run;
            run += v[i];
        }
        uint4 *dst = (uint4 *)(out + base);
#pragma unroll
        for (uint32_t q = 0; q < SCAN_IPT * sizeof(T) / 16; ++q) {
            if (sizeof(T) == 4)
                dst[q] = make_uint4((uint32_t)o[4 * q], (uint32_t)o[4 * q + 1], (uint32_t)o[4 * q + 2],
                                    (uint32_t)o[4 * q + 3]);
            else
                dst[q] = make_uint4((uint32_t)o[2 * q], (uint32_t)((uint64_t)o[2 * q] >> 32), (uint32_t)o[2 * q + 1],
                                    (uint32_t)((uint64_t)o[2 * q + 1] >> 32));
        }
        return;
    }
#pragma unroll
    for (uint32_t i = 0; i < SCAN_IPT; ++i) {
        const uint64_t k = base + i;
        if (SEL) {
            if (v[i]) sel[run] = (uint32_t)k;
        } else if (k < n) {
            out[k] = run;
        }
        run += v[i];
    }
}

uint32_t tiles_of(uint64_t n) { return n ? (uint32_t)((n + SCAN_TILE - 1) / SCAN_TILE) : 1u; }

}  // namespace

size_t idb_scan_workspace_bytes(uint64_t max_n) { return 256 + 24ull * tiles_of(max_n); }

hipError_t idb_scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint64_t n, void *ws, uint32_t gen, hipStream_t s)
{
    ScanState *st = (ScanState *)ws;
    uint64_t *status = (uint64_t *)((uint8_t *)ws + 256);
    hipLaunchKernelGGL((k_scan<uint32_t, false>), dim3(tiles_of(n)), dim3(SCAN_THREADS), 0, s, in, nullptr, out, nullptr,
                       nullptr, n, st, status, (uint64_t)gen);
    return hipGetLastError();
}

hipError_t idb_scan_exclusive_u64(const uint64_t *in, uint64_t *out, uint64_t n, void *ws, uint32_t gen, hipStream_t s)
{
    ScanState *st = (ScanState *)ws;
    uint64_t *status = (uint64_t *)((uint8_t *)ws + 256);
    hipLaunchKernelGGL((k_scan<uint64_t, false>), dim3(tiles_of(n)), dim3(SCAN_THREADS), 0, s, in, nullptr, out, nullptr,
                       nullptr, n, st, status, (uint64_t)gen);
    return hipGetLastError();
}

hipError_t idb_select_flagged(const uint8_t *flags, uint32_t *out, uint32_t *count, uint64_t n, void *ws, uint32_t gen,
                              hipStream_t s)
{
    ScanState *st = (ScanState *)ws;
    uint64_t *status = (uint64_t *)((uint8_t *)ws + 256);
    hipLaunchKernelGGL((k_scan<uint32_t, true>), dim3(tiles_of(n)), dim3(SCAN_THREADS), 0, s, nullptr, flags, nullptr,
                       out, count, n, st, status, (uint64_t)gen);
    return hipGetLastError();
}

uint32_t idb_scan_error_offset() { return (uint32_t)offsetof(ScanState, err); }
