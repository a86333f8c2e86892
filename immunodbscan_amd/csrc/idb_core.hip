// idb_core.hip — host driver and O(N) kernels of libidb_hip.so (gfx950).
//
// Pipeline of one call (stage names follow the reference's timers,
// cuda/PSDBSCAN.hcu:1301-1354):
//  Identify Duplicates      row scan (64-bit content hash, letter set, max
//                           length) -> radix sort of (hash, row) -> each row's
//                           representative = first equal row of its hash run.
//                           Replaces the O(N^2) singlet_find_duplicates_kernel
//                           (:525-573) with O(N log N) work; equality is
//                           checked on the full content, so collisions only
//                           cost time.
//  Determine copies         class sizes, uniques (ascending), count init (:1074-1111)
//  Identify core points     bucket entries (unique row x V gene x J gene),
//                           radix sort by (repertoire, bucket, length), tile
//                           list, count pass (idb_pairs.hip), core flags (:1097-1184)
//  Grow clusters            fuse pass over the same tiles (:1189-1260)
//  Reduction of cluster information tree structure
//                           roots of cores, border roots (:1007-1017)
//  Re-number clusters and identify noise
//                           exclusive scan of root flags -> labels (:1263-1294)
#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <string_view>
#include <array>
#include <map>
#include <mutex>
#include <atomic>
#include <condition_variable>
#include <thread>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../../include/idb.h"
#include "idb_internal.h"
#include "copy_pool.h"

// ---------------------------------------------------------------------------
// Scratch of the allocation-free runs: one static device array, which every
// GPU holds at its own address for the life of the process.  Host code never
// uses its name as an address (in host code it is only a stand-in): the
// device address is looked up once per device (scratch_on) and kept.
static __device__ __align__(256) uint8_t scratch[IDB_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, uint8_t *> g_scratch_at;  // device -> address of its scratch array

// The device address of `device`'s scratch array; the device must be current.
hipError_t scratch_on(int device, uint8_t **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_at.find(device);
    if (it == g_scratch_at.end()) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return e != hipSuccess ? e : hipErrorInvalidSymbol;
        it = g_scratch_at.emplace(device, (uint8_t *)p).first;
    }
    *out = it->second;
    return hipSuccess;
}
}  // namespace

// ---------------------------------------------------------------------------
// Context and device arena
enum Slot {
    S_SEQ, S_OFF, S_LEN, S_V, S_J, S_RID, S_RSTART,
    S_KEY, S_GSLOT, S_HTAB, S_REP, S_COPIES, S_CNT, S_CORE, S_PAR, S_BMIN,
    S_FLAG, S_UNIQ, S_NUM, S_MEMB, S_EOFF, S_EKEY, S_EROW, S_EKEY2, S_EROW2,
    S_ELEN, S_ECP, S_ECHK, S_ECODES, S_EPK, S_BSTART, S_QBOFF, S_TCNT, S_TOFF, S_TILES,
    S_LAB, S_RANK, S_RFLAG, S_QBCNT, S_BFLAG, S_TLUT_C, S_SMALL, S_WORK, S_TMP,
    S_OUT_CID, S_OUT_CORE, S_OUT_COP, S_OUT_UNQ, S_OUT_CNT, S_USTART, S_EDGES, S_ENT, S_FILL, S_ZB, S_PAIRS,
    S_DIST, S_CNT2, S_EQG, S_RSCNT, S_PE, S_PARE, S_CMINE, S_GCALLS, S_GOFF, S_GLEN, S_GOUT, S_GHDR, S_GWORDS,
    S_ELOW, S_SCAN, S_ROWREC, S_COUNT
};

struct idb_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // Outputs that are final early (copies, uniques, core flags) go back to
    // the host on this stream while the clustering runs on `stream`.
    hipStream_t side = nullptr;
    hipEvent_t side_ev[2] = {};
    std::string err;
    void *ptr[S_COUNT] = {};
    size_t cap[S_COUNT] = {};
    uint64_t *pinned = nullptr;   // pinned host staging: read-backs [0, 32 KiB), uploads [32, 64 KiB)
    // Pinned staging of host-memory inputs and outputs (grown on demand).
    uint8_t *xfer = nullptr;
    size_t xfer_cap = 0;
    hipEvent_t ev[IDB_N_STAGES + 1] = {};
    hipEvent_t kev[4] = {};       // around the count and the fuse kernels
    // autotuner: best count-kernel launch shape per kernel and shape class
    struct Tuned {
        uint32_t wpb, grid_mul, no_xcd, trials;
    };
    std::map<std::array<uint32_t, 6>, Tuned> tuned;
    hipGraphExec_t gexec = nullptr;  // the tail graph, updated in place
    void *scan_ws = nullptr;         // the current call's scan workspace (idb_scan.hip)
    uint32_t scan_gen = 0;           // scans launched on it since it was zeroed
    // Allocation-free runs (idb_dbscan_noalloc): every buffer of the call is
    // carved from this device's scratch array instead of the arena.
    uint8_t *scratch = nullptr;   // device address of the scratch array on this device
    bool noalloc = false;         // the current call takes buffers from scratch
    size_t na_used = 0;           // bytes handed out in this call
    void *na_ptr[S_COUNT] = {};
    size_t na_cap[S_COUNT] = {};
};

namespace {

struct Fail {
    int code;
};

void check(idb_ctx *c, hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        c->err = std::string(what) + ": " + hipGetErrorString(e);
        throw Fail{IDB_ERR_HIP};
    }
}

void fail(idb_ctx *c, int code, const std::string &msg)
{
    c->err = msg;
    throw Fail{code};
}

template <typename T>
T *buf(idb_ctx *c, Slot s, size_t count)
{
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = 256;
    if (c->noalloc) {
        // bump allocation in the scratch array, one region per slot and call
        if (c->na_cap[s] >= bytes) return (T *)c->na_ptr[s];
        const size_t at = (c->na_used + 255) & ~(size_t)255;
        if (at + bytes > IDB_NOALLOC_SCRATCH_BYTES)
            fail(c, IDB_ERR_UNSUPPORTED, "the allocation-free run needs more than its " +
                                             std::to_string(IDB_NOALLOC_SCRATCH_BYTES >> 20) +
                                             " MiB scratch for this input (" + std::to_string(at + bytes) +
                                             " bytes at least)");
        c->na_ptr[s] = c->scratch + at;
        c->na_cap[s] = bytes;
        c->na_used = at + bytes;
        return (T *)c->na_ptr[s];
    }
    if (c->cap[s] < bytes) {
        if (c->ptr[s]) (void)hipFree(c->ptr[s]);
        c->ptr[s] = nullptr;
        c->cap[s] = 0;
        size_t want = bytes + bytes / 4;
        void *p = nullptr;
        if (hipMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            if (hipMalloc(&p, bytes) != hipSuccess) {
                (void)hipGetLastError();
                fail(c, IDB_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
            }
            want = bytes;
        }
        c->ptr[s] = p;
        c->cap[s] = want;
    }
    return (T *)c->ptr[s];
}

void *tmp(idb_ctx *c, size_t bytes) { return buf<uint8_t>(c, S_TMP, bytes); }

// ---------------------------------------------------------------------------
// R5 threshold table.  The reference computes truncf(max(float lx, float ly)
// * eps) on the device and converts to int with the device's saturating rule
// (NaN -> 0); it depends on the pair only through max(lx, ly).
int32_t cvt_rzi_sat(float f)
{
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}

int32_t threshold_of(uint32_t L, const idb_params &p)
{
    if (p.abs_k >= 0) return p.abs_k;
    volatile float prod = (float)L * p.eps;
    return cvt_rzi_sat(std::trunc((float)prod));
}

int16_t clamp16(int64_t v)
{
    if (v < -1) return -1;
    if (v > 32767) return 32767;
    return (int16_t)v;
}

// ---------------------------------------------------------------------------
// Kernels
__device__ __forceinline__ uint64_t mix64(uint64_t h, uint64_t v)
{
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    h *= 0xBF58476D1CE4E5B9ull;
    return h ^ (h >> 31);
}

__device__ __forceinline__ uint64_t fin64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Bytes k .. k+15 of a row of length L as one 16-byte value, zero past the
// row's end.  Every load stays inside the row: a short tail is read as the
// row's last 16 bytes and shifted down, rows under 16 bytes byte by byte.
// Unaligned 16-byte global loads are legal on gfx950.
__device__ __forceinline__ uint4 row16(const uint8_t *s, uint32_t L, uint32_t k)
{
    uint4 v;
    if (k + 16 <= L) {
        __builtin_memcpy(&v, s + k, 16);
        return v;
    }
    const uint32_t r = L - k;  // 1 .. 15 bytes left
    uint64_t lo = 0, hi = 0;
    if (L >= 16) {
        uint64_t t[2];
        __builtin_memcpy(t, s + L - 16, 16);
        const uint32_t sh = 8 * (16 - r);  // 8 .. 120 bits
        if (sh >= 64) {
            lo = t[1] >> (sh - 64);
        } else {
            lo = (t[0] >> sh) | (t[1] << (64 - sh));
            hi = t[1] >> sh;
        }
    } else {
        for (uint32_t b = 0; b < r; ++b) {
            const uint64_t c = s[k + b];
            if (b < 8) lo |= c << (8 * b); else hi |= c << (8 * (b - 8));
        }
    }
    v.x = (uint32_t)lo;
    v.y = (uint32_t)(lo >> 32);
    v.z = (uint32_t)hi;
    v.w = (uint32_t)(hi >> 32);
    return v;
}

__device__ __forceinline__ uint32_t byte_of(const uint4 &v, uint32_t b)
{
    const uint32_t w = b < 8 ? (b < 4 ? v.x : v.y) : (b < 12 ? v.z : v.w);
    return (w >> (8 * (b & 3))) & 0xFFu;
}

__global__ void k_fill_rid(const uint32_t *rstart, uint32_t n_reps, uint32_t n, uint32_t *rid)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        // last repertoire whose first row is <= i (empty repertoires are skipped)
        uint32_t lo = 0, hi = n_reps;
        while (lo + 1 < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (rstart[mid] <= i) lo = mid; else hi = mid;
        }
        rid[i] = lo;
    }
}

// Row scan: content hash (R3 key), letter set, longest row; it also clears
// the duplicate table's P slots.  Letters below 128 are collected in
// per-thread masks and merged once per wave.
__global__ void k_rowscan(const uint8_t *seq, const uint64_t *off, const uint32_t *len, const int32_t *vb, uint32_t vw,
                          const int32_t *jb, uint32_t jw, const uint32_t *rid, uint32_t n, uint64_t hmask,
                          uint64_t *key, uint32_t *small, uint2 *tab, uint32_t P)
{
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < P; s += gridDim.x * blockDim.x)
        tab[s] = make_uint2(IDB_NONE, 0u);
    __shared__ uint32_t sym[8];
    __shared__ uint32_t smax, s_vtop, s_jtop;
    if (threadIdx.x < 8) sym[threadIdx.x] = 0;
    if (threadIdx.x == 0) smax = s_vtop = s_jtop = 0;
    __syncthreads();
    uint64_t m_lo = 0, m_hi = 0;
    uint32_t lmax = 0, vtop = 0, jtop = 0;  // longest row; highest V and J gene + 1
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t L = len[i];
        const uint8_t *s = seq + off[i];
        uint64_t h = mix64(0x1234567ull, ((uint64_t)L << 32) | (rid ? rid[i] : 0u));
        uint64_t acc = 0;
        for (uint32_t k0 = 0; k0 < L; k0 += 16) {
            const uint4 v = row16(s, L, k0);
#pragma unroll
            for (uint32_t b = 0; b < 16; ++b) {
                const uint32_t k = k0 + b;
                if (k < L) {
                    const uint32_t c = byte_of(v, b);
                    if (c < 64) m_lo |= 1ull << c;
                    else if (c < 128) m_hi |= 1ull << (c - 64);
                    else atomicOr(&sym[c >> 5], 1u << (c & 31));
                    acc = (acc << 8) | c;
                    if ((k & 7) == 7) {
                        h = mix64(h, acc);
                        acc = 0;
                    }
                }
            }
        }
        h = mix64(h, acc);
        for (uint32_t w = 0; w < vw; ++w) {
            const uint32_t b = (uint32_t)vb[(size_t)i * vw + w];
            h = mix64(h, b);
            if (b) vtop = max(vtop, w * 32 + 32 - (uint32_t)__builtin_clz(b));
        }
        for (uint32_t w = 0; w < jw; ++w) {
            const uint32_t b = (uint32_t)jb[(size_t)i * jw + w];
            h = mix64(h, 0x100000000ull | b);
            if (b) jtop = max(jtop, w * 32 + 32 - (uint32_t)__builtin_clz(b));
        }
        key[i] = fin64(h) & hmask;
        lmax = L > lmax ? L : lmax;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        m_lo |= __shfl_xor(m_lo, o, 64);
        m_hi |= __shfl_xor(m_hi, o, 64);
        const uint32_t t = __shfl_xor(lmax, o, 64);
        lmax = t > lmax ? t : lmax;
        vtop = max(vtop, (uint32_t)__shfl_xor(vtop, o, 64));
        jtop = max(jtop, (uint32_t)__shfl_xor(jtop, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        if (m_lo) {
            atomicOr(&sym[0], (uint32_t)m_lo);
            atomicOr(&sym[1], (uint32_t)(m_lo >> 32));
        }
        if (m_hi) {
            atomicOr(&sym[2], (uint32_t)m_hi);
            atomicOr(&sym[3], (uint32_t)(m_hi >> 32));
        }
        atomicMax(&smax, lmax);
        if (vtop) atomicMax(&s_vtop, vtop);
        if (jtop) atomicMax(&s_jtop, jtop);
    }
    __syncthreads();
    // eight copies of the words, so that blocks do not queue on one word
    uint32_t *dst = small + IDB_SMALL_ROWSCAN + 9 * (blockIdx.x & 7u);
    if (threadIdx.x < 8 && sym[threadIdx.x]) atomicOr(dst + threadIdx.x, sym[threadIdx.x]);
    if (threadIdx.x == 0) {
        atomicMax(dst + 8, smax);
        uint32_t *gd = small + IDB_SMALL_GENES + 2 * (blockIdx.x & 7u);
        if (s_vtop) atomicMax(gd, s_vtop);
        if (s_jtop) atomicMax(gd + 1, s_jtop);
    }
}

__device__ bool rows_equal(const uint8_t *seq, const uint64_t *off, const uint32_t *len, const int32_t *vb, uint32_t vw,
                           const int32_t *jb, uint32_t jw, const uint32_t *rid, uint32_t a, uint32_t b)
{
    // R3 (PSDBSCAN.hcu:542-565): equal length, bitsets and bytes; batched
    // repertoires never share duplicates.
    if (len[a] != len[b]) return false;
    if (rid && rid[a] != rid[b]) return false;
    for (uint32_t w = 0; w < vw; ++w)
        if (vb[(size_t)a * vw + w] != vb[(size_t)b * vw + w]) return false;
    for (uint32_t w = 0; w < jw; ++w)
        if (jb[(size_t)a * jw + w] != jb[(size_t)b * jw + w]) return false;
    const uint8_t *sa = seq + off[a], *sb = seq + off[b];
    const uint32_t L = len[a];
    for (uint32_t k = 0; k < L; k += 16) {
        const uint4 x = row16(sa, L, k), y = row16(sb, L, k);
        if (x.x != y.x || x.y != y.y || x.z != y.z || x.w != y.w) return false;
    }
    return true;
}

// R3/R4 with an open-addressing table of P >= 2N slots, one slot per group of
// equal rows: x = the smallest row of the group seen so far (IDB_NONE while
// the slot is free), y = the group's size.  A row starts at its hash's slot
// and probes linearly.  One 64-bit exchange claims a free slot with size 1
// (most rows are unique and pay only that); a slot belongs to one group from
// then on, and every member of a group is as good as any other to compare
// with, so later members lower x with atomicMin and count themselves in y
// while others compare against it.  Equal rows share a hash, hence a probe
// sequence, and all stop at the first slot their group claimed.  At most N of
// the P slots fill, so every probe ends at a free slot or at the row's group.
__global__ void k_dedup_insert(const uint8_t *seq, const uint64_t *off, const uint32_t *len, const int32_t *vb,
                               uint32_t vw, const int32_t *jb, uint32_t jw, const uint32_t *rid, const uint64_t *key,
                               uint32_t n, uint2 *tab, uint32_t pmask, uint32_t *gslot)
{
    const unsigned long long FREE = (unsigned long long)IDB_NONE;  // x = IDB_NONE, y = 0
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint64_t h = key[i];
        uint32_t s = (uint32_t)(h ^ (h >> 32)) & pmask;
        for (;;) {
            const unsigned long long old =
                atomicCAS((unsigned long long *)(tab + s), FREE, (1ull << 32) | (unsigned long long)i);
            if (old == FREE) break;  // claimed for a new group
            const uint32_t o = (uint32_t)old;
            if (key[o] == h && rows_equal(seq, off, len, vb, vw, jb, jw, rid, i, o)) {
                atomicMin(&tab[s].x, i);
                atomicAdd(&tab[s].y, 1u);
                break;
            }
            s = (s + 1) & pmask;
        }
        gslot[i] = s;
    }
}

// Representatives, copies and the per-row state of the later stages: flag
// (unique), cnt (own copies), par (own root), bmin (no core neighbour yet),
// pe (no entry yet).
// Also, when the row's V and J words fit in four, its 32-byte record for the
// entry records (rrec): offset, copies, then the V words and the J words,
// so that each entry reads one line for its row instead of four scattered
// ones.
__global__ void k_dedup_resolve(const uint2 *tab, const uint32_t *gslot, uint32_t n, uint32_t *rep, uint32_t *copies,
                                uint8_t *flag, int32_t *cnt, uint32_t *par, uint32_t *bmin, uint32_t *rflag,
                                uint32_t *pe, const uint64_t *off, const int32_t *vb, uint32_t vw, const int32_t *jb,
                                uint32_t jw, uint4 *rrec)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint2 g = tab[gslot[i]];
        const uint32_t cp = g.x == i ? g.y : 0u;
        rep[i] = g.x;
        copies[i] = cp;
        if (rrec) {
            uint32_t w[4] = {0u, 0u, 0u, 0u};
            for (uint32_t k = 0; k < vw; ++k) w[k] = (uint32_t)vb[(size_t)i * vw + k];
            for (uint32_t k = 0; k < jw; ++k) w[vw + k] = (uint32_t)jb[(size_t)i * jw + k];
            const uint64_t o = off[i];
            rrec[2 * (size_t)i] = make_uint4((uint32_t)o, (uint32_t)(o >> 32), cp, 0u);
            rrec[2 * (size_t)i + 1] = make_uint4(w[0], w[1], w[2], w[3]);
        }
        flag[i] = g.x == i;
        cnt[i] = (int32_t)cp;
        par[i] = i;
        bmin[i] = IDB_NONE;
        rflag[i] = 0u;
        pe[i] = IDB_NONE;
    }
}

__device__ __forceinline__ uint32_t popc_words(const int32_t *w, uint32_t n)
{
    uint32_t c = 0;
    for (uint32_t k = 0; k < n; ++k) c += __builtin_popcount((uint32_t)w[k]);
    return c;
}

__global__ void k_membership(const uint32_t *uniq, const uint32_t *num, const int32_t *vb, uint32_t vw,
                             const int32_t *jb, uint32_t jw, uint32_t single, uint32_t cap, uint64_t *memb)
{
    const uint32_t U = num[0];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k <= cap; k += gridDim.x * blockDim.x) {
        if (k >= U) {
            memb[k] = 0u;
            continue;
        }
        const uint32_t u = uniq[k];
        memb[k] = single ? 1u : (uint64_t)popc_words(vb + (size_t)u * vw, vw) * popc_words(jb + (size_t)u * jw, jw);
    }
}

// One entry per (unique row, V gene, J gene); key = bucket << lb | length
// (lb = bits of the longest row).
__global__ void k_emit_entries(const uint32_t *uniq, const uint32_t *num, const uint64_t *eoff, const int32_t *vb,
                               uint32_t vw, const int32_t *jb, uint32_t jw, const uint32_t *len, const uint32_t *rid,
                               uint32_t single, uint32_t gj, uint32_t gvj, uint32_t lb, uint64_t *ekey, uint32_t *erow)
{
    const uint32_t U = num[0];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < U; k += gridDim.x * blockDim.x) {
        const uint32_t u = uniq[k];
        const uint64_t r = rid ? rid[u] : 0u;
        const uint64_t L = len[u];
        uint64_t o = eoff[k];
        if (single) {
            ekey[o] = (r << lb) | L;
            erow[o] = u;
            continue;
        }
        for (uint32_t wv = 0; wv < vw; ++wv) {
            uint32_t mv = (uint32_t)vb[(size_t)u * vw + wv];
            while (mv) {
                const uint32_t v = wv * 32 + __builtin_ctz(mv);
                mv &= mv - 1;
                for (uint32_t wj = 0; wj < jw; ++wj) {
                    uint32_t mj = (uint32_t)jb[(size_t)u * jw + wj];
                    while (mj) {
                        const uint32_t j = wj * 32 + __builtin_ctz(mj);
                        mj &= mj - 1;
                        const uint64_t bucket = r * gvj + (uint64_t)v * gj + j;
                        ekey[o] = (bucket << lb) | L;
                        erow[o] = u;
                        ++o;
                    }
                }
            }
        }
    }
}

// Word k of a gene bitset restricted to the genes below gene g.
__device__ __forceinline__ uint32_t below_gene(uint32_t w, uint32_t k, uint32_t g)
{
    if (k * 32 >= g) return 0u;
    return (k * 32 + 32 > g) ? (w & ((1u << (g - k * 32)) - 1u)) : w;
}

// Residues mod 16 of the set bits of `w` below gene `g` (all bits when g is
// ~0u), folded into a 16-bit mask.
__device__ __forceinline__ uint32_t gene_residues(const int32_t *w, uint32_t nw, uint32_t g)
{
    uint32_t r = 0;
    for (uint32_t k = 0; k < nw && k * 32 < g; ++k) {
        uint32_t m = (uint32_t)w[k];
        if (k * 32 + 32 > g) m &= (1u << (g - k * 32)) - 1u;
        r |= (m | (m >> 16)) & 0xFFFFu;
    }
    return r;
}

// Materialise entries in bucket order: the 32-byte record, the length (for
// the tile builder) and the letters as codes or packed words.
__global__ void k_fill_entries(const uint64_t *ekey, const uint32_t *erow, uint32_t E, const uint8_t *seq,
                               const uint64_t *off, const uint32_t *copies, const int32_t *vb, uint32_t vw,
                               const int32_t *jb, uint32_t jw, uint32_t single, uint32_t gj, uint32_t gvj,
                               const uint8_t *cmap, uint32_t pad, uint32_t S, uint32_t HW, uint32_t bits,
                               const int16_t *tlut_c, const int16_t *tlut_f, uint32_t pres, uint32_t group_len, uint32_t lb,
                               uint16_t *elen, idb_entry *ent, uint32_t *ecodes, uint32_t *epk, uint8_t *bflag,
                               uint32_t *eqg, uint32_t *pe, uint32_t *par_e, uint32_t *cmin_e, uint4 *elow,
                               const uint4 *rrec)
{
    __shared__ uint8_t cm[256];
    __shared__ uint32_t qgc[256 * IDB_QGRAM_WORDS];  // one 64-byte 3-gram count row per thread
    for (uint32_t b = threadIdx.x; b < 256; b += blockDim.x) cm[b] = cmap[b];
    __syncthreads();
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const uint64_t k = ekey[e];
        const uint32_t u = erow[e];
        const uint32_t L = (uint32_t)(k & ((1ull << lb) - 1ull));
        const uint64_t bucket = k >> lb;
        elen[e] = (uint16_t)L;
        // cluster growth in entry space: own root, no component minimum yet,
        // and one entry to stand for the row (any of a row's entries serves:
        // the components, and so the rows' roots, do not depend on which)
        par_e[e] = e;
        cmin_e[e] = IDB_NONE;
        pe[u] = e;
        // a group is a bucket, or in Hamming mode a bucket's run of one length
        bflag[e] = (e == 0) || (group_len ? ekey[e - 1] != k : (ekey[e - 1] >> lb) != bucket);
        idb_entry r;
        r.row = u;
        uint64_t ro;
        uint32_t sv, sj;
        const uint32_t bl = single ? 0u : (uint32_t)(bucket % gvj);
        const uint32_t gv = single ? 0xFFFFFFFFu : bl / gj, gjb = single ? 0xFFFFFFFFu : bl % gj;
        if (rrec) {
            // the row's record: offset, copies, V words then J words
            const uint4 r0 = rrec[2 * (size_t)u], r1 = rrec[2 * (size_t)u + 1];
            ro = (uint64_t)r0.x | ((uint64_t)r0.y << 32);
            r.cp = r0.z;
            const uint32_t w[4] = {r1.x, r1.y, r1.z, r1.w};
            sv = 0u;
            sj = 0u;
            uint32_t lw[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                const bool isv = k < vw, isj = k >= vw && k < vw + jw;
                const uint32_t kk = isv ? k : k - vw;
                const uint32_t m = (isv || isj) ? below_gene(w[k], kk, isv ? gv : gjb) : 0u;
                lw[k] = m;
                const uint32_t f = (m | (m >> 16)) & 0xFFFFu;
                if (isv) sv |= f;
                if (isj) sj |= f;
            }
            if (elow) elow[e] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
        } else {
            ro = off[u];
            r.cp = copies[u];
            sv = gene_residues(vb + (size_t)u * vw, vw, gv);
            sj = gene_residues(jb + (size_t)u * jw, jw, gjb);
            if (elow) {
                // the row's V genes below the bucket's V gene, then its J genes
                // below the bucket's J gene (canonical mode, vw + jw <= 4)
                uint32_t w[4] = {0u, 0u, 0u, 0u};
                for (uint32_t k = 0; k < vw; ++k) w[k] = below_gene((uint32_t)vb[(size_t)u * vw + k], k, gv);
                for (uint32_t k = 0; k < jw; ++k) w[vw + k] = below_gene((uint32_t)jb[(size_t)u * jw + k], k, gjb);
                elow[e] = make_uint4(w[0], w[1], w[2], w[3]);
            }
        }
        const uint32_t chk = single ? 1u : ((sv | sj) != 0u);
        r.len_chk = L | (chk << 16);
        r.t2 = (uint32_t)(uint16_t)tlut_c[L] | ((uint32_t)(uint16_t)tlut_f[L] << 16);
        r.sig = sv | (sj << 16);
        const uint8_t *s = seq + ro;
        uint32_t p1 = 0, p2 = 0;
        if (ecodes) {
            // codes 16 letters at a time, each letter looked up once: letters
            // past the row's end carry the sentinel code `pad`, and the
            // letter-set planes come from the same codes
            uint4 *dst = (uint4 *)(ecodes + (size_t)e * (S / 4));
            const uint32_t padw = pad * 0x01010101u;
            for (uint32_t k0 = 0; k0 < S; k0 += 16) {
                if (k0 >= L) {
                    dst[k0 / 16] = make_uint4(padw, padw, padw, padw);
                    continue;
                }
                const uint4 v = row16(s, L, k0);
                uint32_t w[4];
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q) {
                    uint32_t word = 0;
#pragma unroll
                    for (uint32_t b = 0; b < 4; ++b) {
                        const uint32_t pos = k0 + q * 4 + b;
                        const uint32_t code = pos < L ? (uint32_t)cm[byte_of(v, q * 4 + b)] : pad;
                        word |= code << (8 * b);
                        if (pres && pos < L) {
                            const uint32_t bit = 1u << code;
                            p2 |= p1 & bit;
                            p1 |= bit;
                        }
                    }
                    w[q] = word;
                }
                dst[k0 / 16] = make_uint4(w[0], w[1], w[2], w[3]);
            }
        } else if (pres) {
            for (uint32_t k0 = 0; k0 < L; k0 += 16) {
                const uint4 v = row16(s, L, k0);
#pragma unroll
                for (uint32_t b = 0; b < 16; ++b)
                    if (k0 + b < L) {
                        const uint32_t bit = 1u << cm[byte_of(v, b)];
                        p2 |= p1 & bit;
                        p1 |= bit;
                    }
            }
        }
        r.pres1 = p1;
        r.pres2 = p2;
        r.len = L;
        ent[e] = r;
        if (eqg) {
            // 3-gram profile (codes 0..3): counts of c[i-2] c[i-1] c[i],
            // saturating at 255
            uint32_t *row = qgc + threadIdx.x * IDB_QGRAM_WORDS;
            uint8_t *cnt8 = (uint8_t *)row;
            for (uint32_t w = 0; w < IDB_QGRAM_WORDS; ++w) row[w] = 0u;
            uint32_t hist = 0;
            for (uint32_t k0 = 0; k0 < L; k0 += 16) {
                const uint4 v = row16(s, L, k0);
                for (uint32_t b = 0; b < 16 && k0 + b < L; ++b) {
                    hist = ((hist << 2) | (cm[byte_of(v, b)] & 3u)) & 63u;
                    if (k0 + b >= 2 && cnt8[hist] < 255) ++cnt8[hist];
                }
            }
            uint4 *dst = (uint4 *)(eqg + (size_t)e * IDB_QGRAM_WORDS);
            for (uint32_t w = 0; w < IDB_QGRAM_WORDS / 4; ++w)
                dst[w] = make_uint4(row[4 * w], row[4 * w + 1], row[4 * w + 2], row[4 * w + 3]);
        }
        if (epk && bits == 5) {
            // six 5-bit letters per word (bits 0..29), trailing fields zero
            uint32_t *dst = epk + (size_t)e * HW;
            uint32_t cur = 0, fld = 0, wi = 0;
            for (uint32_t k0 = 0; k0 < L; k0 += 16) {
                const uint4 v = row16(s, L, k0);
                for (uint32_t b = 0; b < 16 && k0 + b < L; ++b) {
                    cur |= (uint32_t)cm[byte_of(v, b)] << (5 * fld);
                    if (++fld == 6) {
                        dst[wi++] = cur;
                        cur = 0;
                        fld = 0;
                    }
                }
            }
            if (fld) dst[wi++] = cur;
            for (; wi < HW; ++wi) dst[wi] = 0u;
        } else if (epk) {
            // 16 letters per step: one 2-bit word or four 8-bit words
            uint32_t *dst = epk + (size_t)e * HW;
            const uint32_t per = 32 / bits;
            for (uint32_t k0 = 0; k0 < HW * per; k0 += 16) {
                const uint4 v = k0 < L ? row16(s, L, k0) : make_uint4(0, 0, 0, 0);
                uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (uint32_t b = 0; b < 16; ++b)
                    if (k0 + b < L) {
                        const uint32_t code = cm[byte_of(v, b)];
                        if (bits == 2) w[0] |= code << (2 * b);
                        else w[b >> 2] |= code << (8 * (b & 3));
                    }
                const uint32_t w0 = k0 / per, nw = bits == 2 ? 1u : 4u;
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q)
                    if (q < nw && w0 + q < HW) dst[w0 + q] = w[q];
            }
        }
    }
}

// Query blocks per bucket over all E + 1 slots (0 past the B buckets), so the
// scan's length is known on the host without reading B back.
__global__ void k_bucket_blocks(const uint32_t *bstart, const uint32_t *nb_dev, uint32_t E, uint32_t *qbcnt)
{
    const uint32_t B = nb_dev[0];
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b <= E; b += gridDim.x * blockDim.x) {
        if (b >= B) {
            qbcnt[b] = 0;
            continue;
        }
        const uint32_t s = bstart[b], e = (b + 1 < B) ? bstart[b + 1] : E;
        qbcnt[b] = (e - s + IDB_TILE - 1) / IDB_TILE;
    }
}

__device__ __forceinline__ uint32_t upper_bound_u32(const uint32_t *a, uint32_t n, uint32_t v)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Index of the last entry in [s, e) with length <= Lmax, or s - 1.
__device__ __forceinline__ int64_t last_len_le(const uint16_t *elen, uint32_t s, uint32_t e, uint32_t Lmax)
{
    uint32_t lo = s, hi = e;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (elen[mid] <= Lmax) lo = mid + 1; else hi = mid;
    }
    return (int64_t)lo - 1;
}

struct QBlock {
    uint32_t b, I, s, e, nblk;
};

__device__ __forceinline__ QBlock qblock_of(uint32_t qb, const uint32_t *qboff, uint32_t B, const uint32_t *bstart,
                                            uint32_t E)
{
    QBlock r;
    r.b = upper_bound_u32(qboff, B, qb) - 1;
    r.I = qb - qboff[r.b];
    r.s = bstart[r.b];
    r.e = (r.b + 1 < B) ? bstart[r.b + 1] : E;
    r.nblk = (r.e - r.s + IDB_TILE - 1) / IDB_TILE;
    return r;
}

__device__ __forceinline__ int64_t last_target_block(const QBlock &q, const uint16_t *elen, const uint16_t *hilen,
                                                    uint32_t banded)
{
    if (!banded) return (int64_t)q.nblk - 1;
    const uint32_t qend = min(q.s + (q.I + 1) * IDB_TILE, q.e);
    const uint32_t lmax = elen[qend - 1];
    const int64_t last = last_len_le(elen, q.s, q.e, hilen[lmax]);
    if (last < (int64_t)q.s) return -1;
    return (last - (int64_t)q.s) / IDB_TILE;
}

// Tiles per query block, as 64-bit values: one bucket of a few million
// entries (single-bucket mode at 20M rows) makes more than 2^32 tiles, and
// their total is scanned in 64 bits so that the host sees it whole.
__global__ void k_tile_counts(const uint32_t *qboff, const uint32_t *nb_dev, const uint32_t *bstart, uint32_t E,
                              const uint16_t *elen, const uint16_t *hilen, uint32_t banded, uint32_t split,
                              uint64_t *tcnt)
{
    const uint32_t B = nb_dev[0];
    const uint32_t QB = qboff[B];
    for (uint32_t qb = blockIdx.x * blockDim.x + threadIdx.x; qb <= E; qb += gridDim.x * blockDim.x) {
        if (qb >= QB) {
            tcnt[qb] = 0;
            continue;
        }
        const QBlock q = qblock_of(qb, qboff, B, bstart, E);
        const int64_t jhi = last_target_block(q, elen, hilen, banded);
        uint64_t nt = jhi >= (int64_t)q.I ? (uint64_t)(jhi - q.I + 1) : 0u;
        // split: off-diagonal tiles of more than 32 queries become two tiles
        if (split && nt > 1 && q.e - (q.s + q.I * IDB_TILE) > IDB_TILE / 2) nt = 2 * nt - 1;
        tcnt[qb] = nt;
    }
}

// Sizes the host needs, gathered on the device so that one read-back brings
// them: sz[0..1] = entries (eoff[U], when there are bucket memberships),
// sz[5] = uniques; later sz[2] = buckets, sz[3] = query blocks, sz[4] and
// sz[6] = tiles (low and high words).
__global__ void k_size_entries(const uint64_t *eoff, const uint32_t *num, uint32_t *sz)
{
    const uint32_t U = num[0];
    const uint64_t e = eoff ? eoff[U] : 0u;
    sz[0] = (uint32_t)e;
    sz[1] = (uint32_t)(e >> 32);
    sz[5] = U;
}

__global__ void k_size_tiles(const uint32_t *num, const uint32_t *qboff, const uint64_t *toff, uint32_t *sz)
{
    const uint32_t B = num[1], QB = qboff[B];
    const uint64_t T = toff[QB];
    sz[2] = B;
    sz[3] = QB;
    sz[4] = (uint32_t)T;
    sz[6] = (uint32_t)(T >> 32);
}

__global__ void k_emit_tiles(const uint32_t *qboff, const uint32_t *nb_dev, const uint32_t *bstart, uint32_t E,
                             const uint16_t *elen, const uint16_t *hilen, uint32_t banded, uint32_t split,
                             const uint64_t *ekey, uint32_t lb, uint32_t gvj, const uint64_t *toff, idb_tile *tiles)
{
    const uint32_t B = nb_dev[0];
    const uint32_t QB = qboff[B];
    for (uint32_t qb = blockIdx.x * blockDim.x + threadIdx.x; qb < QB; qb += gridDim.x * blockDim.x) {
        const QBlock q = qblock_of(qb, qboff, B, bstart, E);
        const int64_t jhi = last_target_block(q, elen, hilen, banded);
        const uint32_t q0 = q.s + q.I * IDB_TILE;
        const uint32_t qn = min((uint32_t)IDB_TILE, q.e - q0);
        // bucket id within its repertoire: the canonical test needs only (v, j)
        const uint32_t bucket = (uint32_t)((ekey[q.s] >> lb) % gvj);
        uint64_t o = toff[qb];
        for (int64_t J = q.I; J <= jhi; ++J) {
            const uint32_t t0 = q.s + (uint32_t)J * IDB_TILE;
            const uint32_t tn = min((uint32_t)IDB_TILE, q.e - t0);
            idb_tile t;
            t.q0 = q0;
            t.t0 = t0;
            t.bucket = bucket;
            if (split && J != q.I && qn > IDB_TILE / 2) {
                // two tiles of 32 and qn - 32 queries against the same targets
                t.shape = (IDB_TILE / 2) | (tn << 8);
                tiles[o++] = t;
                t.q0 = q0 + IDB_TILE / 2;
                t.shape = (qn - IDB_TILE / 2) | (tn << 8);
                tiles[o++] = t;
                continue;
            }
            t.shape = qn | (tn << 8) | ((J == q.I ? 1u : 0u) << 16);
            tiles[o++] = t;
        }
    }
}

// Sum over the block, then one atomic per block: a per-thread atomic on one
// address serialises.
__device__ __forceinline__ void block_add(uint32_t v, uint32_t *dst)
{
    __shared__ uint32_t part[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) part[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t k = 0; k < (blockDim.x + 63) / 64; ++k) t += part[k];
        if (t) atomicAdd(dst, t);
    }
}

// R8 core flags; the entry that stands for a core row (pe) also carries the flag in bit 31.
__global__ void k_core_flags(const int32_t *cnt, uint32_t n, int32_t minpts, int32_t *core, uint32_t *pe,
                             uint32_t *small)
{
    uint32_t local = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int32_t c = cnt[i] >= minpts;
        core[i] = c;
        if (c) {
            const uint32_t e = pe[i];
            if (e != IDB_NONE) pe[i] = e | 0x80000000u;
        }
        local += (uint32_t)c;
    }
    block_add(local, small + IDB_SMALL_CORES + (blockIdx.x & 15u));  // 16 counters
}

__device__ __forceinline__ uint32_t find_root(const uint32_t *par, uint32_t x)
{
    uint32_t p = par[x];
    while (p != x) {
        x = p;
        p = par[x];
    }
    return x;
}

// R9/R10: label root of every unique row (its component root, its border
// core's root, or none) and root flags for R12.
__global__ void k_roots(const uint32_t *uniq, const uint32_t *num, const int32_t *core, const uint32_t *par,
                        const uint32_t *bmin, uint32_t *lab, uint32_t *rflag, uint32_t n)
{
    const uint32_t U = num[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) rflag[n] = 0u;  // the renumbering scan's last slot
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < U; k += gridDim.x * blockDim.x) {
        const uint32_t u = uniq[k];
        uint32_t l = IDB_NONE;
        if (core[u]) {
            l = find_root(par, u);
            if (l == u) rflag[u] = 1u;
        } else if (bmin[u] != IDB_NONE) {
            l = find_root(par, bmin[u]);
        }
        lab[u] = l;
    }
}

// Depth histogram over the row forest (the reference's
// tree_depth_evaluation_kernel, cuda/PSDBSCAN.hcu:154-177) for runs whose
// clusters grew in row space: the fuse pass after a pair-list overflow, or
// no bucket entries at all.  Runs grown from the pair list get theirs from
// the entry forest (k_union_depth) and skip this one (only_if_overflow).
__global__ void k_tree_depth(const uint32_t *rep, const int32_t *core, const uint32_t *par, const uint32_t *bmin,
                             uint32_t n, uint32_t *hist, const uint32_t *n_chunks, uint32_t max_chunks,
                             uint32_t only_if_overflow)
{
    if (only_if_overflow && *n_chunks <= max_chunks) return;
    // a block histogram in LDS, added to the run's histogram once per bin
    __shared__ uint32_t h[51];
    for (uint32_t k = threadIdx.x; k < 51; k += blockDim.x) h[k] = 0u;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t d = 0, x = i;
        if (rep[x] != x) { x = rep[x]; ++d; }
        if (!core[x] && bmin[x] != IDB_NONE) { x = bmin[x]; ++d; }
        if (core[x])
            while (par[x] != x && d < 50) { x = par[x]; ++d; }
        atomicAdd(&h[d > 50 ? 50 : d], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < 51; k += blockDim.x)
        if (h[k]) atomicAdd(hist + k, h[k]);
}

// R11/R12 labels plus the per-row outputs.
__global__ void k_labels(const uint32_t *rep, const uint32_t *lab, const uint32_t *rank, const uint32_t *rid,
                         const uint32_t *rstart, const int32_t *core, const uint32_t *copies, const int32_t *cnt,
                         uint32_t n, uint32_t *cid, int32_t *core_out, uint32_t *cop_out, uint32_t *cnt_out,
                         uint32_t *unq_out, uint32_t *small)
{
    uint32_t noise = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t l = lab[rep[i]];
        uint32_t c = 0;
        if (l != IDB_NONE) {
            const uint32_t base = rid ? rank[rstart[rid[i]]] : 0u;
            c = rank[l] - base + 1u;
        }
        noise += c == 0;
        cid[i] = c;
        if (core_out) core_out[i] = core[i];  // (outputs already sent back early are skipped)
        if (cop_out) cop_out[i] = copies[i];
        if (cnt_out) cnt_out[i] = (uint32_t)cnt[i];
        if (unq_out) unq_out[i] = 0u;  // the uniques' positions are written next
    }
    block_add(noise, small + IDB_SMALL_NOISE + (blockIdx.x & 15u));  // 16 counters
}

// First unique position of every repertoire (binary search over uniques).
__global__ void k_ustart(const uint32_t *uniq, const uint32_t *num, const uint32_t *rstart, uint32_t n_reps,
                         uint32_t *ustart)
{
    const uint32_t U = num[0];
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r <= n_reps; r += gridDim.x * blockDim.x) {
        const uint32_t v = rstart[r];
        uint32_t lo = 0, hi = U;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (uniq[mid] < v) lo = mid + 1; else hi = mid;
        }
        ustart[r] = lo;
    }
}

__global__ void k_uniques_out(const uint32_t *uniq, const uint32_t *num, const uint32_t *rid, const uint32_t *rstart,
                              const uint32_t *ustart, uint32_t *out)
{
    const uint32_t U = num[0];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < U; k += gridDim.x * blockDim.x) {
        const uint32_t u = uniq[k];
        if (!rid) {
            out[k] = u;
        } else {
            const uint32_t r = rid[u];
            out[rstart[r] + (k - ustart[r])] = u - rstart[r];
        }
    }
}

// Grid-stride kernels run at most 8 blocks per CU: enough to fill the chip,
// few enough that per-block counter atomics stay cheap.
uint32_t blocks_for(uint64_t n, uint32_t bs = 256)
{
    uint64_t b = (n + bs - 1) / bs;
    if (b > 2048u) b = 2048u;
    return b ? (uint32_t)b : 1u;
}

// ---------------------------------------------------------------------------
// rocPRIM wrappers (temporary storage from the arena)
void sort_pairs_u64(idb_ctx *c, uint64_t *kin, uint64_t *kout, uint32_t *vin, uint32_t *vout, size_t n, unsigned end_bit)
{
    size_t bytes = 0;
    check(c, rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, 0u, end_bit, c->stream), "radix sort size");
    void *t = tmp(c, bytes);
    check(c, rocprim::radix_sort_pairs(t, bytes, kin, kout, vin, vout, n, 0u, end_bit, c->stream), "radix sort");
}

template <typename T>
void exclusive_scan_total(idb_ctx *c, const T *in, T *out, size_t n_plus_1)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
    if (sizeof(T) == 4)
        check(c, idb_scan_exclusive_u32((const uint32_t *)in, (uint32_t *)out, n_plus_1, c->scan_ws, ++c->scan_gen,
                                        c->stream), "scan");
    else
        check(c, idb_scan_exclusive_u64((const uint64_t *)in, (uint64_t *)out, n_plus_1, c->scan_ws, ++c->scan_gen,
                                        c->stream), "scan");
}

// Indices of the set flags, in order, and their number (*count, device).
void select_flagged(idb_ctx *c, size_t n, const uint8_t *flags, uint32_t *out, uint32_t *count)
{
    check(c, idb_select_flagged(flags, out, count, n, c->scan_ws, ++c->scan_gen, c->stream), "select");
}

// The scan workspace of a call, sized for scans of up to max_n values and
// cleared when it is new (allocation-free calls carve it afresh each call).
void *scan_workspace(idb_ctx *c, uint64_t max_n)
{
    const size_t bytes = idb_scan_workspace_bytes(max_n);
    const bool fresh = c->noalloc || c->cap[S_SCAN] < bytes;
    uint8_t *w = buf<uint8_t>(c, S_SCAN, bytes);
    if (fresh || c->scan_gen > 0x3FFFFFFFu) {
        check(c, hipMemsetAsync(w, 0, bytes, c->stream), "memset");
        c->scan_gen = 0;
    }
    return w;
}

// Small read-backs go through the context's pinned staging area (a direct
// copy instead of a staged one into pageable memory).
void readback(idb_ctx *c, void *dst, const void *src, size_t bytes)
{
    const bool staged = bytes <= IDB_STAGE_BYTES;
    check(c, hipMemcpyAsync(staged ? (void *)c->pinned : dst, src, bytes, hipMemcpyDeviceToHost, c->stream), "readback");
    check(c, hipStreamSynchronize(c->stream), "readback sync");
    if (staged) std::memcpy(dst, c->pinned, bytes);
}


// Pinned staging of at least `bytes` (at most IDB_XFER_MAX_BYTES; larger
// transfers are copied from pageable memory directly, where the clustering
// outweighs them).  Allocation-free calls use what an earlier call made.
bool xfer_ready(idb_ctx *c, size_t bytes)
{
    if (c->xfer_cap >= bytes) return true;
    if (c->noalloc || bytes > IDB_XFER_MAX_BYTES) return false;
    check(c, hipStreamSynchronize(c->stream), "staging sync");
    if (c->xfer) (void)hipHostFree(c->xfer);
    c->xfer = nullptr;
    c->xfer_cap = 0;
    const size_t want = std::min<size_t>(IDB_XFER_MAX_BYTES, std::max<size_t>(bytes + bytes / 2, 1u << 20));
    if (hipHostMalloc((void **)&c->xfer, want, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        c->xfer = nullptr;
        return false;
    }
    c->xfer_cap = want;
    return true;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

template <typename T>
const T *stage_in(idb_ctx *c, Slot s, const T *src, size_t count, uint32_t mem)
{
    if (mem == IDB_MEM_DEVICE || count == 0) return src;
    T *d = buf<T>(c, s, count);
    check(c, hipMemcpyAsync(d, src, count * sizeof(T), hipMemcpyHostToDevice, c->stream), "input copy");
    return d;
}

// The five input arrays of a host-memory dataset in one device buffer,
// filled by one pinned transfer (or each copied directly when the staging
// is not available).
struct Inputs {
    const uint8_t *seq;
    const uint64_t *off;
    const uint32_t *len;
    const int32_t *vb, *jb;
};

Inputs stage_inputs(idb_ctx *c, const idb_dataset *d, uint32_t N, uint32_t vw, uint32_t jw)
{
    if (d->mem == IDB_MEM_HOST) {
        const size_t by[5] = {d->seq_bytes, (size_t)N * 8, (size_t)N * 4, (size_t)N * vw * 4, (size_t)N * jw * 4};
        const void *src[5] = {d->seq, d->off, d->len, d->vbits, d->jbits};
        size_t at[5], total = 0;
        for (int k = 0; k < 5; ++k) {
            at[k] = total;
            total += align256(by[k]);
        }
        if (total && xfer_ready(c, total)) {
            uint8_t *dev = buf<uint8_t>(c, S_SEQ, total);
            // in four parts above 2 MiB: the transfer of one part overlaps
            // the staging of the next, and each part goes out as soon as it
            // is staged
            const size_t parts = total > (2u << 20) ? 4 : 1, step = align256((total + parts - 1) / parts);
            CopyPool::Job jobs[4 * 5];
            uint32_t part_of[4 * 5];
            size_t nj = 0;
            uint32_t np = 0;
            for (size_t a = 0; a < total; a += step, ++np) {
                const size_t b = std::min(total, a + step);
                for (int k = 0; k < 5; ++k) {
                    const size_t lo = std::max(a, at[k]), hi = std::min(b, at[k] + by[k]);
                    if (lo < hi) {
                        jobs[nj] = {c->xfer + lo, (const uint8_t *)src[k] + (lo - at[k]), hi - lo};
                        part_of[nj++] = np;
                    }
                }
            }
            CopyPool::get().run_parts(jobs, part_of, nj, np, [&](uint32_t q) {
                const size_t a = q * step, b = std::min(total, a + step);
                check(c, hipMemcpyAsync(dev + a, c->xfer + a, b - a, hipMemcpyHostToDevice, c->stream), "input copy");
            });
            return {by[0] ? dev + at[0] : nullptr, (const uint64_t *)(dev + at[1]), (const uint32_t *)(dev + at[2]),
                    by[3] ? (const int32_t *)(dev + at[3]) : nullptr, by[4] ? (const int32_t *)(dev + at[4]) : nullptr};
        }
    }
    return {stage_in<uint8_t>(c, S_SEQ, d->seq, d->seq_bytes, d->mem),
            stage_in<uint64_t>(c, S_OFF, d->off, N, d->mem),
            stage_in<uint32_t>(c, S_LEN, d->len, N, d->mem),
            stage_in<int32_t>(c, S_V, d->vbits, (size_t)N * vw, d->mem),
            stage_in<int32_t>(c, S_J, d->jbits, (size_t)N * jw, d->mem)};
}

unsigned bits_for(uint64_t v)
{
    unsigned b = 0;
    while (b < 64 && (v >> b)) ++b;
    return b;
}

void record(idb_ctx *c, int k) { check(c, hipEventRecord(c->ev[k], c->stream), "event record"); }

// roctx ranges named after the reference's stage timers
// (cuda/PSDBSCAN.hcu:1301-1354): one outer range for the whole run and one
// for the stage being enqueued; the destructor closes what is open, also
// when a run fails.
struct StageRanges {
    int depth = 0;
    StageRanges() { depth += roctxRangePushA("Clustering core process") >= 0; }
    void stage(const char *name)
    {
        if (depth > 1) {
            roctxRangePop();
            --depth;
        }
        depth += roctxRangePushA(name) >= 0;
    }
    ~StageRanges()
    {
        while (depth-- > 0) roctxRangePop();
    }
};

// Autotuner of the count kernel's launch shape: waves per block (1, 2, 4 as
// LDS allows), grid of 1, 2, 4 or 8 times the resident blocks, XCD-grouped or
// plain block order.  Each candidate runs the real count pass on the real
// tiles twice (counts restored from a snapshot in between) and the fastest by
// HIP events is kept for this kernel and shape class (kernel, words, code
// width, filter, alphabet, log2 tiles) for the context's lifetime.
void autotune(idb_ctx *c, idb_pair_args &a, uint32_t N, uint32_t W, bool ham, bool learned, uint32_t *small,
              unsigned long long *work)
{
    uint32_t lt = 0;
    while (lt < 31 && (1u << (lt + 1)) <= a.n_tiles) ++lt;
    const std::array<uint32_t, 6> key{learned ? 2u : (ham ? 1u : 0u), ham ? a.HW : W, a.bits, a.band_filter, a.K, lt};
    auto it = c->tuned.find(key);
    if (it == c->tuned.end()) {
        hipStream_t s = c->stream;
        int32_t *snap = buf<int32_t>(c, S_CNT2, N);
        check(c, hipMemcpyAsync(snap, a.cnt, sizeof(int32_t) * N, hipMemcpyDeviceToDevice, s), "tune snapshot");
        idb_ctx::Tuned best{0, 0, 0, 0};
        float best_ms = 1e30f;
        uint32_t trials = 0;
        for (uint32_t wpb : {1u, 2u, 4u})
            for (uint32_t mul : {1u, 2u, 4u, 8u})
                for (uint32_t nx : {0u, 1u}) {
                    a.tune_wpb = wpb;
                    a.tune_grid_mul = mul;
                    a.no_xcd_order = nx;
                    check(c, hipMemcpyAsync(a.cnt, snap, sizeof(int32_t) * N, hipMemcpyDeviceToDevice, s), "tune reset");
                    check(c, hipMemsetAsync(small + 12, 0, sizeof(uint32_t), s), "tune reset");
                    check(c, hipMemsetAsync(work, 0, IDB_WORK_U64 * IDB_WORK_COPIES * sizeof(unsigned long long), s), "tune reset");
                    // best of two runs per shape
                    float ms = 1e30f;
                    for (int rep = 0; rep < 2; ++rep) {
                        if (rep) {
                            check(c, hipMemcpyAsync(a.cnt, snap, sizeof(int32_t) * N, hipMemcpyDeviceToDevice, s),
                                  "tune reset");
                            check(c, hipMemsetAsync(small + 12, 0, sizeof(uint32_t), s), "tune reset");
                            check(c, hipMemsetAsync(work, 0, IDB_WORK_U64 * IDB_WORK_COPIES * sizeof(unsigned long long), s), "tune reset");
                        }
                        check(c, hipEventRecord(c->kev[0], s), "event record");
                        const hipError_t e = learned ? idb_launch_count_learned(a, s)
                                                     : (ham ? idb_launch_count_hamming(a, s)
                                                            : idb_launch_count_myers(a, W, s));
                        check(c, e, "autotune launch");
                        check(c, hipEventRecord(c->kev[1], s), "event record");
                        check(c, hipEventSynchronize(c->kev[1]), "event sync");
                        float t = 0.f;
                        check(c, hipEventElapsedTime(&t, c->kev[0], c->kev[1]), "event time");
                        ms = std::min(ms, t);
                    }
                    ++trials;
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = {wpb, mul, nx, 0};
                    }
                }
        best.trials = trials;
        check(c, hipMemcpyAsync(a.cnt, snap, sizeof(int32_t) * N, hipMemcpyDeviceToDevice, s), "tune reset");
        check(c, hipMemsetAsync(small + 12, 0, sizeof(uint32_t), s), "tune reset");
        check(c, hipMemsetAsync(work, 0, IDB_WORK_U64 * IDB_WORK_COPIES * sizeof(unsigned long long), s), "tune reset");
        it = c->tuned.emplace(key, best).first;
    }
    a.tune_wpb = it->second.wpb;
    a.tune_grid_mul = it->second.grid_mul;
    a.no_xcd_order = it->second.no_xcd;
}

// Options beyond the reference's run: the learned candidate filter, and a
// copy of the neighbour pairs for the caller.
struct Extra {
    idb_encoder *enc = nullptr;
    float min_sim = 0.f;
    uint32_t *pairs_out = nullptr;
    uint64_t pairs_cap = 0;
    uint64_t *n_pairs_out = nullptr;
};

// ---------------------------------------------------------------------------
// The pipeline.  rstart (host) has n_reps + 1 row offsets; n_reps == 0 means a
// single repertoire without repertoire ids.
void run(idb_ctx *c, const idb_dataset *d, const uint32_t *rstart_h, uint32_t n_reps, const idb_params *p,
         idb_outputs *o, uint32_t *n_uniques_h, idb_stats *st, const Extra *x = nullptr)
{
    const bool learned = x && x->enc;
    const bool force_large = (p->flags & IDB_FLAG_FORCE_LARGE_PATHS) != 0;
    float encoder_ms = 0.f;
    const uint32_t N = d->n_rows;
    const uint32_t vw = d->v_width, jw = d->j_width;
    if (o->mem != IDB_MEM_HOST && o->mem != IDB_MEM_DEVICE) fail(c, IDB_ERR_ARG, "outputs.mem must be host or device");
    if (d->mem != IDB_MEM_HOST && d->mem != IDB_MEM_DEVICE) fail(c, IDB_ERR_ARG, "dataset.mem must be host or device");
    if (!o->clone_id || !o->core || !o->copies || !o->uniques) fail(c, IDB_ERR_ARG, "missing output array");
    if (N && (!d->off || !d->len || (d->seq_bytes && !d->seq))) fail(c, IDB_ERR_ARG, "missing input array");
    if (N && ((vw && !d->vbits) || (jw && !d->jbits))) fail(c, IDB_ERR_ARG, "missing bitset array");
    if (vw > 64 || jw > 64) fail(c, IDB_ERR_UNSUPPORTED, "gene bitsets wider than 64 words (2048 genes)");
    // the duplicate table holds at least 2N slots under 32-bit indexes
    if (N >= (1u << 31)) fail(c, IDB_ERR_UNSUPPORTED, "2^31 rows or more in one call");
    if (p->debug_hash_bits > 64) fail(c, IDB_ERR_ARG, "debug_hash_bits must be 0..64");
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->n_rows = N;
    }
    if (N == 0) {
        if (n_uniques_h)
            for (uint32_t r = 0; r < n_reps; ++r) n_uniques_h[r] = 0;
        return;
    }
    hipStream_t s = c->stream;
    record(c, 0);
    StageRanges rng;
    // scans run over at most N + 1 rows or E + 1 entries, E <= 4N + 4096
    c->scan_ws = scan_workspace(c, 4ull * N + 4098);
    rng.stage("Identify Duplicates");

    // inputs
    const Inputs in = stage_inputs(c, d, N, vw, jw);
    const uint8_t *seq = in.seq;
    const uint64_t *off = in.off;
    const uint32_t *len = in.len;
    const int32_t *vb = in.vb, *jb = in.jb;
    if (!seq) seq = buf<uint8_t>(c, S_SEQ, 4);
    if (!vb) vb = buf<int32_t>(c, S_V, 1);
    if (!jb) jb = buf<int32_t>(c, S_J, 1);
    uint32_t *rid = nullptr, *rstart = nullptr;
    if (n_reps > 0) {
        rstart = buf<uint32_t>(c, S_RSTART, n_reps + 1);
        check(c, hipMemcpyAsync(rstart, rstart_h, sizeof(uint32_t) * (n_reps + 1), hipMemcpyHostToDevice, s), "rstart");
        rid = buf<uint32_t>(c, S_RID, N);
        hipLaunchKernelGGL(k_fill_rid, dim3(blocks_for(N)), dim3(256), 0, s, rstart, n_reps, N, rid);
    }

    // --- Identify Duplicates
    uint32_t *small = buf<uint32_t>(c, S_SMALL, IDB_SMALL_WORDS);
    check(c, hipMemsetAsync(small, 0, IDB_SMALL_WORDS * sizeof(uint32_t), s), "memset");
    uint64_t *key = buf<uint64_t>(c, S_KEY, N);
    // 32 key bits are plenty to group duplicates (colliding rows are told apart
    // by the byte comparison)
    const unsigned hbits = p->debug_hash_bits == 0 ? 32u : std::min(p->debug_hash_bits, 64u);
    const uint64_t hmask = hbits >= 64 ? ~0ull : ((1ull << hbits) - 1ull);
    uint32_t P = 256;
    while (P < 2ull * N) P <<= 1;
    uint2 *htab = buf<uint2>(c, S_HTAB, P);
    uint32_t *gslot = buf<uint32_t>(c, S_GSLOT, N);
    hipLaunchKernelGGL(k_rowscan, dim3(blocks_for(N)), dim3(256), 0, s, seq, off, len, vb, vw, jb, jw, rid, N, hmask,
                       key, small, htab, P);
    hipLaunchKernelGGL(k_dedup_insert, dim3(blocks_for(N)), dim3(256), 0, s, seq, off, len, vb, vw, jb, jw, rid, key, N,
                       htab, P - 1, gslot);
    uint32_t *rep = buf<uint32_t>(c, S_REP, N);
    uint32_t *copies = buf<uint32_t>(c, S_COPIES, N);
    uint8_t *flag = buf<uint8_t>(c, S_FLAG, N);
    int32_t *cnt = buf<int32_t>(c, S_CNT, N);
    uint32_t *par = buf<uint32_t>(c, S_PAR, N), *bmin = buf<uint32_t>(c, S_BMIN, N);
    uint32_t *rflag = buf<uint32_t>(c, S_RFLAG, N + 1);
    uint32_t *rank = buf<uint32_t>(c, S_RANK, N + 1);
    uint32_t *pe = buf<uint32_t>(c, S_PE, N);
    // row records for the entry records, when a row's gene words fit in four
    uint4 *rrec = (vw + jw <= 4) ? buf<uint4>(c, S_ROWREC, 2 * (size_t)N) : nullptr;
    hipLaunchKernelGGL(k_dedup_resolve, dim3(blocks_for(N)), dim3(256), 0, s, htab, gslot, N, rep, copies, flag, cnt, par,
                       bmin, rflag, pe, off, vb, vw, jb, jw, rrec);
    record(c, 1);
    rng.stage("Determine copies");

    // --- Determine copies: uniques
    uint32_t *uniq = buf<uint32_t>(c, S_UNIQ, N);
    uint32_t *num = buf<uint32_t>(c, S_NUM, 4);
    select_flagged(c, N, flag, uniq, num);
    record(c, 2);
    rng.stage("Identify core points");

    // --- Identify core points: buckets, tiles, count pass
    // Memberships are scanned over N + 1 slots (0 past the U uniques).  The
    // entry count (sz[0..1]), U (sz[5]) and the row scan's letter set and
    // longest row (its 8 copies) come back in one read-back.
    uint64_t *memb = buf<uint64_t>(c, S_MEMB, (size_t)N + 1);
    uint64_t *eoff = buf<uint64_t>(c, S_EOFF, (size_t)N + 1);
    uint32_t *sz = small + 100;
    uint32_t single = 0;
    uint32_t E = 0, U = 0;
    if (vw > 0 && jw > 0) {
        hipLaunchKernelGGL(k_membership, dim3(blocks_for((uint64_t)N + 1)), dim3(256), 0, s, uniq, num, vb, vw, jb, jw,
                           0u, N, memb);
        exclusive_scan_total<uint64_t>(c, memb, eoff, (size_t)N + 1);
    }
    hipLaunchKernelGGL(k_size_entries, dim3(1), dim3(64), 0, s, (vw > 0 && jw > 0) ? eoff : nullptr, num, sz);
    uint32_t hs[IDB_SMALL_GENES + 16];
    readback(c, hs, small, sizeof(hs));
    uint32_t lw[9] = {};  // letter-set words and the longest row, over the row scan's 8 copies
    for (uint32_t cpy = 0; cpy < 8; ++cpy)
        for (uint32_t t = 0; t < 9; ++t) {
            const uint32_t v = hs[IDB_SMALL_ROWSCAN + 9 * cpy + t];
            lw[t] = t < 8 ? (lw[t] | v) : std::max(lw[t], v);
        }
    const uint32_t maxlen = lw[8];
    // Buckets are (V gene, J gene) over the gene positions that occur: bucket
    // = v * gj + j with j < gj, so the entry key holds no more bits than the
    // genes in use need (on 200k protein rows 9 bucket bits instead of 11, and
    // one digit pass less in the entry sort).
    uint32_t vtop = 0, jtop = 0;
    for (uint32_t cpy = 0; cpy < 8; ++cpy) {
        vtop = std::max(vtop, hs[IDB_SMALL_GENES + 2 * cpy]);
        jtop = std::max(jtop, hs[IDB_SMALL_GENES + 2 * cpy + 1]);
    }
    const uint32_t gj = std::max(jtop, 1u);
    const uint64_t gvj64 = (uint64_t)std::max(vtop, 1u) * gj;
    uint8_t cmap_h[256];
    uint32_t K = 0;
    for (uint32_t b = 0; b < 256; ++b) {
        cmap_h[b] = (uint8_t)K;
        if (lw[b >> 5] >> (b & 31) & 1u) ++K;
    }
    if (maxlen > 2048) fail(c, IDB_ERR_UNSUPPORTED, "sequences longer than 2048 letters are not supported");
    const bool ham = p->use_hamming != 0;
    uint32_t W = 0, S = 0, HW = 0, bits = 0;
    if (ham) {
        // letter codes 0..K-1 in 2-, 5- or 8-bit fields: 16, 6 or 4 per word
        bits = K <= 4 ? 2u : (K <= 32 ? 5u : 8u);
        const uint32_t per = 32 / bits;
        HW = std::max(1u, (maxlen + per - 1) / per);
        if (HW > IDB_MAX_HW)
            fail(c, IDB_ERR_UNSUPPORTED, "Hamming rows longer than " + std::to_string(IDB_MAX_HW * per) +
                                             " letters are not supported for this alphabet");
    } else {
        W = std::max(1u, (maxlen + 31) / 32);
        if (W > IDB_MAX_W) fail(c, IDB_ERR_UNSUPPORTED, "Levenshtein rows longer than 256 letters are not supported");
        if (idb_myers_lds_per_wave(std::max(std::min(K + 1, 256u), 1u), W) +
                (K <= 4 ? idb_qgram_lds_per_wave() : 0) > idb_lds_limit())
            fail(c, IDB_ERR_UNSUPPORTED, "alphabet of " + std::to_string(K) + " letters with rows of " +
                                             std::to_string(maxlen) + " letters exceeds the LDS budget");
        S = std::max(16u, (maxlen + 15) / 16 * 16);
    }
    if (learned && (ham || maxlen > idb_encoder_positions(x->enc) || W > 5))
        fail(c, IDB_ERR_UNSUPPORTED,
             "the learned filter needs Levenshtein mode and rows no longer than the encoder's positions (" +
                 std::to_string(idb_encoder_positions(x->enc)) + ")");
    if (learned && idb_encoder_device(x->enc) != c->device)
        fail(c, IDB_ERR_ARG, "the encoder lives on another device");
    {
        U = hs[105];
        uint64_t E64 = (uint64_t)hs[100] | ((uint64_t)hs[101] << 32);
        // With no V or no J genes the gate of R6 fails for every pair: no entries.
        // Rows with many genes are copied into many buckets; past 4x the
        // uniques one bucket per repertoire with an explicit gate is cheaper.
        if (vw > 0 && jw > 0 && (E64 > 4ull * U + 4096 || gvj64 * std::max(n_reps, 1u) >= (1ull << 47))) {
            single = 1;
            hipLaunchKernelGGL(k_membership, dim3(blocks_for((uint64_t)N + 1)), dim3(256), 0, s, uniq, num, vb, vw, jb,
                               jw, 1u, N, memb);
            exclusive_scan_total<uint64_t>(c, memb, eoff, (size_t)N + 1);
            E64 = U;
        }
        E = (uint32_t)E64;
    }
    const uint32_t gvj = single ? 1u : (uint32_t)std::min<uint64_t>(gvj64, 0xFFFFFFFFull);
    // threshold tables: count pass, fuse pass, band end
    std::array<int16_t, 2049> tc, tf, hl;  // maxlen <= 2048
    for (uint32_t L = 0; L <= maxlen; ++L) {
        const int64_t t = threshold_of(L, *p);
        if (ham) {
            tc[L] = clamp16(t < 0 ? 0 : t);   // count pass compares only after a mismatch
            tf[L] = clamp16(t);               // fuse pass compares after every letter
        } else {
            tc[L] = tf[L] = clamp16(t);
        }
    }
    bool banded = true;
    if (!ham)
        for (uint32_t L = 1; L <= maxlen; ++L)
            if ((int64_t)L - tc[L] < (int64_t)(L - 1) - tc[L - 1]) banded = false;
    for (uint32_t l = 0; l <= maxlen; ++l) {
        uint32_t h = l;
        if (!ham) {
            h = 0;
            for (uint32_t L = 0; L <= maxlen; ++L)
                if ((int64_t)L - tc[L] <= (int64_t)l) h = L;
        }
        hl[l] = (int16_t)std::min<uint32_t>(h, 32767u);
    }
    // the three length tables and the letter map in one upload from pinned
    // memory: [count t | fuse t | band end] as int16 per length, then 256 codes
    const size_t tl = sizeof(int16_t) * (maxlen + 1), lut_bytes = 3 * tl + 256;
    uint8_t *lut = buf<uint8_t>(c, S_TLUT_C, lut_bytes);
    int16_t *tlut_c = (int16_t *)lut, *tlut_f = (int16_t *)(lut + tl);
    uint16_t *hilen = (uint16_t *)(lut + 2 * tl);
    uint8_t *cmap = lut + 3 * tl;
    {
        uint8_t *up = (uint8_t *)c->pinned + IDB_STAGE_BYTES;  // maxlen <= 2048: 12.5 KiB at most
        std::memcpy(up, tc.data(), tl);
        std::memcpy(up + tl, tf.data(), tl);
        std::memcpy(up + 2 * tl, hl.data(), tl);
        std::memcpy(up + 3 * tl, cmap_h, 256);
        check(c, hipMemcpyAsync(lut, up, lut_bytes, hipMemcpyHostToDevice, s), "lut");
    }

    unsigned long long *work = buf<unsigned long long>(c, S_WORK, IDB_WORK_U64 * IDB_WORK_COPIES);
    check(c, hipMemsetAsync(work, 0, IDB_WORK_U64 * IDB_WORK_COPIES * sizeof(unsigned long long), s), "memset");
    unsigned long long edge_cap = 0;
    uint32_t B = 0, T = 0;
    idb_pair_args a{};
    idb_union_args ua{};
    if (E > 0) {
        uint64_t *ekey = buf<uint64_t>(c, S_EKEY, E), *ekey2 = buf<uint64_t>(c, S_EKEY2, E);
        uint32_t *erow = buf<uint32_t>(c, S_EROW, E), *erow2 = buf<uint32_t>(c, S_EROW2, E);
        const uint32_t lb = std::max(1u, bits_for(maxlen));
        hipLaunchKernelGGL(k_emit_entries, dim3(blocks_for(U)), dim3(256), 0, s, uniq, num, eoff, vb, vw, jb, jw, len, rid,
                           single, gj, gvj, lb, ekey, erow);
        const uint64_t max_bucket = single ? std::max(n_reps, 1u) : gvj64 * std::max(n_reps, 1u);
        const unsigned ebits = lb + bits_for(max_bucket);
        bool sorted = false, in_first = false;
        if (!force_large && E <= IDB_SORT_SMALL_MAX && ebits <= 64) {
            uint32_t *gc = buf<uint32_t>(c, S_RSCNT, (size_t)256 * idb_sort_small_blocks(E));
            const hipError_t e = idb_sort_small(ekey, ekey2, erow, erow2, E, ebits, gc, s, &in_first);
            if (e == hipSuccess) sorted = true;
            else if (e == hipErrorInvalidValue) (void)hipGetLastError();
            else check(c, e, "entry sort");
        }
        if (!sorted) {
            sort_pairs_u64(c, ekey, ekey2, erow, erow2, E, ebits);
            in_first = false;
        }
        // the sorted entry keys and rows
        const uint64_t *skey = in_first ? ekey : ekey2;
        const uint32_t *srow = in_first ? erow : erow2;
        uint16_t *elen = buf<uint16_t>(c, S_ELEN, E);
        idb_entry *ent = buf<idb_entry>(c, S_ENT, E);
        uint8_t *bflag = buf<uint8_t>(c, S_BFLAG, E);
        uint32_t *ecodes = ham ? nullptr : buf<uint32_t>(c, S_ECODES, (size_t)E * (S / 4));
        uint32_t *epk = ham ? buf<uint32_t>(c, S_EPK, (size_t)E * HW) : nullptr;
        // Levenshtein codes are padded with a sentinel letter K whose Peq row
        // stays zero, so a Peq read for a letter past a row's end stays in
        // range (with all 256 byte values present code 0 is used; only the
        // one-word count kernel reads letters past a row's end, and it runs
        // for alphabets of at most 31 letters)
        const uint32_t pad = (!ham && K < 256) ? K : 0u;
        // The absent-letter filter pays off where random letter matches are
        // rare (protein alphabets); with 4 letters nearly every pair passes.
        const bool filt = !ham && W == 1 && K >= 8 && K <= 31;
        // Nucleotide rows: the 3-gram profile bound (learned runs apply it to their proposals).
        const bool qgram = !ham && K <= 4 && maxlen >= 16;
        uint32_t *eqg = qgram ? buf<uint32_t>(c, S_EQG, (size_t)E * IDB_QGRAM_WORDS) : nullptr;
        uint32_t *par_e = buf<uint32_t>(c, S_PARE, E), *cmin_e = buf<uint32_t>(c, S_CMINE, E);
        // The count passes test the canonical bucket on the entries' own
        // genes below the bucket's (V words, then J words), when they fit
        // four words.
        uint4 *elow = (!single && !learned && vw + jw <= 4) ? buf<uint4>(c, S_ELOW, E) : nullptr;
        hipLaunchKernelGGL(k_fill_entries, dim3(blocks_for(E)), dim3(256), 0, s, skey, srow, E, seq, off, copies, vb, vw,
                           jb, jw, single, gj, gvj, cmap, pad, S, HW, bits, tlut_c, tlut_f, filt ? 1u : 0u,
                           ham ? 1u : 0u, lb, elen, ent, ecodes, epk, bflag, eqg, pe, par_e, cmin_e, elow, rrec);
        ua.pe = pe;
        ua.par_e = par_e;
        ua.cmin_e = cmin_e;
        ua.n_entries = E;
        ua.n_rows = N;
        ua.par = par;
        ua.bmin = bmin;
        ua.rep = rep;
        ua.depth_hist = p->tree_stats ? small + 16 : nullptr;
        uint32_t *bstart = buf<uint32_t>(c, S_BSTART, E);
        select_flagged(c, E, bflag, bstart, num + 1);
        // Buckets B <= E and query blocks QB <= E stay on the device: the
        // block and tile counts are scanned over E + 1 slots, and only the
        // tile total comes back.
        uint32_t *qbcnt = buf<uint32_t>(c, S_QBCNT, (size_t)E + 1);
        uint32_t *qboff = buf<uint32_t>(c, S_QBOFF, (size_t)E + 1);
        hipLaunchKernelGGL(k_bucket_blocks, dim3(blocks_for((uint64_t)E + 1)), dim3(256), 0, s, bstart, num + 1, E, qbcnt);
        exclusive_scan_total<uint32_t>(c, qbcnt, qboff, (size_t)E + 1);
        uint64_t *tcount = buf<uint64_t>(c, S_TCNT, (size_t)E + 1);
        const uint32_t qsplit = (p->flags & IDB_FLAG_SPLIT_TILES) ? 1u : 0u;
        uint64_t *toff = buf<uint64_t>(c, S_TOFF, (size_t)E + 1);
        hipLaunchKernelGGL(k_tile_counts, dim3(blocks_for((uint64_t)E + 1)), dim3(256), 0, s, qboff, num + 1, bstart, E,
                           elen, hilen, banded ? 1u : 0u, qsplit, tcount);
        exclusive_scan_total<uint64_t>(c, tcount, toff, (size_t)E + 1);
        hipLaunchKernelGGL(k_size_tiles, dim3(1), dim3(64), 0, s, num, qboff, toff, sz);
        uint32_t bqt[5];  // sz[2 .. 6]: buckets, query blocks, tiles (low), uniques, tiles (high)
        readback(c, bqt, sz + 2, sizeof(bqt));
        B = bqt[0];
        const uint64_t T64 = (uint64_t)bqt[2] | ((uint64_t)bqt[4] << 32);
        // Tile indices are 32-bit in the pair kernels, and a wave's next
        // index (its current one plus the waves of the grid) must not wrap:
        // keep 2^24 indices of headroom below 2^32.
        if (T64 >= (1ull << 32) - (1ull << 24))
            fail(c, IDB_ERR_UNSUPPORTED, std::to_string(T64) + " tiles in one run: more than " +
                                             std::to_string((1ull << 32) - (1ull << 24)) +
                                             " are not supported (split the largest V/J bucket)");
        T = (uint32_t)T64;
        idb_tile *tiles = buf<idb_tile>(c, S_TILES, std::max<uint32_t>(T, 1u));
        hipLaunchKernelGGL(k_emit_tiles, dim3(blocks_for(bqt[1])), dim3(256), 0, s, qboff, num + 1, bstart, E, elen, hilen,
                           banded ? 1u : 0u, qsplit, skey, lb, gvj, toff, tiles);
        a.tiles = tiles;
        a.n_tiles = T;
        a.ent = ent;
        a.e_codes = (const uint8_t *)ecodes;
        a.e_pk = epk;
        a.S = S;
        a.HW = HW;
        a.K = ham ? std::max(K, 1u) : std::max(std::min(K + 1, 256u), 1u);
        a.bits = bits;
        a.vbits = vb;
        a.jbits = jb;
        a.vw = vw;
        a.jw = jw;
        a.gj = gj;
        a.single_bucket = single;
        a.band_filter = filt ? 1u : 0u;
        // the absolute mode gives every pair the same count-pass threshold
        a.t_const = (p->abs_k >= 0 && !ham) ? (int32_t)tc[0] : -1;
        a.qgram_filter = qgram ? 1u : 0u;
        a.e_qg = eqg;
        a.e_low = elow;
        a.cnt = cnt;
        a.work = work;
        // Neighbour pairs found by the count pass are kept for the fuse step;
        // if they overflow the list, the fuse pass re-evaluates the tiles.
        // Waves own chunks of 256 pairs (64 for small requested capacities,
        // 1,024 for lists of 16M pairs or more: every new chunk is one atomic
        // on one counter, and on the 512-repertoire batch 4x fewer claims
        // took the count pass from 3.40 to 3.15 ms).
        edge_cap = p->edge_capacity ? p->edge_capacity : std::max<unsigned long long>(1ull << 22, 64ull * E);
        edge_cap = std::min<unsigned long long>(edge_cap, 1ull << 31);
        if (c->noalloc) {
            // Allocation-free run: the pair list gets what the scratch holds
            // after the buffers still to come (core flags, labels, output
            // staging: 28 bytes per row, plus 1 MiB for the scans and
            // alignment), at 12 bytes per pair (the pair and its chunk's
            // share of the fill counts); a smaller list only sends cluster
            // growth to the fuse pass.
            const size_t used = (c->na_used + 255) & ~(size_t)255;
            const size_t later = 28ull * N + (1ull << 20);
            const size_t room = IDB_NOALLOC_SCRATCH_BYTES > used + later ? IDB_NOALLOC_SCRATCH_BYTES - used - later : 0;
            edge_cap = std::min<unsigned long long>(edge_cap, room / 12);
            if (edge_cap < IDB_TILE) fail(c, IDB_ERR_UNSUPPORTED, "the allocation-free run's scratch is too small for this input");
        }
        a.chunk_log2 = (force_large || edge_cap >= (1ull << 24)) ? 10u : (edge_cap >= (1ull << 20) ? 8u : 6u);
        a.max_chunks = c->noalloc ? (uint32_t)std::max<unsigned long long>(1ull, edge_cap >> a.chunk_log2)
                                  : (uint32_t)((edge_cap + (1ull << a.chunk_log2) - 1) >> a.chunk_log2);
        edge_cap = (unsigned long long)a.max_chunks << a.chunk_log2;
        a.edges = buf<uint2>(c, S_EDGES, edge_cap);
        a.chunk_fill = buf<uint32_t>(c, S_FILL, a.max_chunks);
        a.n_chunks = small + 12;
        if (learned) {
            // embeddings of every row, then proposals verified by Myers
            uint16_t *zb = buf<uint16_t>(c, S_ZB, (size_t)N * 64);
            check(c, hipEventRecord(c->kev[2], s), "event record");
            if (idb_encoder_embed_rows(x->enc, seq, off, len, N, zb, s) != hipSuccess)
                fail(c, IDB_ERR_HIP, std::string("encoder: ") + idb_encoder_error(x->enc));
            check(c, hipEventRecord(c->kev[3], s), "event record");
            check(c, hipEventSynchronize(c->kev[3]), "event sync");
            check(c, hipEventElapsedTime(&encoder_ms, c->kev[2], c->kev[3]), "event time");
            a.zb = zb;
            a.min_sim = x->min_sim;
        }
    }
    int32_t *core = buf<int32_t>(c, S_CORE, N);
    uint32_t *lab = buf<uint32_t>(c, S_LAB, N);
    const bool dev_out = o->mem == IDB_MEM_DEVICE;
    // host outputs: one device buffer, so that they come back in one transfer
    const size_t ob = align256((size_t)N * 4);
    const uint32_t n_out = o->neighbour_count ? 5u : 4u;
    uint8_t *oall = dev_out ? nullptr : buf<uint8_t>(c, S_OUT_CID, ob * n_out);
    uint32_t *cid = dev_out ? o->clone_id : (uint32_t *)oall;
    int32_t *core_o = dev_out ? o->core : (int32_t *)(oall + ob);
    uint32_t *cop_o = dev_out ? o->copies : (uint32_t *)(oall + 2 * ob);
    uint32_t *unq_o = dev_out ? o->uniques : (uint32_t *)(oall + 3 * ob);
    uint32_t *cnt_o = o->neighbour_count ? (dev_out ? o->neighbour_count : (uint32_t *)(oall + 4 * ob)) : nullptr;
    uint32_t *ustart = n_reps > 0 ? buf<uint32_t>(c, S_USTART, n_reps + 1) : nullptr;
    // Host outputs: copies and uniques are final already, and the core flags
    // are after the core-point stage (outside graph mode); they go back on
    // the side stream into their places in the pinned staging while the
    // clustering runs, so that only clone ids (and counts) are left for the
    // end.
    const bool early_out = !dev_out && c->side && xfer_ready(c, ob * n_out);
    if (early_out) {
        check(c, hipMemsetAsync(unq_o, 0, sizeof(uint32_t) * N, s), "memset");
        if (n_reps > 0)
            hipLaunchKernelGGL(k_ustart, dim3(blocks_for((uint64_t)n_reps + 1)), dim3(256), 0, s, uniq, num, rstart,
                               n_reps, ustart);
        hipLaunchKernelGGL(k_uniques_out, dim3(blocks_for(U)), dim3(256), 0, s, uniq, num, rid, rstart, ustart, unq_o);
        check(c, hipEventRecord(c->side_ev[0], s), "event record");
        check(c, hipStreamWaitEvent(c->side, c->side_ev[0], 0), "side wait");
        check(c, hipMemcpyAsync(c->xfer + 2 * ob, copies, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, c->side),
              "output copy");
        check(c, hipMemcpyAsync(c->xfer + 3 * ob, unq_o, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, c->side),
              "output copy");
    }
    if (E > 0 && T > 0 && (p->flags & IDB_FLAG_AUTOTUNE)) autotune(c, a, N, W, ham, learned, small, work);
    // Learned runs stay uncaptured: when their pair list overflows, the host
    // grows it and runs the learned pass again (below).
    const bool graph = (p->flags & IDB_FLAG_GRAPH) && !(x && x->pairs_out) && !learned;
    bool from_edges = false;

    // Everything after the sizing steps.  As a graph (g) the union / fuse
    // choice is made on the device (each kernel checks the chunk counter) and
    // no event is recorded inside.
    auto tail = [&](bool g) {
        if (E > 0) {
            if (!g) check(c, hipEventRecord(c->kev[0], s), "event record");
            if (learned) check(c, idb_launch_count_learned(a, s), "learned count kernel launch");
            else check(c, ham ? idb_launch_count_hamming(a, s) : idb_launch_count_myers(a, W, s), "count kernel launch");
            if (!g) check(c, hipEventRecord(c->kev[1], s), "event record");
        }
        hipLaunchKernelGGL(k_core_flags, dim3(blocks_for(N)), dim3(256), 0, s, cnt, N, (int32_t)p->min_pts, core, pe,
                           small);
        if (!g && early_out) {
            check(c, hipEventRecord(c->side_ev[1], s), "event record");
            check(c, hipStreamWaitEvent(c->side, c->side_ev[1], 0), "side wait");
            check(c, hipMemcpyAsync(c->xfer + ob, core, sizeof(int32_t) * N, hipMemcpyDeviceToHost, c->side),
                  "output copy");
        }
        if (!g) record(c, 3);
        rng.stage("Grow clusters");

        // --- Grow clusters: union over the recorded neighbour pairs, or a
        // fuse pass over the tiles when the pair list overflowed
        if (E > 0) {
            idb_pair_args f = a;
            f.core = core;
            f.par = par;
            f.bmin = bmin;
            f.cnt = nullptr;
            f.edges = nullptr;
            // The union or the fuse pass is chosen on the device (each kernel
            // reads the chunk counter and returns when it is not the one to
            // run), so no read-back waits between the count pass and the
            // growth; the host learns which ran from the final statistics.
            // Learned runs and callers of the pair list decide on the host.
            if (g || !(learned || (x && x->pairs_out))) {
                if (!g) check(c, hipEventRecord(c->kev[2], s), "event record");
                check(c, idb_launch_union_edges(a, ua, s), "union kernel launch");
                if (!learned) {
                    f.fuse_if_overflow = 1;
                    check(c, ham ? idb_launch_fuse_hamming(f, s) : idb_launch_fuse_myers(f, W, s), "fuse kernel launch");
                }
                if (!g) check(c, hipEventRecord(c->kev[3], s), "event record");
            } else {
                uint32_t n_chunks = 0;
                readback(c, &n_chunks, small + 12, sizeof(n_chunks));
                for (int tries = 0; learned && n_chunks > a.max_chunks && tries < 3; ++tries) {
                    // The learned pass found more neighbour pairs than the
                    // list holds.  A second pass over the tiles would have
                    // to repeat its proposals, so instead the list grows to
                    // what the pass claimed (plus a margin: the claim order
                    // of a new run differs) and the pass runs again from its
                    // starting state: counts back to each row's own copies,
                    // chunk and statistics counters cleared.  The counts of
                    // the first pass were complete, so the core flags stand.
                    const uint64_t want = (uint64_t)n_chunks + n_chunks / 4 + 64;
                    a.max_chunks = (uint32_t)std::min<uint64_t>(want, (1ull << 31) >> a.chunk_log2);
                    a.edges = buf<uint2>(c, S_EDGES, (size_t)a.max_chunks << a.chunk_log2);
                    a.chunk_fill = buf<uint32_t>(c, S_FILL, a.max_chunks);
                    check(c, hipMemcpyAsync(cnt, copies, sizeof(int32_t) * N, hipMemcpyDeviceToDevice, s), "count reset");
                    check(c, hipMemsetAsync(small + 12, 0, 4 * sizeof(uint32_t), s), "chunk reset");
                    check(c, hipMemsetAsync(work, 0, IDB_WORK_U64 * IDB_WORK_COPIES * sizeof(unsigned long long), s),
                          "statistics reset");
                    check(c, hipEventRecord(c->kev[0], s), "event record");
                    check(c, idb_launch_count_learned(a, s), "learned count kernel launch");
                    check(c, hipEventRecord(c->kev[1], s), "event record");
                    readback(c, &n_chunks, small + 12, sizeof(n_chunks));
                }
                from_edges = n_chunks <= a.max_chunks;
                if (learned && !from_edges)
                    fail(c, IDB_ERR_OOM, "neighbour pair list overflow in learned mode after growing it");
                if (x && x->pairs_out) {
                    if (!from_edges) fail(c, IDB_ERR_OOM, "neighbour pair list overflow: raise edge_capacity");
                    std::vector<uint32_t> fill(n_chunks);
                    std::vector<uint2> ed((size_t)n_chunks << a.chunk_log2);
                    if (n_chunks) {
                        check(c, hipMemcpyAsync(fill.data(), a.chunk_fill, sizeof(uint32_t) * n_chunks,
                                                hipMemcpyDeviceToHost, s), "pairs copy");
                        check(c, hipMemcpyAsync(ed.data(), a.edges, sizeof(uint2) * ed.size(), hipMemcpyDeviceToHost, s),
                              "pairs copy");
                        check(c, hipStreamSynchronize(s), "pairs sync");
                    }
                    uint64_t k = 0;
                    for (uint32_t ch = 0; ch < n_chunks; ++ch)
                        for (uint32_t i = 0; i < fill[ch]; ++i, ++k) {
                            if (k >= x->pairs_cap) continue;
                            const uint2 e = ed[((size_t)ch << a.chunk_log2) + i];
                            x->pairs_out[2 * k] = std::min(e.x, e.y);
                            x->pairs_out[2 * k + 1] = std::max(e.x, e.y);
                        }
                    *x->n_pairs_out = k;
                }
                check(c, hipEventRecord(c->kev[2], s), "event record");
                if (from_edges) check(c, idb_launch_union_edges(a, ua, s), "union kernel launch");
                else check(c, ham ? idb_launch_fuse_hamming(f, s) : idb_launch_fuse_myers(f, W, s), "fuse kernel launch");
                check(c, hipEventRecord(c->kev[3], s), "event record");
            }
        }
        if (!g) record(c, 4);
        rng.stage("Reduction of cluster information tree structure");

        // --- Reduction of cluster information tree structure
        hipLaunchKernelGGL(k_roots, dim3(blocks_for(U)), dim3(256), 0, s, uniq, num, core, par, bmin, lab, rflag, N);
        if (p->tree_stats)
            hipLaunchKernelGGL(k_tree_depth, dim3(blocks_for(N)), dim3(256), 0, s, rep, core, par, bmin, N, small + 16,
                               small + 12, a.max_chunks, E > 0 ? 1u : 0u);
        if (!g) record(c, 5);
        rng.stage("Re-number clusters and identify noise");

        // --- Re-number clusters and identify noise
        exclusive_scan_total<uint32_t>(c, rflag, rank, (size_t)N + 1);
        hipLaunchKernelGGL(k_labels, dim3(blocks_for(N)), dim3(256), 0, s, rep, lab, rank, rid, rstart, core, copies, cnt,
                           N, cid, (early_out && !g) ? nullptr : core_o, early_out ? nullptr : cop_o, cnt_o,
                           early_out ? nullptr : unq_o, small);
        if (!early_out) {
            if (n_reps > 0)
                hipLaunchKernelGGL(k_ustart, dim3(blocks_for((uint64_t)n_reps + 1)), dim3(256), 0, s, uniq, num, rstart,
                                   n_reps, ustart);
            hipLaunchKernelGGL(k_uniques_out, dim3(blocks_for(U)), dim3(256), 0, s, uniq, num, rid, rstart, ustart,
                               unq_o);
        }
        check(c, hipGetLastError(), "tail launch");
        if (!g) record(c, 6);
    };

    if (graph) {
        hipGraph_t gr = nullptr;
        check(c, hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), "begin capture");
        try {
            tail(true);
        } catch (...) {
            (void)hipStreamEndCapture(s, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            throw;
        }
        check(c, hipStreamEndCapture(s, &gr), "end capture");
        bool updated = false;
        if (c->gexec) {
            hipGraphNode_t err_node = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            updated = hipGraphExecUpdate(c->gexec, gr, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
            (void)hipGetLastError();
            if (!updated) {
                (void)hipGraphExecDestroy(c->gexec);
                c->gexec = nullptr;
            }
        }
        if (!c->gexec) {
            const hipError_t e = hipGraphInstantiate(&c->gexec, gr, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(gr);
                check(c, e, "graph instantiate");
            }
        }
        (void)hipGraphDestroy(gr);
        check(c, hipGraphLaunch(c->gexec, s), "graph launch");
        for (int k = 3; k <= 6; ++k) record(c, k);
    } else {
        tail(false);
    }
    // host outputs: one pinned transfer, copied out after the final sync
    const bool staged_out = early_out || (!dev_out && xfer_ready(c, ob * n_out));
    if (early_out) {
        // clone ids, the core flags in graph mode, and the counts
        check(c, hipMemcpyAsync(c->xfer, cid, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, s), "output copy");
        if (graph)
            check(c, hipMemcpyAsync(c->xfer + ob, core_o, sizeof(int32_t) * N, hipMemcpyDeviceToHost, s), "output copy");
        if (cnt_o)
            check(c, hipMemcpyAsync(c->xfer + 4 * ob, cnt_o, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, s),
                  "output copy");
    } else if (staged_out) {
        check(c, hipMemcpyAsync(c->xfer, oall, ob * n_out, hipMemcpyDeviceToHost, s), "output copy");
    } else if (!dev_out) {
        check(c, hipMemcpyAsync(o->clone_id, cid, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, s), "output copy");
        check(c, hipMemcpyAsync(o->core, core_o, sizeof(int32_t) * N, hipMemcpyDeviceToHost, s), "output copy");
        check(c, hipMemcpyAsync(o->copies, cop_o, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, s), "output copy");
        check(c, hipMemcpyAsync(o->uniques, unq_o, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, s), "output copy");
        if (cnt_o)
            check(c, hipMemcpyAsync(o->neighbour_count, cnt_o, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, s),
                  "output copy");
    }
    record(c, 7);
    std::vector<uint32_t> us;
    if (n_reps > 0) {
        us.resize(n_reps + 1);
        check(c, hipMemcpyAsync(us.data(), ustart, sizeof(uint32_t) * (n_reps + 1), hipMemcpyDeviceToHost, s), "ustart");
    }
    uint32_t fin[IDB_SMALL_WORDS];
    unsigned long long wkc[IDB_WORK_U64 * IDB_WORK_COPIES], wk[IDB_WORK_U64] = {};
    check(c, hipMemcpyAsync(fin, small, sizeof(fin), hipMemcpyDeviceToHost, s), "stats");
    check(c, hipMemcpyAsync(wkc, work, sizeof(wkc), hipMemcpyDeviceToHost, s), "stats");
    uint32_t scan_err = 0;
    check(c, hipMemcpyAsync(&scan_err, (uint8_t *)c->scan_ws + idb_scan_error_offset(), sizeof(scan_err),
                            hipMemcpyDeviceToHost, s), "stats");
    check(c, hipStreamSynchronize(s), "final sync");
    if (early_out) check(c, hipStreamSynchronize(c->side), "final sync");
    if (scan_err) fail(c, IDB_ERR_HIP, "a scan's look-back did not end (the run's sizes are not valid)");
    if (staged_out) {
        CopyPool::Job jobs[5] = {{o->clone_id, c->xfer, (size_t)N * 4}, {o->core, c->xfer + ob, (size_t)N * 4},
                                 {o->copies, c->xfer + 2 * ob, (size_t)N * 4}, {o->uniques, c->xfer + 3 * ob, (size_t)N * 4},
                                 {o->neighbour_count, c->xfer + 4 * ob, (size_t)N * 4}};
        CopyPool::get().run(jobs, n_out);
    }
    for (uint32_t k = 0; k < IDB_WORK_U64 * IDB_WORK_COPIES; ++k) wk[k % IDB_WORK_U64] += wkc[k];
    if (E > 0 && (graph || !(learned || (x && x->pairs_out)))) {
        from_edges = fin[12] <= a.max_chunks;
        if (learned && !from_edges)
            fail(c, IDB_ERR_OOM, "neighbour pair list overflow in learned mode: raise edge_capacity");
    }
    if (n_uniques_h)
        for (uint32_t r = 0; r < n_reps; ++r) n_uniques_h[r] = us[r + 1] - us[r];
    if (st) {
        st->n_uniques = U;
        st->n_cores = 0;
        st->n_noise = 0;
        for (uint32_t k = 0; k < 16; ++k) {
            st->n_cores += fin[IDB_SMALL_CORES + k];
            st->n_noise += fin[IDB_SMALL_NOISE + k];
        }
        st->max_len = maxlen;
        st->alphabet_size = K;
        st->n_buckets = B;
        st->bucket_entries = E;
        st->tiles = T;
        st->gated_pairs = wk[0];
        st->neighbour_pairs = wk[1];
        st->fuse_pairs = wk[2];
        st->gated_letters = wk[3];
        st->myers_pairs = wk[5];
        st->myers_letters = wk[6];
        st->band_filter = a.band_filter ? 1u : (a.qgram_filter ? 2u : 0u);
        st->fuse_from_edges = from_edges ? 1u : 0u;
        st->tuned_wpb = a.tune_wpb;
        st->tuned_grid_mul = a.tune_grid_mul;
        st->tuned_xcd_order = a.tune_wpb ? (a.no_xcd_order ? 0u : 1u) : 0u;
        st->tuned_trials = 0;
        if (p->flags & IDB_FLAG_AUTOTUNE)
            for (const auto &kv : c->tuned)
                if (kv.second.wpb == a.tune_wpb && kv.second.grid_mul == a.tune_grid_mul) st->tuned_trials = kv.second.trials;
        st->graph_used = graph ? 1u : 0u;
        st->encoder_ms = encoder_ms;
        if (E > 0 && T > 0 && !graph) {
            (void)hipEventElapsedTime(&st->count_kernel_ms, c->kev[0], c->kev[1]);
            (void)hipEventElapsedTime(&st->fuse_kernel_ms, c->kev[2], c->kev[3]);
        }
        if (p->tree_stats)
            for (int k = 0; k <= 50; ++k) st->tree_depth[k] = fin[16 + k];
        // Stage slots follow the reference's timers; duplicate-tree reduction
        // is part of the duplicate stage here (representatives are direct).
        const int from[8] = {0, 1, 1, 2, 3, 4, 5, 0}, to[8] = {1, 1, 2, 3, 4, 5, 6, 7};
        for (int k = 0; k < 8; ++k) {
            float ms = 0.f;
            if (from[k] != to[k]) (void)hipEventElapsedTime(&ms, c->ev[from[k]], c->ev[to[k]]);
            st->stage_ms[k] = ms;
        }
    }
    if (st) {
        uint32_t ncl = 0;
        check(c, hipMemcpy(&ncl, rank + N, sizeof(uint32_t), hipMemcpyDeviceToHost), "stats");
        st->n_clusters = ncl;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
extern "C" {

int idb_create(idb_ctx **out, int device)
{
    if (!out) return IDB_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
        (void)hipGetLastError();
        return IDB_ERR_HIP;
    }
    idb_ctx *c = new idb_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->side_ev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->side_ev[1], hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void **)&c->pinned, 2 * IDB_STAGE_BYTES, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        return IDB_ERR_HIP;
    }
    for (int k = 0; k <= IDB_N_STAGES; ++k)
        if (hipEventCreate(&c->ev[k]) != hipSuccess) {
            delete c;
            return IDB_ERR_HIP;
        }
    for (int k = 0; k < 4; ++k)
        if (hipEventCreate(&c->kev[k]) != hipSuccess) {
            delete c;
            return IDB_ERR_HIP;
        }
    // the scratch array's address on this device (the device is current)
    if (scratch_on(device, &c->scratch) != hipSuccess) {
        (void)hipGetLastError();
        idb_destroy(c);
        return IDB_ERR_HIP;
    }
    *out = c;
    return IDB_OK;
}

void idb_destroy(idb_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int s = 0; s < S_COUNT; ++s)
        if (c->ptr[s]) (void)hipFree(c->ptr[s]);
    for (int k = 0; k <= IDB_N_STAGES; ++k)
        if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
    for (int k = 0; k < 4; ++k)
        if (c->kev[k]) (void)hipEventDestroy(c->kev[k]);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->xfer) (void)hipHostFree(c->xfer);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->side) {
        (void)hipStreamSynchronize(c->side);
        (void)hipStreamDestroy(c->side);
    }
    for (int k = 0; k < 2; ++k)
        if (c->side_ev[k]) (void)hipEventDestroy(c->side_ev[k]);
    delete c;
}

const char *idb_last_error(const idb_ctx *c) { return c ? c->err.c_str() : "null context"; }

int idb_gene_bitsets(uint32_t n, const uint8_t *calls, const uint64_t *call_off, const uint32_t *call_len,
                     uint32_t n_genes, const uint8_t *genes, const uint64_t *gene_off, const uint32_t *gene_len,
                     int32_t *out, uint32_t width)
{
    if ((n && (!call_off || !call_len || !out)) || (n_genes && (!gene_off || !gene_len)))
        return IDB_ERR_ARG;
    if (width < (n_genes + 31) / 32) return IDB_ERR_ARG;
    std::memset(out, 0, sizeof(int32_t) * (size_t)n * width);
    for (uint32_t g = 0; g < n_genes; ++g) {
        const uint32_t gl = gene_len[g];
        // The reference launches one block of len(gene) threads per gene
        // (cuda/BitsetVector.inl:204); 0 or more than 1024 threads fails the
        // launch silently and the bit stays clear.
        if (gl == 0 || gl > 1024) continue;
        const std::string_view needle((const char *)genes + gene_off[g], gl);
        for (uint32_t i = 0; i < n; ++i) {
            const std::string_view hay((const char *)calls + call_off[i], call_len[i]);
            if (hay.find(needle) != std::string_view::npos) out[(size_t)i * width + (g >> 5)] |= (int32_t)(1u << (g & 31));
        }
    }
    return IDB_OK;
}

static int run_guarded(idb_ctx *c, const idb_dataset *d, const uint32_t *rs, uint32_t nr, const idb_params *p,
                       idb_outputs *o, uint32_t *nu, idb_stats *st, const Extra *x = nullptr, bool noalloc = false)
{
    if (!c) return IDB_ERR_ARG;
    if (!d || !p || !o) {
        c->err = "null argument";
        return IDB_ERR_ARG;
    }
    c->err.clear();
    // allocation-free mode lasts for this call only
    struct Mode {
        idb_ctx *c;
        Mode(idb_ctx *cc, bool on) : c(cc)
        {
            c->noalloc = on;
            c->na_used = 0;
            for (int k = 0; k < S_COUNT; ++k) {
                c->na_ptr[k] = nullptr;
                c->na_cap[k] = 0;
            }
        }
        ~Mode() { c->noalloc = false; }
    } mode(c, noalloc);
    try {
        check(c, hipSetDevice(c->device), "hipSetDevice");
        run(c, d, rs, nr, p, o, nu, st, x);
    } catch (const Fail &f) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(c->stream);
        return f.code;
    } catch (const std::exception &e) {
        c->err = e.what();
        return IDB_ERR_HIP;
    }
    return IDB_OK;
}

int idb_dbscan(idb_ctx *c, const idb_dataset *d, const idb_params *p, idb_outputs *o, idb_stats *st)
{
    return run_guarded(c, d, nullptr, 0, p, o, nullptr, st);
}

int idb_dbscan_noalloc(idb_ctx *c, const idb_dataset *d, const idb_params *p, idb_outputs *o, idb_stats *st)
{
    if (!c) return IDB_ERR_ARG;
    if (!d || !p || !o) {
        c->err = "null argument";
        return IDB_ERR_ARG;
    }
    if (p->flags != 0) {
        c->err = "the allocation-free run takes no flags (autotuning and graph capture allocate)";
        return IDB_ERR_ARG;
    }
    if (d->n_rows > IDB_NOALLOC_MAX_ROWS) {
        c->err = "the allocation-free run takes at most " + std::to_string(IDB_NOALLOC_MAX_ROWS) + " rows, got " +
                 std::to_string(d->n_rows);
        return IDB_ERR_UNSUPPORTED;
    }
    if (!c->scratch) {
        c->err = "no scratch array on this device";
        return IDB_ERR_HIP;
    }
    return run_guarded(c, d, nullptr, 0, p, o, nullptr, st, nullptr, true);
}

int idb_dbscan_batch(idb_ctx *c, const idb_dataset *d, const uint32_t *rep_start, uint32_t n_reps, const idb_params *p,
                     idb_outputs *o, uint32_t *n_uniques, idb_stats *st)
{
    if (!c) return IDB_ERR_ARG;
    if (!rep_start || !n_uniques || n_reps == 0 || !d) {
        c->err = "batch needs rep_start, n_uniques and n_reps > 0";
        return IDB_ERR_ARG;
    }
    if (rep_start[0] != 0 || rep_start[n_reps] != d->n_rows) {
        c->err = "rep_start must run from 0 to n_rows";
        return IDB_ERR_ARG;
    }
    for (uint32_t r = 0; r < n_reps; ++r)
        if (rep_start[r + 1] < rep_start[r]) {
            c->err = "rep_start must be non-decreasing";
            return IDB_ERR_ARG;
        }
    return run_guarded(c, d, rep_start, n_reps, p, o, n_uniques, st);
}

int idb_dbscan_learned(idb_ctx *c, idb_encoder *enc, float min_similarity, const idb_dataset *d, const idb_params *p,
                       idb_outputs *o, idb_stats *st)
{
    if (!c) return IDB_ERR_ARG;
    if (!enc) {
        c->err = "null encoder";
        return IDB_ERR_ARG;
    }
    Extra x;
    x.enc = enc;
    x.min_sim = min_similarity;
    return run_guarded(c, d, nullptr, 0, p, o, nullptr, st, &x);
}

int idb_neighbour_pairs(idb_ctx *c, const idb_dataset *d, const idb_params *p, uint32_t *pairs, uint64_t cap,
                        uint64_t *n_pairs)
{
    if (!c) return IDB_ERR_ARG;
    if (!d || !p || !n_pairs || (cap && !pairs)) {
        c->err = "null argument";
        return IDB_ERR_ARG;
    }
    const uint32_t N = d->n_rows;
    std::vector<uint32_t> cid(N), cp(N), uq(N);
    std::vector<int32_t> core(N);
    idb_outputs o{IDB_MEM_HOST, cid.data(), core.data(), cp.data(), uq.data(), nullptr};
    Extra x;
    std::vector<uint32_t> sink(2);
    x.pairs_out = pairs ? pairs : sink.data();
    x.pairs_cap = pairs ? cap : 0;
    x.n_pairs_out = n_pairs;
    const int rc = run_guarded(c, d, nullptr, 0, p, &o, nullptr, nullptr, &x);
    if (rc == IDB_OK && *n_pairs > cap) {
        c->err = "more neighbour pairs than cap";
        return IDB_ERR_OOM;
    }
    return rc;
}

int idb_gene_bitsets_gpu(idb_ctx *c, uint32_t mem, uint32_t n, const uint8_t *calls, const uint64_t *call_off,
                         const uint32_t *call_len, uint32_t n_genes, const uint8_t *genes, const uint64_t *gene_off,
                         const uint32_t *gene_len, int32_t *out, uint32_t width)
{
    if (!c) return IDB_ERR_ARG;
    c->err.clear();
    if ((n && (!call_off || !call_len || !out)) || (n_genes && (!genes || !gene_off || !gene_len)) ||
        (mem != IDB_MEM_HOST && mem != IDB_MEM_DEVICE)) {
        c->err = "null or inconsistent argument";
        return IDB_ERR_ARG;
    }
    if (width < (n_genes + 31) / 32) {
        c->err = "width below ceil(n_genes / 32)";
        return IDB_ERR_ARG;
    }
    if (width > IDB_GENE_MAX_WIDTH) {
        c->err = "more than 256 bit words per row";
        return IDB_ERR_UNSUPPORTED;
    }
    if (n == 0 || width == 0) return IDB_OK;
    try {
        check(c, hipSetDevice(c->device), "hipSetDevice");
        hipStream_t s = c->stream;
        // headers and 4-byte words of the names (names that cannot match get
        // len = 0xFFFFFFFF: empty, or longer than the reference's 1024-thread
        // launch allows, cuda/BitsetVector.inl:204)
        std::vector<idb_gene_hdr> hdr(std::max(n_genes, 1u));
        std::vector<uint32_t> words(1, 0u);
        for (uint32_t g = 0; g < n_genes; ++g) {
            const uint32_t gl = gene_len[g];
            if (gl == 0 || gl > 1024) {
                hdr[g] = {0u, 0u, 0xFFFFFFFFu, 0u};
                continue;
            }
            const uint32_t woff = (uint32_t)words.size();
            for (uint32_t k = 0; k < gl; k += 4) {
                uint32_t v = 0;
                for (uint32_t b = 0; b < 4 && k + b < gl; ++b) v |= (uint32_t)genes[gene_off[g] + k + b] << (8 * b);
                words.push_back(v);
            }
            uint32_t last = 0;
            for (uint32_t b = 0; b < 4 && b < gl; ++b) last |= (uint32_t)genes[gene_off[g] + gl - std::min(gl, 4u) + b] << (8 * b);
            hdr[g] = {words[woff], last, gl, woff};
        }
        idb_gene_hdr *hd = buf<idb_gene_hdr>(c, S_GHDR, hdr.size());
        uint32_t *wd = buf<uint32_t>(c, S_GWORDS, words.size());
        check(c, hipMemcpyAsync(hd, hdr.data(), sizeof(idb_gene_hdr) * hdr.size(), hipMemcpyHostToDevice, s), "gene copy");
        check(c, hipMemcpyAsync(wd, words.data(), sizeof(uint32_t) * words.size(), hipMemcpyHostToDevice, s), "gene copy");
        const uint8_t *cd = calls;
        const uint64_t *od = call_off;
        const uint32_t *ld = call_len;
        int32_t *outd = out;
        if (mem == IDB_MEM_HOST) {
            uint64_t bytes = 0;
            for (uint32_t i = 0; i < n; ++i) bytes = std::max<uint64_t>(bytes, call_off[i] + call_len[i]);
            if (bytes && !calls) fail(c, IDB_ERR_ARG, "null call bytes");
            cd = bytes ? stage_in<uint8_t>(c, S_GCALLS, calls, bytes, mem) : buf<uint8_t>(c, S_GCALLS, 4);
            od = stage_in<uint64_t>(c, S_GOFF, call_off, n, mem);
            ld = stage_in<uint32_t>(c, S_GLEN, call_len, n, mem);
            outd = buf<int32_t>(c, S_GOUT, (size_t)n * width);
        }
        check(c, idb_launch_gene_bitsets(cd, od, ld, n, hd, wd, (uint32_t)words.size(), n_genes, outd, width, s),
              "gene bitset launch");
        if (mem == IDB_MEM_HOST)
            check(c, hipMemcpyAsync(out, outd, sizeof(int32_t) * (size_t)n * width, hipMemcpyDeviceToHost, s),
                  "bitset copy");
        check(c, hipStreamSynchronize(s), "sync");
    } catch (const Fail &f) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(c->stream);
        return f.code;
    }
    return IDB_OK;
}

int idb_pair_distances(idb_ctx *c, const idb_dataset *d, const uint32_t *pairs, uint64_t n_pairs, uint32_t use_hamming,
                       uint32_t *dist)
{
    if (!c) return IDB_ERR_ARG;
    if (!d || (n_pairs && (!pairs || !dist))) {
        c->err = "null argument";
        return IDB_ERR_ARG;
    }
    c->err.clear();
    try {
        check(c, hipSetDevice(c->device), "hipSetDevice");
        hipStream_t s = c->stream;
        const uint32_t N = d->n_rows;
        for (uint64_t i = 0; i < 2 * n_pairs; ++i)
            if (pairs[i] >= N) fail(c, IDB_ERR_ARG, "pair row out of range");
        if (n_pairs == 0) return IDB_OK;
        std::vector<uint32_t> lh;
        const uint32_t *len_h = d->len;
        if (d->mem == IDB_MEM_DEVICE) {
            lh.resize(N);
            check(c, hipMemcpy(lh.data(), d->len, sizeof(uint32_t) * N, hipMemcpyDeviceToHost), "length copy");
            len_h = lh.data();
        }
        uint32_t maxlen = 0;
        for (uint64_t i = 0; i < 2 * n_pairs; ++i) maxlen = std::max(maxlen, len_h[pairs[i]]);
        const uint32_t W = std::max(1u, (maxlen + 31) / 32);
        if (!use_hamming && W > IDB_MAX_W) fail(c, IDB_ERR_UNSUPPORTED, "rows longer than 256 letters");
        const uint8_t *seq = stage_in<uint8_t>(c, S_SEQ, d->seq, d->seq_bytes, d->mem);
        const uint64_t *off = stage_in<uint64_t>(c, S_OFF, d->off, N, d->mem);
        const uint32_t *len = stage_in<uint32_t>(c, S_LEN, d->len, N, d->mem);
        if (!seq) seq = buf<uint8_t>(c, S_SEQ, 4);
        uint2 *pd = buf<uint2>(c, S_PAIRS, n_pairs);
        uint32_t *dd = buf<uint32_t>(c, S_DIST, n_pairs);
        check(c, hipMemcpyAsync(pd, pairs, sizeof(uint2) * n_pairs, hipMemcpyHostToDevice, s), "pairs copy");
        check(c, idb_launch_pair_distances(seq, off, len, pd, n_pairs, W, use_hamming, dd, s), "distance launch");
        check(c, hipMemcpyAsync(dist, dd, sizeof(uint32_t) * n_pairs, hipMemcpyDeviceToHost, s), "distance copy");
        check(c, hipStreamSynchronize(s), "sync");
    } catch (const Fail &f) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(c->stream);
        return f.code;
    }
    return IDB_OK;
}

}  // extern "C"
