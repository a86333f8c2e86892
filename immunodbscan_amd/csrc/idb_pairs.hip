// idb_pairs.hip — the O(pairs) kernels: neighbour counting (R8) and cluster
// fusion (R9/R10) over bucket tiles, for gfx950.
//
// Replaces the reference's pair kernels half_core_points / fuse_clusters
// (Levenshtein, cuda/PSDBSCAN.hcu:751-931) and half_hamming_core_points /
// fuse_hamming_clusters (Hamming, :585-742).  The reference evaluates every
// unique pair, one block per pair in Levenshtein mode with a barrier per DP
// anti-diagonal.  Here a wave owns one tile of <= 64 target entries x <= 64
// query entries of one V/J bucket:
//   * each lane holds one target (the Myers pattern) and builds its match
//     masks Peq[letter][word] in its own LDS column ([letter][word][lane]
//     layout: a wave's 64 lanes touch 64 consecutive dwords, no bank
//     conflicts);
//   * lane l also holds query l's 32-byte entry record for the whole tile; the
//     query loop broadcasts it with readlane, so the loop itself issues no
//     memory loads in the common case (one round trip per tile fetches both
//     records; code words come in 16-byte loads);
//   * gate work per pair is the length band, the triangle order inside the
//     diagonal tile and, only when both rows carry a gene below the bucket's,
//     the canonical bucket test (lowest common V gene and J gene equal the
//     bucket's), so each unordered pair is evaluated in exactly one bucket.
//     Gene residues mod 16 decide most of those tests without loading the
//     bitsets.
// Count pass: copies-weighted counts and the neighbour pairs (kept for the
// fuse step).  Fuse pass (fallback when the pair list overflows): pairs with at
// least one core; core-core pairs are joined in a lock-free union-find that
// links the larger root under the smaller (roots are component minima), and a
// non-core row records its minimum-index core neighbour with atomicMin (the
// canonical form of the reference's first-come atomicCAS at :113-122).
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "idb_internal.h"
#include "pairdist.h"

namespace {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint32_t rl(uint32_t v, uint32_t lane) { return __builtin_amdgcn_readlane(v, lane); }

__device__ __forceinline__ uint32_t ld_agent(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ uint32_t uf_find(uint32_t *par, uint32_t x)
{
    uint32_t p = ld_agent(par + x);
    while (p != x) {
        const uint32_t gp = ld_agent(par + p);
        if (gp != p) {
            // path halving: gp is an ancestor of x, so the shortcut is safe
            __hip_atomic_store(par + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        x = p;
        p = gp;
    }
    return x;
}

// Link the larger root under the smaller one.  After a failed exchange the
// larger of the two roots strictly decreases, so the loop ends.
__device__ void uf_union(uint32_t *par, uint32_t a, uint32_t b)
{
    a = uf_find(par, a);
    b = uf_find(par, b);
    while (a != b) {
        if (a < b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        uint32_t expected = a;
        if (__hip_atomic_compare_exchange_strong(par + a, &expected, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
        a = uf_find(par, expected);
        b = uf_find(par, b);
    }
}

__device__ __forceinline__ int lowest_common(const int32_t *a, const int32_t *b, uint32_t w)
{
    for (uint32_t k = 0; k < w; ++k) {
        const uint32_t m = (uint32_t)(a[k] & b[k]);
        if (m) return (int)(k * 32 + __builtin_ctz(m));
    }
    return -1;
}

// Exact canonical-bucket test from the bitsets (single-bucket mode: the V/J
// gate of cuda/BitsetVector.inl:214-237).
__device__ bool pair_in_bucket(const idb_pair_args &a, uint32_t x, uint32_t y, uint32_t bucket)
{
    const int32_t *vx = a.vbits + (size_t)x * a.vw, *vy = a.vbits + (size_t)y * a.vw;
    const int32_t *jx = a.jbits + (size_t)x * a.jw, *jy = a.jbits + (size_t)y * a.jw;
    const int lv = lowest_common(vx, vy, a.vw), lj = lowest_common(jx, jy, a.jw);
    if (a.single_bucket) return lv >= 0 && lj >= 0;
    return lv == (int)(bucket / a.gj) && lj == (int)(bucket % a.gj);
}

// The signature holds the residues mod 16 of the genes in question (canonical
// mode: genes below the bucket's; single-bucket mode: all genes).  Canonical
// mode: no common V residue and no common J residue -> no common lower gene ->
// canonical.  Single-bucket mode: no common V (or J) residue -> no common gene
// -> the gate fails.  Otherwise the bitsets decide.
__device__ __forceinline__ bool gate_pair(const idb_pair_args &a, bool chk_x, bool chk_y, uint32_t sx, uint32_t sy,
                                          uint32_t x, uint32_t y, uint32_t bucket)
{
    if (!(a.single_bucket || (chk_x && chk_y))) return true;
    const uint32_t c = sx & sy;
    if (a.single_bucket) {
        if ((c & 0xFFFFu) == 0u || (c >> 16) == 0u) return false;
    } else if (c == 0u) {
        return true;
    }
    return pair_in_bucket(a, x, y, bucket);
}

// Neighbour pairs are staged per wave in LDS (IDB_TILE entries) and appended
// to a chunk of the edge list that the wave owns; a new chunk costs one
// atomic on the chunk counter, so that counter sees one atomic per chunk
// instead of one per hit or per stage.  Past the capacity the counter keeps
// growing and the pairs are dropped, so the overflow is visible and the fuse
// pass over the tiles runs instead.
struct EdgeStage {
    uint2 *buf;       // this wave's IDB_TILE staging slots
    uint32_t n;       // staged pairs (wave-uniform)
    uint32_t chunk;   // the wave's current chunk, IDB_NONE before the first
    uint32_t fill;    // pairs written into it
    uint32_t total;   // pairs this wave produced
};

// This block's copy of the run's statistics counters.
__device__ __forceinline__ unsigned long long *work_copy(const idb_pair_args &a)
{
    return a.work + IDB_WORK_U64 * (blockIdx.x & (IDB_WORK_COPIES - 1u));
}

__device__ __forceinline__ void close_chunk(const idb_pair_args &a, const EdgeStage &st, uint32_t lane)
{
    if (st.chunk != IDB_NONE && st.chunk < a.max_chunks && lane == 0) a.chunk_fill[st.chunk] = st.fill;
}

__device__ __forceinline__ void flush_edges(const idb_pair_args &a, EdgeStage &st, uint32_t lane)
{
    if (st.n == 0) return;
    if (st.chunk == IDB_NONE || st.fill + st.n > (1u << a.chunk_log2)) {
        close_chunk(a, st, lane);
        uint32_t c = 0;
        if (lane == 0) c = atomicAdd(a.n_chunks, 1u);
        st.chunk = __shfl(c, 0, 64);
        st.fill = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    if (st.chunk < a.max_chunks && lane < st.n)
        a.edges[((size_t)st.chunk << a.chunk_log2) + st.fill + lane] = st.buf[lane];
    st.fill += st.n;
    st.total += st.n;
    st.n = 0;
}

__device__ __forceinline__ void stage_edges(const idb_pair_args &a, EdgeStage &st, bool keep, uint32_t x, uint32_t y,
                                            uint32_t lane)
{
    const unsigned long long km = __ballot(keep);
    if (!km || !a.edges) return;
    const uint32_t cnt = (uint32_t)__popcll(km);
    if (st.n + cnt > IDB_TILE) flush_edges(a, st, lane);
    if (keep) {
        const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(km >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)km, 0u));
        st.buf[st.n + below] = make_uint2(x, y);
    }
    st.n += cnt;
}

// End of a count-pass wave: last stage, last chunk's fill, pair total.
__device__ __forceinline__ void finish_edges(const idb_pair_args &a, EdgeStage &st, uint32_t lane)
{
    if (!a.edges) return;
    flush_edges(a, st, lane);
    close_chunk(a, st, lane);
    if (lane == 0 && st.total && a.work) atomicAdd(work_copy(a) + 4, (unsigned long long)st.total);
}

// Fuse pass launched behind the union step without a host decision: it runs
// only when the pair list overflowed.
__device__ __forceinline__ bool fuse_skipped(const idb_pair_args &a)
{
    return a.fuse_if_overflow && *(volatile const uint32_t *)a.n_chunks <= a.max_chunks;
}

// The grid holds only as many waves as can be resident at once (no partly
// empty second round of waves trails the launch); wave w takes tiles w,
// w + waves, w + 2 waves, ...  A shared tile counter balances better in
// principle, but its device-scope atomics wait behind the pass's other
// atomics: measured on the 200k-row workload, claiming tiles (one counter per
// XCD, half of the tiles static) made the count pass slower, not faster.
__device__ __forceinline__ uint32_t next_tile(uint32_t ti, uint32_t total_waves) { return ti + total_waves; }

// The Myers count pass deals the rounds back and forth: wave w takes
// position w of even rounds and position waves - 1 - w of odd ones, so a wave
// that drew from a heavy stretch of one round draws from the other end of
// the next (3% off the 200k-row count pass).  n_tiles past the end.
__device__ __forceinline__ uint32_t step_tile(uint32_t ti, uint32_t total_waves, uint32_t n_tiles)
{
    if (ti >= n_tiles) return n_tiles;
    const uint32_t r = ti / total_waves, pos = ti % total_waves;
    const uint64_t nx = (uint64_t)(r + 1) * total_waves + (total_waves - 1 - pos);
    return nx < n_tiles ? (uint32_t)nx : n_tiles;
}

// Blocks are dealt to the 8 XCDs round-robin; renumber them so that each XCD
// starts on a contiguous run of tiles (neighbouring tiles share query blocks).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb)
{
    const uint32_t per = nb / 8, rem = nb % 8, x = b % 8, k = b / 8;
    return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

__device__ __forceinline__ int32_t t_of(uint32_t t2, bool fuse)
{
    return fuse ? (int32_t)(int16_t)(t2 >> 16) : (int32_t)(int16_t)(t2 & 0xFFFFu);
}

__device__ __forceinline__ idb_entry load_entry(const idb_entry *e)
{
    const uint4 *p = (const uint4 *)e;
    const uint4 a = p[0], b = p[1];
    idb_entry r;
    r.row = a.x;
    r.cp = a.y;
    r.len_chk = a.z;
    r.t2 = a.w;
    r.len = b.x;
    r.pres1 = b.y;
    r.pres2 = b.z;
    r.sig = b.w;
    return r;
}

// A tile record: the list is read-only during a pass and the index is
// wave-uniform, so this is a scalar load (waited on by lgkmcnt, which keeps
// it clear of the vector loads in flight).
__device__ __forceinline__ idb_tile tile_rec(const idb_pair_args &a, uint32_t ti)
{
    typedef const __attribute__((address_space(4))) uint32_t *const_words;
    const const_words w = (const_words)a.tiles + 4 * (size_t)ti;
    idb_tile t;
    t.q0 = w[0];
    t.t0 = w[1];
    t.bucket = w[2];
    t.shape = w[3];
    return t;
}

// What a Myers tile needs from global memory, loaded one tile ahead: the
// record, this lane's query and target entries and, for one-word rows
// (CODES), both rows' codes (16 or 32 bytes).
template <bool CODES>
struct TilePre {
    idb_tile tile;
    uint4 te0, te1, qe0, qe1;  // the two halves of the target's and the query's records
    uint4 qc0, qc1, tc0, tc1;
};

// A record from its two 16-byte halves (kept as loaded, so that the loads
// write their registers directly).
__device__ __forceinline__ idb_entry entry_of(const uint4 &p0, const uint4 &p1)
{
    idb_entry r;
    r.row = p0.x;
    r.cp = p0.y;
    r.len_chk = p0.z;
    r.t2 = p0.w;
    r.len = p1.x;
    r.pres1 = p1.y;
    r.pres2 = p1.z;
    r.sig = p1.w;
    return r;
}

template <bool CODES>
__device__ __forceinline__ void tile_load(const idb_pair_args &a, const idb_tile &t, uint32_t lane, TilePre<CODES> &d)
{
    const uint32_t qn = t.shape & 0xFFu, tn = (t.shape >> 8) & 0xFFu;
    d.tile = t;
    d.te0 = d.te1 = d.qe0 = d.qe1 = make_uint4(0u, 0u, 0u, 0u);
    d.qc0 = d.qc1 = d.tc0 = d.tc1 = make_uint4(0u, 0u, 0u, 0u);
    if (lane < tn) {
        const uint4 *p = (const uint4 *)(a.ent + t.t0 + lane);
        d.te0 = p[0];
        d.te1 = p[1];
    }
    if (lane < qn) {
        const uint4 *p = (const uint4 *)(a.ent + t.q0 + lane);
        d.qe0 = p[0];
        d.qe1 = p[1];
    }
    if (CODES) {
        const bool wide = a.S > 16;
        if (lane < qn) {
            const uint4 *q = (const uint4 *)(a.e_codes + (size_t)(t.q0 + lane) * a.S);
            d.qc0 = q[0];
            if (wide) d.qc1 = q[1];
        }
        if (lane < tn) {
            const uint4 *c = (const uint4 *)(a.e_codes + (size_t)(t.t0 + lane) * a.S);
            d.tc0 = c[0];
            if (wide) d.tc1 = c[1];
        }
    }
}

// Hit bookkeeping of the batched Myers path (one pair per lane).
// Counts go to the tile's LDS accumulators of query qi and target tl.
template <bool FUSE>
__device__ __forceinline__ void record_hit(const idb_pair_args &a, EdgeStage &est, bool hit, uint32_t x, uint32_t y,
                                           uint32_t cx, uint32_t cy, int32_t core_x, int32_t core_y, uint32_t lane,
                                           int32_t *qacc, int32_t *tacc, uint32_t qi, uint32_t tl)
{
    if (!FUSE) {
        if (hit) {
            atomicAdd(qacc + qi, (int32_t)cy);
            atomicAdd(tacc + tl, (int32_t)cx);
        }
        stage_edges(a, est, hit, x, y, lane);
    } else if (hit) {
        if (core_x && core_y) uf_union(a.par, x, y);
        else if (core_x) atomicMin(a.bmin + y, x);
        else atomicMin(a.bmin + x, y);
    }
}

// Myers over up to 64 queued (query, target) pairs of the current tile, one
// pair per lane.  The target's Peq column belongs to lane tl; query and
// target fields come from the lanes that hold them.
template <int W, bool FUSE>
__device__ __forceinline__ void myers_batch(const idb_pair_args &a, const idb_tile &tile, const uint32_t *peq,
                                            const uint16_t *queue, uint32_t n, uint32_t lane, const idb_entry &qe,
                                            const idb_entry &te, int32_t core_y, int32_t q_core, EdgeStage &est,
                                            int32_t *qacc, int32_t *tacc, unsigned long long &hits,
                                            unsigned long long &mlet)
{
    // Every cross-lane read runs with the whole wave active: a cross-lane
    // read from a lane that is switched off returns 0.
    const bool act = lane < n;
    const uint32_t e = queue[lane < n ? lane : 0u];
    const int qi = (int)(e >> 8), tl = (int)(e & 0xFFu);
    const uint32_t ty = (uint32_t)__shfl((int)te.row, tl, 64);
    const uint32_t tm = (uint32_t)__shfl((int)te.len_chk, tl, 64) & 0xFFFFu;
    const uint32_t tcy = (uint32_t)__shfl((int)te.cp, tl, 64);
    const uint32_t tt2 = (uint32_t)__shfl((int)te.t2, tl, 64);
    const int32_t tcore = __shfl(core_y, tl, 64);
    const uint32_t x = (uint32_t)__shfl((int)qe.row, qi, 64);
    const uint32_t lxq = (uint32_t)__shfl((int)qe.len_chk, qi, 64) & 0xFFFFu;
    const uint32_t cx = (uint32_t)__shfl((int)qe.cp, qi, 64);
    const uint32_t qt2 = (uint32_t)__shfl((int)qe.t2, qi, 64);
    const int32_t core_x = __shfl(q_core, qi, 64);
    const uint32_t lx = act ? lxq : 0u;
    const int32_t t = !act ? -1 : (lx >= tm ? t_of(qt2, FUSE) : t_of(tt2, FUSE));
    mlet += lx;
    idb_myers<W> st;
    idb_myers_init<W>(st, tm);
    const uint32_t top_w = tm ? (tm - 1) >> 5 : 0xFFFFFFFFu, top_b = tm ? (tm - 1) & 31 : 0u;
    const uint32_t *xc = (const uint32_t *)(a.e_codes + (size_t)(tile.q0 + qi) * a.S);
    for (uint32_t j = 0; j < lx; j += 4) {
        const uint32_t word = xc[j >> 2];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            if (j + k < lx) {
                const uint32_t c = (word >> (8 * k)) & 0xFFu;
                uint32_t eq[W];
#pragma unroll
                for (int w = 0; w < W; ++w) eq[w] = peq[(c * W + w) * 64 + tl];
                idb_myers_step<W>(st, eq, top_w, top_b);
            }
        }
    }
    const int32_t dist = tm ? st.score : (int32_t)lx;
    const bool hit = act && dist <= t;
    if (!FUSE) hits += __popcll(__ballot(hit));
    record_hit<FUSE>(a, est, hit, x, ty, cx, tcy, core_x, tcore, lane, qacc, tacc, (uint32_t)qi, (uint32_t)tl);
}

// Myers restricted to a diagonal band (patterns of 3 or more words, pair
// thresholds of at most 15).  Only lev(x, y) <= t is asked, and E. Ukkonen
// ("Algorithms for approximate string matching", Inf. Control 1985) showed
// that cells with |i - j| > t cannot lie on a path of cost <= t; so at text
// letter j only the pattern rows j - K .. j + K matter (K = the largest
// threshold of the batch, uniform over the wave), which lie in two
// consecutive 32-row words.  Each step runs the usual multi-word recurrence
// on that two-word window only:
//   * the words below the window were never touched and still hold the
//     first column's vertical deltas (all +1), which is a valid assignment
//     of values larger than K to cells outside the band; when the window
//     moves down a word, the new lower word starts from that state;
//   * the word above the window is out of the band, and the boundary row
//     above the window gets horizontal delta +1 per letter (shift-in 1, add
//     carry 0), as row 0 does in the unbanded recurrence;
//   * the tracked score is D[B][j] + (m - B) for the window's last row
//     B = min(m, 32 (lo + 2)) (rows below B count +1 each), which moves by
//     the horizontal delta at row B, and equals D[m][n] at the end because
//     row m is inside the band at the last letter (|m - n| <= t).
// Every value computed this way is at least min(true value, K + 1), and a
// cell whose true value is at most K is exact (its optimal path stays in the
// band), so dist <= t decides exactly as the unbanded recurrence does.
template <int W, bool FUSE>
__device__ __forceinline__ void myers_batch_band(const idb_pair_args &a, const idb_tile &tile, const uint32_t *peq,
                                                 const uint16_t *queue, uint32_t n, uint32_t lane,
                                                 const idb_entry &qe, const idb_entry &te, int32_t core_y,
                                                 int32_t q_core, EdgeStage &est, int32_t *qacc, int32_t *tacc,
                                                 unsigned long long &hits, unsigned long long &mlet, uint32_t K)
{
    const bool act = lane < n;
    const uint32_t e = queue[lane < n ? lane : 0u];
    const int qi = (int)(e >> 8), tl = (int)(e & 0xFFu);
    const uint32_t ty = (uint32_t)__shfl((int)te.row, tl, 64);
    const uint32_t tm = (uint32_t)__shfl((int)te.len_chk, tl, 64) & 0xFFFFu;
    const uint32_t tcy = (uint32_t)__shfl((int)te.cp, tl, 64);
    const uint32_t tt2 = (uint32_t)__shfl((int)te.t2, tl, 64);
    const int32_t tcore = __shfl(core_y, tl, 64);
    const uint32_t x = (uint32_t)__shfl((int)qe.row, qi, 64);
    const uint32_t lxq = (uint32_t)__shfl((int)qe.len_chk, qi, 64) & 0xFFFFu;
    const uint32_t cx = (uint32_t)__shfl((int)qe.cp, qi, 64);
    const uint32_t qt2 = (uint32_t)__shfl((int)qe.t2, qi, 64);
    const int32_t core_x = __shfl(q_core, qi, 64);
    const uint32_t lx = act ? lxq : 0u;
    const int32_t t = !act ? -1 : (lx >= tm ? t_of(qt2, FUSE) : t_of(tt2, FUSE));
    mlet += lx;
    // two-word window [lo, lo + 1] over the pattern rows (lo <= W - 2)
    uint32_t pa = 0xFFFFFFFFu, ma = 0u, pb = 0xFFFFFFFFu, mb = 0u;
    int32_t score = (int32_t)tm;
    uint32_t lo = 0;
    uint32_t lmax = lx;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lmax = max(lmax, (uint32_t)__shfl_xor((int)lmax, o, 64));
    lmax = __builtin_amdgcn_readfirstlane(lmax);
    // Peq address of letter c in the window: c * (W * 64) + lo * 64 + tl words
    uint32_t pbase = (uint32_t)tl;
    // the tracked row B = min(m, 32 (lo + 2)): in word b or word a, at bit
    // (B - 1) & 31; fixed while the window stays.  While the band lies
    // inside word a, word b is untouched (its rows still hold the first
    // column's +1 deltas), only word a is updated, and the tracked row is
    // B' = min(m, 32 (lo + 1)) at bit bit_a: with word b untouched,
    // D[B][j] = D[B'][j] + (B - B'), so both rows give the same tracked score
    // and the switch between them is exact.
    bool in_b = false;
    uint32_t bit = 0, bit_a = 0;
    auto place = [&]() {
        const uint32_t B = min(tm, 32u * (lo + 2u));
        in_b = B > 32u * (lo + 1u);
        bit = (B - 1u) & 31u;
        bit_a = (min(tm, 32u * (lo + 1u)) - 1u) & 31u;
    };
    place();
    // One text letter against the two-word window (two = true) or against
    // word a only (the band inside word a, word b untouched).
    auto step = [&](uint32_t c, uint32_t j, bool two) {
        const uint32_t *pe = peq + c * (W * 64u) + pbase;
        if (!two) {
            const uint32_t ea = pe[0];
            const uint32_t xva = ea | ma;
            const uint32_t xha = (((ea & pa) + pa) ^ pa) | ea;
            const uint32_t pha = ma | ~(xha | pa);
            const uint32_t mha = pa & xha;
            const int32_t dsc = (int32_t)((pha >> bit_a) & 1u) - (int32_t)((mha >> bit_a) & 1u);
            const uint32_t pha_s = (pha << 1) | 1u, mha_s = mha << 1;
            // letters past the row's end: selects keep the state
            const bool upd = j <= lx;
            score += upd ? dsc : 0;
            pa = upd ? (mha_s | ~(xva | pha_s)) : pa;
            ma = upd ? (pha_s & xva) : ma;
            return;
        }
        const uint32_t ea = pe[0], eb = pe[64];
        // word a: boundary row above it (shift-in +1, no add carry)
        const uint32_t xva = ea | ma;
        const uint64_t sa64 = (uint64_t)(ea & pa) + pa;
        const uint32_t xha = ((uint32_t)sa64 ^ pa) | ea;
        const uint32_t pha = ma | ~(xha | pa);
        const uint32_t mha = pa & xha;
        // word b: carries from word a
        const uint32_t cin = (uint32_t)(sa64 >> 32);
        const uint32_t xvb = eb | mb;
        const uint32_t sb = (eb & pb) + pb + cin;
        const uint32_t xhb = (sb ^ pb) | eb;
        const uint32_t phb = mb | ~(xhb | pb);
        const uint32_t mhb = pb & xhb;
        const uint32_t phs_ = in_b ? phb : pha, mhs_ = in_b ? mhb : mha;
        const int32_t dsc = (int32_t)((phs_ >> bit) & 1u) - (int32_t)((mhs_ >> bit) & 1u);
        const uint32_t pha_s = (pha << 1) | 1u, mha_s = mha << 1;
        const uint32_t phb_s = (phb << 1) | (pha >> 31), mhb_s = (mhb << 1) | (mha >> 31);
        const bool upd = j <= lx;
        score += upd ? dsc : 0;
        pa = upd ? (mha_s | ~(xva | pha_s)) : pa;
        ma = upd ? (pha_s & xva) : ma;
        pb = upd ? (mhb_s | ~(xvb | phb_s)) : pb;
        mb = upd ? (phb_s & xvb) : mb;
    };
    // window of column j: the word of row j - K, kept within [0, W - 2]
    auto want_of = [&](uint32_t j) { return j > K + 1 ? min((j - K - 1) >> 5, (uint32_t)W - 2u) : 0u; };
    // Four letters at a time, each block of four with no branch per letter
    // when the window stays put over it (uniform tests): one-word steps
    // while the band stays inside word a, else two-word steps; a block over
    // which the window moves takes the per-letter path.
    auto group = [&](const uint4 &cur, uint32_t j0) {
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            if (j0 + 4 * q >= lmax) break;  // uniform
            const uint32_t word = q == 0 ? cur.x : (q == 1 ? cur.y : (q == 2 ? cur.z : cur.w));
            const uint32_t jf = j0 + 4 * q + 1;  // 1-based text column of the block's first letter
            if (want_of(jf + 3u) == lo) {
                if (jf + 3u + K <= 32u * (lo + 1u)) {
#pragma unroll
                    for (uint32_t k = 0; k < 4; ++k) step((word >> (8 * k)) & 0xFFu, jf + k, false);
                } else {
#pragma unroll
                    for (uint32_t k = 0; k < 4; ++k) step((word >> (8 * k)) & 0xFFu, jf + k, true);
                }
                continue;
            }
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                const uint32_t j = jf + k;  // past lmax: no lane updates
                const uint32_t want = want_of(j);
                if (want != lo) {  // uniform: one word further down
                    pa = pb;
                    ma = mb;
                    pb = 0xFFFFFFFFu;
                    mb = 0u;
                    lo = want;
                    pbase = (uint32_t)tl + lo * 64u;
                    place();
                }
                step((word >> (8 * k)) & 0xFFu, j, true);
            }
        }
    };
    // Sixteen letters of the query per step group, from two register
    // buffers: while one group runs, the next-but-one is loaded into the
    // buffer just used (S is a multiple of 16, so every load is in the row).
    const uint4 *xq = (const uint4 *)(a.e_codes + (size_t)(tile.q0 + qi) * a.S);
    const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
    uint4 bufa = lx ? xq[0] : zero4, bufb = lx > 16 ? xq[1] : zero4;
    for (uint32_t j0 = 0; j0 < lmax; j0 += 32) {
        group(bufa, j0);
        bufa = j0 + 32 < lx ? xq[(j0 >> 4) + 2] : zero4;
        if (j0 + 16 >= lmax) break;
        group(bufb, j0 + 16);
        bufb = j0 + 48 < lx ? xq[(j0 >> 4) + 3] : zero4;
    }
    const int32_t dist = tm ? score : (int32_t)lx;
    const bool hit = act && dist <= t;
    if (!FUSE) hits += __popcll(__ballot(hit));
    record_hit<FUSE>(a, est, hit, x, ty, cx, tcy, core_x, tcore, lane, qacc, tacc, (uint32_t)qi, (uint32_t)tl);
}

// The batch with the banded recurrence when the pattern words and the
// batch's largest threshold allow it, else the full one.
template <int W, bool FUSE>
__device__ __forceinline__ void myers_batch_any(const idb_pair_args &a, const idb_tile &tile, const uint32_t *peq,
                                                const uint16_t *queue, uint32_t n, uint32_t lane, const idb_entry &qe,
                                                const idb_entry &te, int32_t core_y, int32_t q_core, EdgeStage &est,
                                                int32_t *qacc, int32_t *tacc, unsigned long long &hits,
                                                unsigned long long &mlet)
{
    if (W >= 3) {
        // the batch's largest pair threshold, uniform over the wave
        const uint32_t e = queue[lane < n ? lane : 0u];
        const int qi = (int)(e >> 8), tl = (int)(e & 0xFFu);
        const uint32_t tm = (uint32_t)__shfl((int)te.len_chk, tl, 64) & 0xFFFFu;
        const uint32_t tt2 = (uint32_t)__shfl((int)te.t2, tl, 64);
        const uint32_t lxq = (uint32_t)__shfl((int)qe.len_chk, qi, 64) & 0xFFFFu;
        const uint32_t qt2 = (uint32_t)__shfl((int)qe.t2, qi, 64);
        int32_t t = lane < n ? (lxq >= tm ? t_of(qt2, FUSE) : t_of(tt2, FUSE)) : 0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) t = max(t, __shfl_xor(t, o, 64));
        const int32_t K = __builtin_amdgcn_readfirstlane(t);
        if (K >= 0 && K <= 15) {
            myers_batch_band<W, FUSE>(a, tile, peq, queue, n, lane, qe, te, core_y, q_core, est, qacc, tacc, hits,
                                      mlet, (uint32_t)K);
            return;
        }
    }
    myers_batch<W, FUSE>(a, tile, peq, queue, n, lane, qe, te, core_y, q_core, est, qacc, tacc, hits, mlet);
}

// The same for rows of at most 32 letters (one word, code rows of 16 or 32
// bytes): each lane holds its query row's codes in registers from the start
// of the tile (qc0, qc1), so a batch takes them from lane qi with cross-lane
// reads instead of waiting on global loads, and the unrolled letter loop
// leaves the compiler free to issue the Peq reads of a word together.
template <bool FUSE>
__device__ __forceinline__ void myers_batch1(const idb_pair_args &a, const uint32_t *peq, const uint16_t *queue,
                                             uint32_t n, uint32_t lane, const idb_entry &qe, const idb_entry &te,
                                             int32_t core_y, int32_t q_core, EdgeStage &est, int32_t *qacc,
                                             int32_t *tacc, unsigned long long &hits, unsigned long long &mlet,
                                             const uint4 &qc0, const uint4 &qc1, bool wide,
                                             const uint4 *qcodes = nullptr)
{
    const bool act = lane < n;
    const uint32_t e = queue[lane < n ? lane : 0u];
    const int qi = (int)(e >> 8), tl = (int)(e & 0xFFu);
    const uint32_t ty = (uint32_t)__shfl((int)te.row, tl, 64);
    const uint32_t tm = (uint32_t)__shfl((int)te.len_chk, tl, 64) & 0xFFFFu;
    const uint32_t tcy = (uint32_t)__shfl((int)te.cp, tl, 64);
    const uint32_t tt2 = (uint32_t)__shfl((int)te.t2, tl, 64);
    const int32_t tcore = __shfl(core_y, tl, 64);
    const uint32_t x = (uint32_t)__shfl((int)qe.row, qi, 64);
    const uint32_t lxq = (uint32_t)__shfl((int)qe.len_chk, qi, 64) & 0xFFFFu;
    const uint32_t cx = (uint32_t)__shfl((int)qe.cp, qi, 64);
    const uint32_t qt2 = (uint32_t)__shfl((int)qe.t2, qi, 64);
    const int32_t core_x = __shfl(q_core, qi, 64);
    uint32_t w[8];
    if (qcodes) {
        // the tile's query codes staged in LDS, two uint4 per query
        const uint4 c0 = qcodes[2 * qi], c1 = wide ? qcodes[2 * qi + 1] : make_uint4(0u, 0u, 0u, 0u);
        w[0] = c0.x, w[1] = c0.y, w[2] = c0.z, w[3] = c0.w;
        w[4] = c1.x, w[5] = c1.y, w[6] = c1.z, w[7] = c1.w;
    } else {
        w[0] = (uint32_t)__shfl((int)qc0.x, qi, 64);
        w[1] = (uint32_t)__shfl((int)qc0.y, qi, 64);
        w[2] = (uint32_t)__shfl((int)qc0.z, qi, 64);
        w[3] = (uint32_t)__shfl((int)qc0.w, qi, 64);
        if (wide) {
            w[4] = (uint32_t)__shfl((int)qc1.x, qi, 64);
            w[5] = (uint32_t)__shfl((int)qc1.y, qi, 64);
            w[6] = (uint32_t)__shfl((int)qc1.z, qi, 64);
            w[7] = (uint32_t)__shfl((int)qc1.w, qi, 64);
        } else {
            w[4] = w[5] = w[6] = w[7] = 0u;
        }
    }
    const uint32_t lx = act ? lxq : 0u;
    const int32_t t = !act ? -1 : (lx >= tm ? t_of(qt2, FUSE) : t_of(tt2, FUSE));
    mlet += lx;
    idb_myers<1> st;
    idb_myers_init<1>(st, tm);
    const uint32_t top_w = tm ? 0u : 0xFFFFFFFFu, top_b = tm ? (tm - 1) & 31 : 0u;
    const uint32_t *pcol = peq + tl;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (!__ballot(4u * k < lx)) break;
        const uint32_t word = w[k];
        uint32_t eq[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) eq[b] = pcol[((word >> (8 * b)) & 0xFFu) * 64];
        if (!__ballot(4u * k + 4u > lx)) {
            // every lane's row covers the whole word: no per-letter select
#pragma unroll
            for (int b = 0; b < 4; ++b) idb_myers_step<1>(st, eq + b, top_w, top_b);
        } else {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                idb_myers<1> nx = st;
                idb_myers_step<1>(nx, eq + b, top_w, top_b);
                if (4u * k + b < lx) st = nx;
            }
        }
    }
    const int32_t dist = tm ? st.score : (int32_t)lx;
    const bool hit = act && dist <= t;
    if (!FUSE) hits += __popcll(__ballot(hit));
    record_hit<FUSE>(a, est, hit, x, ty, cx, tcy, core_x, tcore, lane, qacc, tacc, (uint32_t)qi, (uint32_t)tl);
}

// FILT selects an exact lower bound that a pair must pass before Myers;
// surviving pairs are queued in LDS and verified by Myers 64 at a time.
//   1 (alphabets of 8-31 letters): the absent-letter bound, a few bit
//     operations per pair;
//   2 (alphabets of at most 4 letters, longer rows): the 3-gram profile bound
//     of E. Ukkonen ("Approximate string-matching with q-grams and maximal
//     matches", TCS 1992): an edit changes at most q q-grams on either side,
//     so lev(x, y) >= |P_x - P_y|_1 / (2 q).  Profiles are 64 saturating byte
//     counts per row (saturation only lowers the L1 distance, so the bound
//     stays a lower bound); the L1 distance is 16 v_sad_u8 per pair, query
//     profiles are read from LDS by broadcast.
// Occupancy: 4 waves per SIMD (128 VGPRs) in general; the 3-gram path with
// banded Myers (3 or more words) needs about 150 and runs at 3, which on 20M
// nucleotide rows is 12% faster than 4 waves with registers spilled.
template <int W, bool FUSE, int FILT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FILT == 2 && W >= 3 ? 3 : 4))) void k_pairs_myers(
    idb_pair_args a, uint32_t wpb)
{
    extern __shared__ uint32_t lds[];
    const uint32_t lane = lane_id();
    const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform, kept in a scalar register
    const uint32_t K = a.K;
    constexpr uint32_t QG_WORDS = FILT == 2 ? IDB_TILE * IDB_QGRAM_WORDS : 0u;
    uint32_t *peq = lds + (size_t)wib * (idb_myers_lds_per_wave(K, W) / sizeof(uint32_t) + QG_WORDS);
    uint32_t *qg = peq + idb_myers_lds_per_wave(K, W) / sizeof(uint32_t);
    EdgeStage est{(uint2 *)(peq + K * W * 64), 0u, IDB_NONE, 0u, 0u};
    uint16_t *queue = (uint16_t *)(peq + K * W * 64 + 2 * IDB_TILE);
    int32_t *qacc = (int32_t *)(peq + K * W * 64 + 3 * IDB_TILE);
    int32_t *tacc = qacc + IDB_TILE;
    idb_entry *qrec = (idb_entry *)(tacc + IDB_TILE);
    if (FUSE && fuse_skipped(a)) return;
    const uint32_t total_waves = gridDim.x * wpb;
    // consecutive tiles (often of one query block, with similar work) go to
    // different blocks and so to different CUs, not to the waves of one block
    const uint32_t wave0 = (a.no_xcd_order ? blockIdx.x : xcd_block(blockIdx.x, gridDim.x)) + wib * gridDim.x;
    unsigned long long gated = 0, hits = 0, gletters = 0, mpairs = 0, mletters = 0;
    unsigned long long glane = 0, gllane = 0;  // per-lane gated pairs and target letters (constant threshold)

    // static order; the next tile's data and the record of the one after are
    // loaded while the current tile runs
    const uint32_t T = a.n_tiles;
    uint32_t ti = wave0;
    uint32_t ti1 = step_tile(ti, total_waves, T);
    idb_tile rec1 = ti1 < T ? tile_rec(a, ti1) : idb_tile{};
    TilePre<FILT == 1> cur{};
    if (ti < T) tile_load<FILT == 1>(a, tile_rec(a, ti), lane, cur);
    const bool wide = a.S > 16;
    while (ti < T) {
        // The tile's records (loaded one tile ago) are waited for once; its
        // other loads go out together below, and the next tile's records
        // only once this tile's loaded data is in registers or LDS (the wait
        // counter is in order: a first use after later loads waits for them).
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        const idb_tile tile = cur.tile;
        const uint32_t qn = tile.shape & 0xFFu, tn = (tile.shape >> 8) & 0xFFu;
        const bool diag = (tile.shape >> 16) & 1u;
        const bool tvalid = lane < tn;
        const uint32_t p = tile.t0 + (tvalid ? lane : 0u);
        const idb_entry te = entry_of(cur.te0, cur.te1), qe = entry_of(cur.qe0, cur.qe1);
        const uint4 qc0 = cur.qc0, qc1 = cur.qc1;
        const uint32_t y = te.row, m = te.len_chk & 0xFFFFu, cy = te.cp;
        const bool chk_y = (te.len_chk >> 16) & 1u;
        const int32_t t_m = tvalid ? t_of(te.t2, FUSE) : -1;
        int32_t core_y = 0, q_core = 0;
        if (FUSE) {
            if (tvalid) core_y = a.core[y];
            if (lane < qn) q_core = a.core[qe.row];
        }
        // Peq for this lane's pattern, in its own LDS column (LDS atomic ORs:
        // no read-back between letters)
        for (uint32_t c = 0; c < K; ++c)
#pragma unroll
            for (int w = 0; w < W; ++w) peq[(c * W + w) * 64 + lane] = 0u;
        if (FILT == 1) {
            if (tvalid) {
                const uint32_t ws[8] = {cur.tc0.x, cur.tc0.y, cur.tc0.z, cur.tc0.w,
                                        cur.tc1.x, cur.tc1.y, cur.tc1.z, cur.tc1.w};
#pragma unroll
                for (uint32_t i = 0; i < 32; ++i)
                    if (i < m) atomicOr(peq + ((ws[i >> 2] >> (8 * (i & 3))) & 0xFFu) * 64 + lane, 1u << i);
            }
        } else if (tvalid) {
            // code words 16 bytes at a time (S is a multiple of 16), five
            // loads in flight at once
            const uint4 *codes = (const uint4 *)(a.e_codes + (size_t)p * a.S);
            constexpr uint32_t CG = 5;
            for (uint32_t i0 = 0; i0 < m; i0 += 16 * CG) {
                uint4 v[CG];
#pragma unroll
                for (uint32_t g = 0; g < CG; ++g) v[g] = i0 + 16 * g < m ? codes[(i0 >> 4) + g] : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
                for (uint32_t g = 0; g < CG; ++g) {
                    const uint32_t ws[4] = {v[g].x, v[g].y, v[g].z, v[g].w};
#pragma unroll
                    for (uint32_t k = 0; k < 16; ++k) {
                        const uint32_t i = i0 + 16 * g + k;
                        if (i < m) {
                            const uint32_t c = (ws[k >> 2] >> (8 * (k & 3))) & 0xFFu;
                            atomicOr(peq + (c * W + (i >> 5)) * 64 + lane, 1u << (i & 31));
                        }
                    }
                }
            }
        }
        const uint32_t top_w = m ? (m - 1) >> 5 : 0xFFFFFFFFu, top_b = m ? (m - 1) & 31 : 0u;
        int32_t acc_y = 0;
        uint32_t qlen = 0;  // queued pairs (FILT)
        // query qi pairs with this lane iff qi < dlim (upper triangle on the
        // diagonal tile, nothing past the tile's targets)
        const uint32_t dlim = !tvalid ? 0u : (diag ? lane : 64u);
        const uint32_t ny1 = ~te.pres1;
        // per-tile counters fit 32 bits: at most 4096 pairs of 16-bit lengths
        uint32_t tg = 0, tgl = 0, tm = 0;
        if (!FUSE) {
            qacc[lane] = 0;
            tacc[lane] = 0;
        }
        // query records for the loop: one LDS broadcast per query instead of
        // a chain of lane reads into scalar registers
        qrec[lane] = qe;
        // genes below the bucket's (exact gene test), when the entries have them
        const bool exact_low = a.e_low != nullptr && !a.single_bucket;
        uint4 tlow = make_uint4(0u, 0u, 0u, 0u), qlow = tlow;
        if (exact_low) {
            if (tvalid) tlow = a.e_low[tile.t0 + lane];
            if (lane < qn) qlow = a.e_low[tile.q0 + lane];
        }
        uint32_t tq[FILT == 2 ? IDB_QGRAM_WORDS : 1];
        if (FILT == 2) {
            const uint4 *src = (const uint4 *)(a.e_qg + (size_t)p * IDB_QGRAM_WORDS);
#pragma unroll
            for (int k = 0; k < (int)IDB_QGRAM_WORDS / 4; ++k) {
                const uint4 v = tvalid ? src[k] : make_uint4(0, 0, 0, 0);
                tq[4 * k] = v.x;
                tq[4 * k + 1] = v.y;
                tq[4 * k + 2] = v.z;
                tq[4 * k + 3] = v.w;
            }
            if (lane < qn) {
                const uint4 *qs = (const uint4 *)(a.e_qg + (size_t)(tile.q0 + lane) * IDB_QGRAM_WORDS);
                uint4 *qd = (uint4 *)(qg + lane * IDB_QGRAM_WORDS);
#pragma unroll
                for (int k = 0; k < (int)IDB_QGRAM_WORDS / 4; ++k) qd[k] = qs[k];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        }
        // (used here, so that they are waited for before the next loads)
        const bool tlow_any = (tlow.x | tlow.y | tlow.z | tlow.w) != 0u;
        const bool qlow_set = __ballot((qlow.x | qlow.y | qlow.z | qlow.w) != 0u) != 0ull;
        // the next tile's records
        TilePre<FILT == 1> nxt{};
        if (ti1 < T) tile_load<FILT == 1>(a, rec1, lane, nxt);
        const uint32_t ti2 = step_tile(ti1, total_waves, T);
        const idb_tile rec2 = ti2 < T ? tile_rec(a, ti2) : idb_tile{};

        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        if (FILT == 1 && !FUSE) {
            // Count pass over one-word rows: the query's fields come from
            // lane qi by v_readlane (no LDS round trip), and every test runs
            // for every lane and becomes a mask, so a query costs one branch
            // unless it may share a lower gene or has a pair for Myers.
            const uint32_t y1 = te.pres1;
            for (uint32_t qi = 0; qi < qn; ++qi) {
                const uint32_t lc = __builtin_amdgcn_readlane(qe.len_chk, qi);
                const uint32_t lx = lc & 0xFFFFu;
                const int32_t t_q = t_of(__builtin_amdgcn_readlane(qe.t2, qi), false);
                const uint32_t p1x = __builtin_amdgcn_readlane(qe.pres1, qi);
                const int32_t t = lx >= m ? t_q : t_m;
                const int32_t dl = (int32_t)__builtin_amdgcn_sad_u16(lx, m, 0u);
                // absent-letter bound, both directions (see below)
                const uint32_t lbx = __builtin_popcount(p1x & ny1);
                const uint32_t lby = __builtin_popcount(y1 & ~p1x);
                const int32_t lb = (int32_t)(lbx > lby ? lbx : lby);
                unsigned long long vm = __ballot(qi < dlim && dl <= t);
                if (vm && (a.single_bucket || ((lc >> 16) & 1u))) {
                    const uint32_t c = __builtin_amdgcn_readlane(qe.sig, qi) & te.sig;
                    const bool valid = (vm >> lane) & 1ull;
                    bool ok = true, need;
                    if (a.single_bucket) {
                        ok = (c & 0xFFFFu) != 0u && (c >> 16) != 0u;
                        need = valid && ok;
                    } else {
                        need = valid && chk_y && c != 0u;
                    }
                    const uint32_t x = __builtin_amdgcn_readlane(qe.row, qi);
                    if (__ballot(need) && need) ok = pair_in_bucket(a, x, y, tile.bucket);
                    vm &= __ballot(ok);
                }
                const uint32_t nv = (uint32_t)__popcll(vm);
                tg += nv;
                tgl += nv * lx;
                const unsigned long long pm = vm & __ballot(lb <= t);
                if (!pm) continue;
                if ((pm >> lane) & 1ull) {
                    const uint32_t below =
                        __builtin_amdgcn_mbcnt_hi((uint32_t)(pm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)pm, 0u));
                    queue[qlen + below] = (uint16_t)((qi << 8) | lane);
                }
                qlen += (uint32_t)__popcll(pm);
                tm += (uint32_t)__popcll(pm);
                if (qlen >= IDB_TILE) {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    myers_batch1<FUSE>(a, peq, queue, IDB_TILE, lane, qe, te, core_y, q_core, est, qacc, tacc, hits,
                                       mletters, qc0, qc1, wide);
                    const uint32_t rest = qlen - IDB_TILE;
                    const uint16_t moved = lane < rest ? queue[IDB_TILE + lane] : (uint16_t)0;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    if (lane < rest) queue[lane] = moved;
                    qlen = rest;
                }
            }
        }
        // Constant threshold over long rows with the 3-gram bound (the
        // count pass of the nucleotide workloads): the band and the bound are
        // tested with no branch per query, as in k_count1.  Queries run in
        // windows of up to 32, four at a time; the second half of the
        // profiles is read only when some pair of the four passes the first
        // half.  A pair that passes sets bit r - 1 - k of the lane's mask mb
        // (query w0 + k, r queries in the window); the exact gene test
        // corrects each window; the bits then become Myers queue entries.
        // Gated pairs are counted per lane (letters: the target's).
        const bool fast2 = FILT == 2 && !FUSE && a.t_const >= 0 && !a.single_bucket;
        if (FILT == 2 && !FUSE && fast2) {
            const int32_t t = a.t_const;
            const uint32_t mlane = tvalid ? m : 0xFFFFu;
            const uint64_t chkq = __ballot(lane < qn && ((qe.len_chk >> 16) & 1u));
            uint32_t gcnt = 0;
            constexpr uint32_t QG = 4, HALF = IDB_QGRAM_WORDS / 8;  // uint4 reads per half profile
            for (uint32_t w0 = 0; w0 < qn; w0 += 32) {
                const uint32_t r = min(32u, (qn - w0 + QG - 1u) & ~(QG - 1u));
                uint32_t mb = 0;
                for (uint32_t g = 0; g < r; g += QG) {
                    uint32_t dl[QG], l1[QG];
#pragma unroll
                    for (uint32_t k = 0; k < QG; ++k) {
                        const uint32_t qi = w0 + g + k;
                        const uint32_t ql = qi < qn ? qrec[qi].len : 0x7FFFu;  // past qn: fails the band
                        dl[k] = __builtin_amdgcn_sad_u16(ql, mlane, 0u);
                        const uint4 *qv = (const uint4 *)(qg + qi * IDB_QGRAM_WORDS);
                        uint32_t l = 0;
#pragma unroll
                        for (uint32_t h = 0; h < HALF; ++h) {
                            const uint4 v = qv[h];
                            l = __builtin_amdgcn_sad_u8(v.x, tq[4 * h], l);
                            l = __builtin_amdgcn_sad_u8(v.y, tq[4 * h + 1], l);
                            l = __builtin_amdgcn_sad_u8(v.z, tq[4 * h + 2], l);
                            l = __builtin_amdgcn_sad_u8(v.w, tq[4 * h + 3], l);
                        }
                        l1[k] = l;
                    }
                    bool any = false;
#pragma unroll
                    for (uint32_t k = 0; k < QG; ++k)
                        any = any || ((int32_t)dl[k] <= t && (int32_t)l1[k] <= 6 * t && w0 + g + k < dlim);
                    if (__ballot(any)) {
#pragma unroll
                        for (uint32_t k = 0; k < QG; ++k) {
                            const uint4 *qv = (const uint4 *)(qg + (w0 + g + k) * IDB_QGRAM_WORDS);
                            uint32_t l = l1[k];
#pragma unroll
                            for (uint32_t h = HALF; h < 2 * HALF; ++h) {
                                const uint4 v = qv[h];
                                l = __builtin_amdgcn_sad_u8(v.x, tq[4 * h], l);
                                l = __builtin_amdgcn_sad_u8(v.y, tq[4 * h + 1], l);
                                l = __builtin_amdgcn_sad_u8(v.z, tq[4 * h + 2], l);
                                l = __builtin_amdgcn_sad_u8(v.w, tq[4 * h + 3], l);
                            }
                            l1[k] = l;
                        }
                    }
#pragma unroll
                    for (uint32_t k = 0; k < QG; ++k) {
                        const bool band = (int32_t)dl[k] <= t && w0 + g + k < dlim;
                        gcnt += band ? 1u : 0u;
                        mb = mb + mb + ((band && (int32_t)l1[k] <= 6 * t) ? 1u : 0u);
                    }
                }
                // exact gene test for the window's queries that need it
                for (uint64_t cm = (chkq >> w0) & (r == 32 ? 0xFFFFFFFFull : ((1ull << r) - 1ull)); cm;
                     cm &= cm - 1ull) {
                    const uint32_t k = (uint32_t)__builtin_ctzll(cm), qi = w0 + k;
                    const idb_entry q = qrec[qi];
                    const bool band = (int32_t)__builtin_amdgcn_sad_u16(q.len, mlane, 0u) <= t && qi < dlim;
                    bool ok = true;
                    if (exact_low) {
                        const uint32_t common = (__builtin_amdgcn_readlane(qlow.x, qi) & tlow.x) |
                                                (__builtin_amdgcn_readlane(qlow.y, qi) & tlow.y) |
                                                (__builtin_amdgcn_readlane(qlow.z, qi) & tlow.z) |
                                                (__builtin_amdgcn_readlane(qlow.w, qi) & tlow.w);
                        ok = !(band && tlow_any && common != 0u);
                    } else {
                        const bool need = band && chk_y && (__builtin_amdgcn_readfirstlane(q.sig) & te.sig) != 0u;
                        if (!__ballot(need)) continue;
                        if (need) ok = pair_in_bucket(a, q.row, y, tile.bucket);
                    }
                    if (!ok) {
                        --gcnt;
                        mb &= ~(1u << (r - 1u - k));
                    }
                }
                for (;;) {
                    const uint64_t anyb = __ballot(mb != 0u);
                    if (!anyb) break;
                    // the earliest query still set in each lane
                    const uint32_t hb = 31u - (uint32_t)__builtin_clz(mb | 1u);
                    if (mb != 0u) {
                        const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(anyb >> 32),
                                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)anyb, 0u));
                        queue[qlen + below] = (uint16_t)(((w0 + r - 1u - hb) << 8) | lane);
                        mb &= ~(1u << hb);
                    }
                    const uint32_t np = (uint32_t)__popcll(anyb);
                    qlen += np;
                    tm += np;
                    if (qlen >= IDB_TILE) {
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        myers_batch_any<W, FUSE>(a, tile, peq, queue, IDB_TILE, lane, qe, te, core_y, q_core, est,
                                                 qacc, tacc, hits, mletters);
                        const uint32_t rest = qlen - IDB_TILE;
                        const uint16_t moved = lane < rest ? queue[IDB_TILE + lane] : (uint16_t)0;
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        if (lane < rest) queue[lane] = moved;
                        qlen = rest;
                    }
                }
            }
            glane += gcnt;
            gllane += (unsigned long long)gcnt * m;
        }
        for (uint32_t qi = ((FILT == 1 && !FUSE) || fast2) ? qn : 0u; qi < qn; ++qi) {
            const idb_entry q = qrec[qi];
            const uint32_t x = q.row;
            // the query's length and thresholds are wave-uniform: scalar registers
            const uint32_t lc = __builtin_amdgcn_readfirstlane(q.len_chk);
            const uint32_t lx = lc & 0xFFFFu;
            const bool chk_x = (lc >> 16) & 1u;
            const int32_t t_q = t_of(__builtin_amdgcn_readfirstlane(q.t2), FUSE);
            // the pair's threshold is the longer row's; |lx - m| in one v_sad_u16
            const int32_t t = lx >= m ? t_q : t_m;
            const int32_t dl = (int32_t)__builtin_amdgcn_sad_u16(lx, m, 0u);
            bool valid = qi < dlim && dl <= t;
            int32_t core_x = 0;
            if (FUSE) {
                core_x = (int32_t)rl((uint32_t)q_core, qi);
                valid = valid && (core_x | core_y);
            }
            // V/J gate: only a query that may share a lower gene (or every
            // query in single-bucket mode) needs it, a wave-uniform test
            if (exact_low && chk_x) {
                // a common gene below the bucket's: not this pair's bucket
                if (qlow_set) {
                    const uint32_t common = (__builtin_amdgcn_readlane(qlow.x, qi) & tlow.x) |
                                            (__builtin_amdgcn_readlane(qlow.y, qi) & tlow.y) |
                                            (__builtin_amdgcn_readlane(qlow.z, qi) & tlow.z) |
                                            (__builtin_amdgcn_readlane(qlow.w, qi) & tlow.w);
                    valid = valid && !(tlow_any && common != 0u);
                }
            } else if (a.single_bucket || chk_x) {
                const uint32_t c = __builtin_amdgcn_readfirstlane(q.sig) & te.sig;
                bool need;
                if (a.single_bucket) {
                    valid = valid && (c & 0xFFFFu) != 0u && (c >> 16) != 0u;
                    need = valid;
                } else {
                    need = valid && chk_y && c != 0u;
                }
                if (__ballot(need) && need) valid = pair_in_bucket(a, x, y, tile.bucket);
            }
            const unsigned long long vm = __ballot(valid);
            if (!vm) continue;
            const uint32_t nv = (uint32_t)__popcll(vm);
            tg += nv;
            tgl += nv * lx;

            if (FILT) {
                bool pass;
                if (FILT == 1) {
                    // Absent-letter bound: a letter of x whose symbol never
                    // occurs in y must be substituted or deleted, one edit per
                    // position, so lev(x, y) >= #{i : x_i not in y} >= the
                    // number of symbols of x absent from y (and symmetrically):
                    // one AND and one popcount of the letter-set planes per
                    // direction.  Counting each symbol's first two occurrences
                    // (a second plane) let 16% fewer pairs through to Myers
                    // on 200k protein rows but cost 4 more operations per
                    // pair, and the pass ran 3.5% slower.
                    const uint32_t p1x = __builtin_amdgcn_readfirstlane(q.pres1);
                    const uint32_t lbx = __builtin_popcount(p1x & ny1);
                    const uint32_t lby = __builtin_popcount(te.pres1 & ~p1x);
                    pass = valid && (int32_t)lbx <= t && (int32_t)lby <= t;
                } else {
                    // half the profile first: its L1 distance is already a
                    // lower bound, and most pairs fail it
                    const uint4 *qv = (const uint4 *)(qg + qi * IDB_QGRAM_WORDS);
                    uint32_t l1 = 0;
#pragma unroll
                    for (int k = 0; k < (int)IDB_QGRAM_WORDS / 4; ++k) {
                        if (k == (int)IDB_QGRAM_WORDS / 8 && !__ballot(valid && (int32_t)l1 <= 6 * t)) break;
                        const uint4 v = qv[k];
                        l1 = __builtin_amdgcn_sad_u8(v.x, tq[4 * k], l1);
                        l1 = __builtin_amdgcn_sad_u8(v.y, tq[4 * k + 1], l1);
                        l1 = __builtin_amdgcn_sad_u8(v.z, tq[4 * k + 2], l1);
                        l1 = __builtin_amdgcn_sad_u8(v.w, tq[4 * k + 3], l1);
                    }
                    pass = valid && (int32_t)l1 <= 6 * t;
                }
                const unsigned long long pm = __ballot(pass);
                if (!pm) continue;
                const uint32_t below =
                    __builtin_amdgcn_mbcnt_hi((uint32_t)(pm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)pm, 0u));
                if (pass) queue[qlen + below] = (uint16_t)((qi << 8) | lane);
                qlen += (uint32_t)__popcll(pm);
                tm += (uint32_t)__popcll(pm);
                if (qlen >= IDB_TILE) {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    if (FILT == 1)
                        myers_batch1<FUSE>(a, peq, queue, IDB_TILE, lane, qe, te, core_y, q_core, est, qacc, tacc,
                                           hits, mletters, qc0, qc1, wide);
                    else
                        myers_batch_any<W, FUSE>(a, tile, peq, queue, IDB_TILE, lane, qe, te, core_y, q_core,
                                                 est, qacc, tacc, hits, mletters);
                    const uint32_t rest = qlen - IDB_TILE;
                    const uint16_t moved = lane < rest ? queue[IDB_TILE + lane] : (uint16_t)0;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    if (lane < rest) queue[lane] = moved;
                    qlen = rest;
                }
                continue;
            }

            tm += nv;
            mletters += valid ? lx : 0u;
            idb_myers<W> st;
            idb_myers_init<W>(st, m);
            const uint32_t *xc = (const uint32_t *)(a.e_codes + (size_t)(tile.q0 + qi) * a.S);
            for (uint32_t j = 0; j < lx; j += 4) {
                const uint32_t word = xc[j >> 2];
#pragma unroll
                for (uint32_t k = 0; k < 4; ++k) {
                    if (j + k < lx) {
                        const uint32_t c = (word >> (8 * k)) & 0xFFu;
                        uint32_t eq[W];
#pragma unroll
                        for (int w = 0; w < W; ++w) eq[w] = peq[(c * W + w) * 64 + lane];
                        idb_myers_step<W>(st, eq, top_w, top_b);
                    }
                }
            }
            const int32_t dist = m ? st.score : (int32_t)lx;
            const bool hit = valid && dist <= t;
            const unsigned long long hm = __ballot(hit);
            if (!hm) continue;
            if (!FUSE) {
                hits += __popcll(hm);
                if (hit) {
                    acc_y += (int32_t)q.cp;
                    atomicAdd(qacc + qi, (int32_t)cy);
                }
                stage_edges(a, est, hit, x, y, lane);
            } else if (hit) {
                if (core_x && core_y) uf_union(a.par, x, y);
                else if (core_x) atomicMin(a.bmin + y, x);
                else atomicMin(a.bmin + x, y);
            }
        }
        if (FILT && qlen) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            if (FILT == 1)
                myers_batch1<FUSE>(a, peq, queue, qlen, lane, qe, te, core_y, q_core, est, qacc, tacc, hits,
                                   mletters, qc0, qc1, wide);
            else
                myers_batch_any<W, FUSE>(a, tile, peq, queue, qlen, lane, qe, te, core_y, q_core, est, qacc, tacc,
                                         hits, mletters);
        }
        gated += tg;
        gletters += tgl;
        mpairs += tm;
        if (!FUSE) {
            // one atomic per row and tile
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            const int32_t qa = qacc[lane];
            acc_y += tacc[lane];
            if (lane < qn && qa) atomicAdd(a.cnt + qe.row, qa);
            if (acc_y) atomicAdd(a.cnt + y, acc_y);
        }
        cur = nxt;
        ti = ti1;
        ti1 = ti2;
        rec1 = rec2;
    }
    if (!FUSE) finish_edges(a, est, lane);
    mletters = wave_sum_u64(mletters);
    if (FILT == 2 && !FUSE) {
        // the constant-threshold loop's per-lane counts
        gated += wave_sum_u64(glane);
        gletters += wave_sum_u64(gllane);
    }
    if (lane == 0 && a.work) {
        atomicAdd(work_copy(a) + (FUSE ? 2 : 0), gated);
        if (!FUSE) {
            atomicAdd(work_copy(a) + 1, hits);
            atomicAdd(work_copy(a) + 3, gletters);
            atomicAdd(work_copy(a) + 5, mpairs);
            atomicAdd(work_copy(a) + 6, mletters);
        }
    }
}

// Count pass over one-word rows (at most 32 letters) of alphabets of 8-31
// letters, outside single-bucket mode: the 200k- and 2M-row protein
// workloads.  Lane l holds target l (its Peq column, length, threshold and
// presence planes); the loop walks the tile's queries, two per iteration.
// A query's gate record (length, gene check flag, count threshold, presence
// planes, gene residues) is one 16-byte LDS word read by broadcast, so a
// query costs its pair tests and no cross-lane traffic: the length band, the
// triangle of a diagonal tile (a lane mask built in scalar registers), the
// absent-letter bound (two AND / bit-count pairs) and, only for a query
// that may share a lower gene, the canonical-bucket test.  Survivors are
// queued in LDS and verified by Myers 64 at a time (myers_batch1).  TCONST:
// every pair has the same count threshold (the absolute "k edits" mode), held
// in a scalar register.
__host__ __device__ inline size_t idb_count1_lds_per_wave(uint32_t K)
{
    // Peq columns, edge staging, the Myers queue, count accumulators of
    // queries and targets, the queries' 16-byte gate records and their letter
    // codes (32 bytes)
    return ((size_t)K * 64 + 2 * IDB_TILE + IDB_TILE + 2 * IDB_TILE) * sizeof(uint32_t) + 3 * IDB_TILE * sizeof(uint4);
}

template <bool TCONST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_count1(idb_pair_args a, uint32_t wpb)
{
    extern __shared__ uint32_t lds[];
    const uint32_t lane = lane_id();
    const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t K = a.K;
    uint32_t *peq = lds + (size_t)wib * (idb_count1_lds_per_wave(K) / sizeof(uint32_t));
    EdgeStage est{(uint2 *)(peq + K * 64), 0u, IDB_NONE, 0u, 0u};
    uint16_t *queue = (uint16_t *)(peq + K * 64 + 2 * IDB_TILE);
    int32_t *qacc = (int32_t *)(peq + K * 64 + 3 * IDB_TILE);
    int32_t *tacc = qacc + IDB_TILE;
    uint4 *gq = (uint4 *)(tacc + IDB_TILE);
    uint4 *gqc = gq + IDB_TILE;  // query codes, two uint4 per query
    const uint32_t total_waves = gridDim.x * wpb;
    const uint32_t wave0 = (a.no_xcd_order ? blockIdx.x : xcd_block(blockIdx.x, gridDim.x)) + wib * gridDim.x;
    const int32_t t_const = a.t_const;
    const bool exact_low = a.e_low != nullptr;
    const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
    unsigned long long gated = 0, hits = 0, gletters = 0, mpairs = 0, mletters = 0;
    unsigned long long glane = 0, gllane = 0;  // per-lane gated pairs and their target letters (TCONST)

    const uint32_t T = a.n_tiles;
    uint32_t ti = wave0;
    uint32_t ti1 = step_tile(ti, total_waves, T);
    idb_tile rec1 = ti1 < T ? tile_rec(a, ti1) : idb_tile{};
    TilePre<true> cur{};
    uint4 cur_tlow = zero4, cur_qlow = zero4;
    // a tile's genes below the bucket's (exact gene test), loaded with it
    auto low_load = [&](const idb_tile &t, uint4 &tl, uint4 &ql) {
        tl = ql = zero4;
        if (!exact_low) return;
        if (lane < ((t.shape >> 8) & 0xFFu)) tl = a.e_low[t.t0 + lane];
        if (lane < (t.shape & 0xFFu)) ql = a.e_low[t.q0 + lane];
    };
    if (ti < T) {
        const idb_tile r0 = tile_rec(a, ti);
        tile_load<true>(a, r0, lane, cur);
        low_load(r0, cur_tlow, cur_qlow);
    }
    const bool wide = a.S > 16;
    // the previous tile's counts (row, amount) for queries and targets
    int32_t p_qa = 0, p_ta = 0;
    uint32_t p_qrow = 0, p_trow = 0;
    auto flush_counts = [&]() {
        if (p_qa) atomicAdd(a.cnt + p_qrow, p_qa);
        if (p_ta) atomicAdd(a.cnt + p_trow, p_ta);
        p_qa = p_ta = 0;
    };
    while (ti < T) {
        // The tile's loads (issued one tile ago) are waited for here, once,
        // and the next tile's are issued only after everything of this tile
        // is in registers or LDS: the wait counter is in order, so a first
        // use after the next loads would wait for those too.  The count
        // atomics of a tile go out after the next tile's loads, so that they
        // do not hold up this wait.
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        const idb_tile tile = cur.tile;
        const uint32_t qn = tile.shape & 0xFFu, tn = (tile.shape >> 8) & 0xFFu;
        const bool diag = (tile.shape >> 16) & 1u;
        const bool tvalid = lane < tn;
        const idb_entry te = entry_of(cur.te0, cur.te1), qe = entry_of(cur.qe0, cur.qe1);
        const uint32_t y = te.row, m = te.len;  // every field of the record is used: no register of a load in flight is reused
        const bool chk_y = (te.len_chk >> 16) & 1u;
        const int32_t t_m = t_of(te.t2, false);
        // Peq of this lane's target
        for (uint32_t c = 0; c < K; ++c) peq[c * 64 + lane] = 0u;
        if (tvalid) {
            const uint32_t ws[8] = {cur.tc0.x, cur.tc0.y, cur.tc0.z, cur.tc0.w,
                                    cur.tc1.x, cur.tc1.y, cur.tc1.z, cur.tc1.w};
#pragma unroll
            for (uint32_t i = 0; i < 32; ++i)
                if (i < m) atomicOr(peq + ((ws[i >> 2] >> (8 * (i & 3))) & 0xFFu) * 64 + lane, 1u << i);
        }
        // gate records: length | check flag << 16 | count threshold (int8) << 24,
        // presence planes, gene residues; the queries' codes and low genes
        if (lane < qn) {
            if (TCONST) {
                gq[lane] = cur.qe1;  // length, presence planes, gene residues
            } else {
                const int32_t tq = t_of(qe.t2, false);
                const uint32_t t8 = (uint32_t)(uint8_t)(int8_t)(tq > 127 ? 127 : tq);
                gq[lane] = make_uint4((qe.len_chk & 0x1FFFFu) | (t8 << 24), qe.pres1, qe.pres2, qe.sig);
            }
        }
        gqc[2 * lane] = cur.qc0;
        gqc[2 * lane + 1] = cur.qc1;
        const uint4 tlow = cur_tlow, qlow = cur_qlow;
        // with the exact test, a target takes part only if it has low genes
        const bool chk_yl = chk_y && (!exact_low || (tlow.x | tlow.y | tlow.z | tlow.w) != 0u);
        qacc[lane] = 0;
        tacc[lane] = 0;
        // the next tile's loads
        TilePre<true> nxt{};
        uint4 nxt_tlow = zero4, nxt_qlow = zero4;
        if (ti1 < T) {
            tile_load<true>(a, rec1, lane, nxt);
            low_load(rec1, nxt_tlow, nxt_qlow);
        }
        const uint32_t ti2 = step_tile(ti1, total_waves, T);
        const idb_tile rec2 = ti2 < T ? tile_rec(a, ti2) : idb_tile{};
        flush_counts();
        const uint64_t tmask = __ballot(tvalid);
        const uint32_t y1 = te.pres1, ny1 = ~te.pres1;
        uint32_t tg = 0, tgl = 0, tmc = 0, qlen = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");

        // One query against the tile's targets: the lanes whose pair passes
        // the band, the triangle and the gate (vm), and of those the lanes
        // that also pass the absent-letter bound (returned).
        auto test = [&](uint32_t qi, const uint4 &q, uint64_t &vm) -> uint64_t {
            const uint32_t qx = __builtin_amdgcn_readfirstlane(q.x);
            const uint32_t lx = qx & 0xFFFFu;
            int32_t t;
            if (TCONST) {
                t = t_const;
            } else {
                const int32_t tq = (int32_t)(int8_t)(qx >> 24);
                t = lx >= m ? tq : t_m;
            }
            const int32_t dl = (int32_t)__builtin_amdgcn_sad_u16(lx, m, 0u);
            // absent-letter bound, both directions: every letter of x whose
            // symbol never occurs in y is substituted or deleted, one edit at
            // least per such symbol
            const uint32_t lbx = __builtin_popcount(q.y & ny1);
            const uint32_t lby = __builtin_popcount(y1 & ~q.y);
            const uint64_t lanes = diag ? (tmask & ((~0ull << qi) << 1)) : tmask;
            vm = lanes & __ballot(dl <= t);
            if (((qx >> 16) & 1u) && vm) {
                // the query may share a lower gene with the target: the
                // canonical-bucket test for targets that may too
                if (exact_low) {
                    const uint32_t common = (__builtin_amdgcn_readlane(qlow.x, qi) & tlow.x) |
                                            (__builtin_amdgcn_readlane(qlow.y, qi) & tlow.y) |
                                            (__builtin_amdgcn_readlane(qlow.z, qi) & tlow.z) |
                                            (__builtin_amdgcn_readlane(qlow.w, qi) & tlow.w);
                    vm &= ~__ballot(common != 0u);
                } else {
                    // residues overlap: exact test on the bitsets
                    const bool need = __builtin_amdgcn_inverse_ballot_w64(vm) && chk_y && (q.w & te.sig) != 0u;
                    const uint64_t nm = __ballot(need);
                    if (nm) {
                        const uint32_t x = __builtin_amdgcn_readlane(qe.row, qi);
                        bool ok = true;
                        if (need) ok = pair_in_bucket(a, x, y, tile.bucket);
                        vm &= ~nm | __ballot(ok);
                    }
                }
            }
            const uint32_t nv = (uint32_t)__popcll(vm);
            tg += nv;
            tgl += nv * lx;
            return vm & __ballot((int32_t)(lbx > lby ? lbx : lby) <= t);
        };
        // A full batch (the first IDB_TILE queue entries) is verified, and
        // the entries past it move to the front of the queue.
        auto full_batch = [&]() {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            myers_batch1<false>(a, peq, queue, IDB_TILE, lane, qe, te, 0, 0, est, qacc, tacc, hits, mletters, zero4,
                                zero4, wide, gqc);
            const uint32_t rest = qlen - IDB_TILE;
            const uint16_t moved = lane < rest ? queue[IDB_TILE + lane] : (uint16_t)0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            if (lane < rest) queue[lane] = moved;
            qlen = rest;
        };
        auto enqueue = [&](uint32_t qi, uint64_t pm) {
            if (!pm) return;
            if (__builtin_amdgcn_inverse_ballot_w64(pm)) {
                const uint32_t below =
                    __builtin_amdgcn_mbcnt_hi((uint32_t)(pm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)pm, 0u));
                queue[qlen + below] = (uint16_t)((qi << 8) | lane);
            }
            const uint32_t np = (uint32_t)__popcll(pm);
            qlen += np;
            tmc += np;
            if (qlen >= IDB_TILE) full_batch();
        };
        if (TCONST) {
            // Constant threshold: the band and both letter bounds fold into
            // one maximum per pair, with no branch per query.  Queries run in
            // windows of up to 32, QG at a time; a pair that passes sets
            // bit r - 1 - k of the lane's mask mb (query w0 + k, r queries in
            // the window), and once per window the masks become Myers queue
            // entries.  gq holds the bare lengths; records past qn are
            // sentinels of length 0x7FFF, and targets past tn take length
            // 0xFFFF, so both fail the band.  Gated pairs are counted per
            // lane.  Queries that may share a lower gene with their targets
            // (chkq) are corrected after each window by the exact gene test.
            if (lane >= qn) gq[lane] = make_uint4(0x7FFFu, 0u, 0u, 0u);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            const uint64_t chkq = __ballot(lane < qn && ((qe.len_chk >> 16) & 1u));
            const uint32_t mlane = tvalid ? m : 0xFFFFu;
            const int32_t t = t_const;
            uint32_t gcnt = 0;  // gated pairs of this lane's target in this tile
            constexpr uint32_t QG = 8;
            auto group = [&](auto diag_c, uint32_t q0, uint32_t &mb) {
                constexpr bool DIAG = decltype(diag_c)::value;
#pragma unroll
                for (uint32_t k = 0; k < QG; ++k) {
                    const uint4 q = gq[q0 + k];
                    const uint32_t dl = __builtin_amdgcn_sad_u16(q.x, mlane, 0u);
                    const uint32_t lbx = __builtin_popcount(q.y & ny1);
                    const uint32_t lby = __builtin_popcount(y1 & ~q.y);
                    const uint32_t mx = max(dl, max(lbx, lby));
                    bool band = (int32_t)dl <= t, pass = (int32_t)mx <= t;
                    if (DIAG) {
                        band = band && lane > q0 + k;
                        pass = pass && lane > q0 + k;
                    }
                    gcnt += band ? 1u : 0u;
                    mb = mb + mb + (pass ? 1u : 0u);
                }
            };
            for (uint32_t w0 = 0; w0 < qn; w0 += 32) {
                const uint32_t r = min(32u, (qn - w0 + QG - 1u) & ~(QG - 1u));
                uint32_t mb = 0;
                for (uint32_t g = 0; g < r; g += QG) {
                    if (diag) group(std::true_type{}, w0 + g, mb);
                    else group(std::false_type{}, w0 + g, mb);
                }
                // exact gene test for the window's queries that need it:
                // pairs it rejects leave the gated count and the masks
                for (uint64_t cm = (chkq >> w0) & (r == 32 ? 0xFFFFFFFFull : ((1ull << r) - 1ull)); cm;
                     cm &= cm - 1ull) {
                    const uint32_t k = (uint32_t)__builtin_ctzll(cm), qi = w0 + k;
                    const uint4 q = gq[qi];
                    const uint32_t dl = __builtin_amdgcn_sad_u16(q.x, mlane, 0u);
                    bool ok = true;
                    if (exact_low) {
                        // a common gene below the bucket's: not this pair's bucket
                        const uint32_t common = (__builtin_amdgcn_readlane(qlow.x, qi) & tlow.x) |
                                                (__builtin_amdgcn_readlane(qlow.y, qi) & tlow.y) |
                                                (__builtin_amdgcn_readlane(qlow.z, qi) & tlow.z) |
                                                (__builtin_amdgcn_readlane(qlow.w, qi) & tlow.w);
                        ok = !((int32_t)dl <= t && (!diag || lane > qi) && chk_yl && common != 0u);
                    } else {
                        const bool need = (int32_t)dl <= t && (!diag || lane > qi) && chk_y && (q.w & te.sig) != 0u;
                        if (!__ballot(need)) continue;
                        const uint32_t x = __builtin_amdgcn_readlane(qe.row, qi);
                        if (need) ok = pair_in_bucket(a, x, y, tile.bucket);
                    }
                    if (!ok) {
                        --gcnt;
                        mb &= ~(1u << (r - 1u - k));
                    }
                }
                for (;;) {
                    const uint64_t any = __ballot(mb != 0u);
                    if (!any) break;
                    // the earliest query still set in each lane
                    const uint32_t hb = 31u - (uint32_t)__builtin_clz(mb | 1u);
                    if (mb != 0u) {
                        const uint32_t below =
                            __builtin_amdgcn_mbcnt_hi((uint32_t)(any >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)any, 0u));
                        queue[qlen + below] = (uint16_t)(((w0 + r - 1u - hb) << 8) | lane);
                        mb &= ~(1u << hb);
                    }
                    const uint32_t np = (uint32_t)__popcll(any);
                    qlen += np;
                    tmc += np;
                    if (qlen >= IDB_TILE) full_batch();
                }
            }
            glane += gcnt;
            gllane += (unsigned long long)gcnt * m;
        } else {
            // two queries per iteration, both records read before either is
            // used (gq[qi + 1] is inside the 64 records even past qn)
            for (uint32_t qi = 0; qi < qn; qi += 2) {
                const uint4 q0 = gq[qi], q1 = gq[qi + 1];
                uint64_t v0 = 0, v1 = 0;
                const uint64_t p0 = test(qi, q0, v0);
                const uint64_t p1 = qi + 1 < qn ? test(qi + 1, q1, v1) : 0ull;
                enqueue(qi, p0);
                enqueue(qi + 1, p1);
            }
        }
        if (qlen) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            myers_batch1<false>(a, peq, queue, qlen, lane, qe, te, 0, 0, est, qacc, tacc, hits, mletters, zero4, zero4,
                                wide, gqc);
        }
        gated += tg;
        gletters += tgl;
        mpairs += tmc;
        // one atomic per row and tile, issued during the next tile
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        p_qa = lane < qn ? qacc[lane] : 0;
        p_ta = tvalid ? tacc[lane] : 0;
        p_qrow = qe.row;
        p_trow = y;
        cur = nxt;
        cur_tlow = nxt_tlow;
        cur_qlow = nxt_qlow;
        ti = ti1;
        ti1 = ti2;
        rec1 = rec2;
    }
    flush_counts();
    finish_edges(a, est, lane);
    mletters = wave_sum_u64(mletters);
    if (TCONST) {
        gated = wave_sum_u64(glane);
        gletters = wave_sum_u64(gllane);
    }
    if (lane == 0 && a.work) {
        atomicAdd(work_copy(a) + 0, gated);
        atomicAdd(work_copy(a) + 1, hits);
        atomicAdd(work_copy(a) + 3, gletters);
        atomicAdd(work_copy(a) + 5, mpairs);
        atomicAdd(work_copy(a) + 6, mletters);
    }
}

// Learned filter (rows of at most 32 letters): the tile's 64 x 64 embedding
// dot products come from 16 MFMAs (v_mfma_f32_32x32x16_bf16, 2 x 2 output
// tiles, 4 k-steps over the 64 dimensions); a pair is proposed when its dot
// product reaches min_sim and it passes the same gate as the exact pass (and,
// for alphabets of 8-31 letters, the exact absent-letter bound), and every
// proposal is verified by Myers exactly as there.  Lane l of output tile
// (i, j) holds target j * 32 + l % 32 and queries i * 32 + 8 (r / 4) +
// 4 (l / 32) + r % 4, r = 0 .. 15; query records are staged in LDS for the
// gate, target fields come from the owning lane once per tile.  The target
// Peq columns are built only for tiles that have a proposal.
typedef __bf16 idb_bf16x8 __attribute__((ext_vector_type(8)));
typedef float idb_f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ idb_bf16x8 load_z8(const uint16_t *zb, uint32_t row, uint32_t k, bool valid)
{
    uint4 v = make_uint4(0, 0, 0, 0);
    if (valid) v = *(const uint4 *)(zb + (size_t)row * 64 + k);
    return __builtin_bit_cast(idb_bf16x8, v);
}

template <int W>
__global__ __launch_bounds__(256) void k_pairs_learned(idb_pair_args a, uint32_t wpb)
{
    extern __shared__ uint32_t lds[];
    const uint32_t lane = lane_id();
    const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform, kept in a scalar register
    const uint32_t K = a.K;
    // per wave: Peq, edge stage, queue, count accumulators, query records,
    // then (3-gram bound) the tile's query and target profiles
    const uint32_t QGW = a.qgram_filter ? 2u * IDB_TILE * IDB_QGRAM_WORDS : 0u;
    uint32_t *peq = lds + (size_t)wib * (idb_myers_lds_per_wave(K, W) / sizeof(uint32_t) + QGW);
    uint32_t *qgp = peq + idb_myers_lds_per_wave(K, W) / sizeof(uint32_t);
    uint32_t *tgp = qgp + IDB_TILE * IDB_QGRAM_WORDS;
    EdgeStage est{(uint2 *)(peq + K * W * 64), 0u, IDB_NONE, 0u, 0u};
    uint16_t *queue = (uint16_t *)(peq + K * W * 64 + 2 * IDB_TILE);
    int32_t *qacc = (int32_t *)(peq + K * W * 64 + 3 * IDB_TILE);
    int32_t *tacc = qacc + IDB_TILE;
    idb_entry *qrec = (idb_entry *)(tacc + IDB_TILE);
    const uint32_t total_waves = gridDim.x * wpb;
    const uint32_t wave0 = (a.no_xcd_order ? blockIdx.x : xcd_block(blockIdx.x, gridDim.x)) * wpb + wib;
    unsigned long long gated = 0, hits = 0, gletters = 0, mletters = 0, mpairs = 0;

    for (uint32_t ti = wave0; ti < a.n_tiles; ti = next_tile(ti, total_waves)) {
        const idb_tile tile = tile_rec(a, ti);
        const uint32_t qn = tile.shape & 0xFFu, tn = (tile.shape >> 8) & 0xFFu;
        const bool diag = (tile.shape >> 16) & 1u;
        const bool tvalid = lane < tn;
        const uint32_t p = tile.t0 + (tvalid ? lane : 0u);
        idb_entry te{}, qe{};
        if (tvalid) te = load_entry(a.ent + p);
        if (lane < qn) qe = load_entry(a.ent + tile.q0 + lane);
        qrec[lane] = qe;
        qacc[lane] = 0;
        tacc[lane] = 0;
        if (a.qgram_filter) {
            // profiles of the tile's queries and targets (64 bytes each)
            const uint4 *qs = (const uint4 *)(a.e_qg + (size_t)(tile.q0 + (lane < qn ? lane : 0u)) * IDB_QGRAM_WORDS);
            const uint4 *ts = (const uint4 *)(a.e_qg + (size_t)p * IDB_QGRAM_WORDS);
#pragma unroll
            for (int k = 0; k < (int)IDB_QGRAM_WORDS / 4; ++k) {
                ((uint4 *)(qgp + lane * IDB_QGRAM_WORDS))[k] = qs[k];
                ((uint4 *)(tgp + lane * IDB_QGRAM_WORDS))[k] = ts[k];
            }
        }
        // rows of the 32-row halves this lane feeds into the MFMAs
        const uint32_t h = lane & 31u;
        const uint32_t rq0 = (uint32_t)__shfl((int)qe.row, (int)h, 64), rq1 = (uint32_t)__shfl((int)qe.row, (int)(32 + h), 64);
        const uint32_t rt0 = (uint32_t)__shfl((int)te.row, (int)h, 64), rt1 = (uint32_t)__shfl((int)te.row, (int)(32 + h), 64);
        const bool vq0 = h < qn, vq1 = 32 + h < qn, vt0 = h < tn, vt1 = 32 + h < tn;
        idb_f32x16 sim[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) sim[i][j][r] = 0.f;
#pragma unroll
        for (uint32_t kk = 0; kk < 4; ++kk) {
            const uint32_t k = kk * 16 + (lane >> 5) * 8;
            const idb_bf16x8 qa0 = load_z8(a.zb, rq0, k, vq0), qa1 = load_z8(a.zb, rq1, k, vq1);
            const idb_bf16x8 tb0 = load_z8(a.zb, rt0, k, vt0), tb1 = load_z8(a.zb, rt1, k, vt1);
            sim[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(qa0, tb0, sim[0][0], 0, 0, 0);
            sim[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(qa0, tb1, sim[0][1], 0, 0, 0);
            sim[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(qa1, tb0, sim[1][0], 0, 0, 0);
            sim[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(qa1, tb1, sim[1][1], 0, 0, 0);
        }
        // candidate mask: bit (2 i + j) * 16 + r for output tile (i, j), row r
        unsigned long long cand = 0;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const uint32_t qi = i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                    const uint32_t tl = j * 32 + h;
                    if (sim[i][j][r] >= a.min_sim && qi < qn && tl < tn && (!diag || qi < tl))
                        cand |= 1ull << ((2 * i + j) * 16 + r);
                }
        // target fields of the two target columns this lane holds
        uint32_t tm[2], tt2[2], tsig[2], ty[2], tp1[2];
        bool tchk[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int src = (int)(j * 32 + h);
            const uint32_t lc = (uint32_t)__shfl((int)te.len_chk, src, 64);
            tm[j] = lc & 0xFFFFu;
            tchk[j] = (lc >> 16) & 1u;
            tt2[j] = (uint32_t)__shfl((int)te.t2, src, 64);
            tsig[j] = (uint32_t)__shfl((int)te.sig, src, 64);
            ty[j] = (uint32_t)__shfl((int)te.row, src, 64);
            tp1[j] = (uint32_t)__shfl((int)te.pres1, src, 64);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        bool peq_ready = false;
        uint32_t qlen = 0;
        auto build_peq = [&]() {
            for (uint32_t c = 0; c < K; ++c)
#pragma unroll
                for (int w = 0; w < W; ++w) peq[(c * W + w) * 64 + lane] = 0u;
            if (tvalid) {
                const uint4 *codes = (const uint4 *)(a.e_codes + (size_t)p * a.S);
                const uint32_t m = te.len_chk & 0xFFFFu;
                for (uint32_t i0 = 0; i0 < m; i0 += 16) {
                    const uint4 v = codes[i0 >> 4];
                    const uint32_t ws[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (uint32_t k = 0; k < 16; ++k) {
                        const uint32_t i = i0 + k;
                        if (i < m) peq[(((ws[k >> 2] >> (8 * (k & 3))) & 0xFFu) * W + (i >> 5)) * 64 + lane] |= 1u << (i & 31);
                    }
                }
            }
            peq_ready = true;
        };
        // every lane takes its candidates one per round: gate, the exact
        // absent-letter bound where the alphabet allows it, then the queue
        while (__ballot(cand != 0ull)) {
            bool ok = false, proposed = false;
            uint32_t qi = 0, tl = 0, lx = 0;
            if (cand) {
                const uint32_t b = (uint32_t)__builtin_ctzll(cand);
                cand &= cand - 1ull;
                const uint32_t i = b >> 5, j = (b >> 4) & 1u, r = b & 15u;
                qi = i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                tl = j * 32 + h;
                const uint32_t m = j ? tm[1] : tm[0];
                const idb_entry &q = qrec[qi];
                lx = q.len_chk & 0xFFFFu;
                const int32_t t = lx >= m ? t_of(q.t2, false) : t_of(j ? tt2[1] : tt2[0], false);
                const int32_t dl = lx > m ? (int32_t)(lx - m) : (int32_t)(m - lx);
                ok = dl <= t && gate_pair(a, (q.len_chk >> 16) & 1u, j ? tchk[1] : tchk[0], q.sig,
                                          j ? tsig[1] : tsig[0], q.row, j ? ty[1] : ty[0], tile.bucket);
                proposed = ok;
                if (ok && a.band_filter) {
                    const uint32_t p1 = j ? tp1[1] : tp1[0];
                    const uint32_t lbx = __builtin_popcount(q.pres1 & ~p1);
                    const uint32_t lby = __builtin_popcount(p1 & ~q.pres1);
                    ok = (int32_t)(lbx > lby ? lbx : lby) <= t;
                } else if (ok && a.qgram_filter) {
                    // the exact 3-gram bound, as the exact pass applies it
                    const uint4 *qv = (const uint4 *)(qgp + qi * IDB_QGRAM_WORDS);
                    const uint4 *tv = (const uint4 *)(tgp + tl * IDB_QGRAM_WORDS);
                    uint32_t l1 = 0;
#pragma unroll
                    for (int k = 0; k < (int)IDB_QGRAM_WORDS / 4; ++k) {
                        const uint4 x4 = qv[k], y4 = tv[k];
                        l1 = __builtin_amdgcn_sad_u8(x4.x, y4.x, l1);
                        l1 = __builtin_amdgcn_sad_u8(x4.y, y4.y, l1);
                        l1 = __builtin_amdgcn_sad_u8(x4.z, y4.z, l1);
                        l1 = __builtin_amdgcn_sad_u8(x4.w, y4.w, l1);
                    }
                    ok = (int32_t)l1 <= 6 * t;
                }
            }
            gated += __popcll(__ballot(proposed));
            gletters += proposed ? lx : 0u;
            const unsigned long long pm = __ballot(ok);
            if (!pm) continue;
            mpairs += __popcll(pm);
            const uint32_t below =
                __builtin_amdgcn_mbcnt_hi((uint32_t)(pm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)pm, 0u));
            if (ok) queue[qlen + below] = (uint16_t)((qi << 8) | tl);
            qlen += (uint32_t)__popcll(pm);
            if (qlen >= IDB_TILE) {
                if (!peq_ready) build_peq();
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                myers_batch_any<W, false>(a, tile, peq, queue, IDB_TILE, lane, qe, te, 0, 0, est, qacc, tacc, hits,
                                          mletters);
                const uint32_t rest = qlen - IDB_TILE;
                const uint16_t moved = lane < rest ? queue[IDB_TILE + lane] : (uint16_t)0;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                if (lane < rest) queue[lane] = moved;
                qlen = rest;
            }
        }
        if (qlen) {
            if (!peq_ready) build_peq();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            myers_batch_any<W, false>(a, tile, peq, queue, qlen, lane, qe, te, 0, 0, est, qacc, tacc, hits, mletters);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        const int32_t qa = qacc[lane], ta = tacc[lane];
        if (lane < qn && qa) atomicAdd(a.cnt + qe.row, qa);
        if (tvalid && ta) atomicAdd(a.cnt + te.row, ta);
    }
    finish_edges(a, est, lane);
    gletters = wave_sum_u64(gletters);
    mletters = wave_sum_u64(mletters);
    if (lane == 0 && a.work) {
        atomicAdd(work_copy(a) + 0, gated);
        atomicAdd(work_copy(a) + 1, hits);
        atomicAdd(work_copy(a) + 3, gletters);
        atomicAdd(work_copy(a) + 5, mpairs);
        atomicAdd(work_copy(a) + 6, mletters);
    }
}

// Exact distance of arbitrary row pairs over raw bytes (labelling): one pair
// per lane, pattern y in W words, match masks built per text letter from y's
// bytes.  Hamming: mismatches at equal length, UINT32_MAX otherwise.
template <int W>
__global__ void k_pair_dist(const uint8_t *seq, const uint64_t *off, const uint32_t *len, const uint2 *pairs,
                            unsigned long long n, uint32_t ham, uint32_t *dist)
{
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const uint2 pr = pairs[i];
        const uint8_t *x = seq + off[pr.x], *y = seq + off[pr.y];
        const uint32_t lx = len[pr.x], m = len[pr.y];
        if (ham) {
            uint32_t d = 0xFFFFFFFFu;
            if (lx == m) {
                d = 0;
                for (uint32_t k = 0; k < m; ++k) d += x[k] != y[k];
            }
            dist[i] = d;
            continue;
        }
        if (m == 0) {
            dist[i] = lx;
            continue;
        }
        idb_myers<W> st;
        idb_myers_init<W>(st, m);
        const uint32_t top_w = (m - 1) >> 5, top_b = (m - 1) & 31;
        for (uint32_t j = 0; j < lx; ++j) {
            const uint8_t c = x[j];
            uint32_t eq[W];
#pragma unroll
            for (int w = 0; w < W; ++w) eq[w] = 0u;
            for (uint32_t k = 0; k < m; ++k)
                if (y[k] == c) eq[k >> 5] |= 1u << (k & 31);
            idb_myers_step<W>(st, eq, top_w, top_b);
        }
        dist[i] = (uint32_t)st.score;
    }
}

// Hamming mode: equal lengths only; packed words (2-bit codes for <= 4
// letters, else 8-bit) compared by XOR, one bit per mismatching field and a
// bit count.  Query words are preloaded per lane when they fit registers.
template <int HWMAX, int BITS, bool FUSE>
__global__ __launch_bounds__(256) void k_pairs_hamming(idb_pair_args a, uint32_t wpb)
{
    const uint32_t lane = lane_id();
    const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform, kept in a scalar register
    const uint32_t total_waves = gridDim.x * wpb;
    const uint32_t wave0 = (a.no_xcd_order ? blockIdx.x : xcd_block(blockIdx.x, gridDim.x)) * wpb + wib;
    const uint32_t HW = a.HW;
    __shared__ uint2 stage[4][IDB_TILE];
    __shared__ int32_t qacc_s[4][IDB_TILE];
    EdgeStage est{stage[wib], 0u, IDB_NONE, 0u, 0u};
    int32_t *qacc = qacc_s[wib];
    if (FUSE && fuse_skipped(a)) return;
    constexpr bool QREG = HWMAX <= 8;
    unsigned long long gated = 0, hits = 0, gletters = 0;

    for (uint32_t ti = wave0; ti < a.n_tiles; ti = next_tile(ti, total_waves)) {
        const idb_tile tile = tile_rec(a, ti);
        const uint32_t qn = tile.shape & 0xFFu, tn = (tile.shape >> 8) & 0xFFu;
        const bool diag = (tile.shape >> 16) & 1u;
        const bool tvalid = lane < tn;
        const uint32_t p = tile.t0 + (tvalid ? lane : 0u);
        idb_entry te{}, qe{};
        if (tvalid) te = load_entry(a.ent + p);
        if (lane < qn) qe = load_entry(a.ent + tile.q0 + lane);
        uint32_t tw[HWMAX], qw[QREG ? HWMAX : 1];
#pragma unroll
        for (int w = 0; w < HWMAX; ++w) tw[w] = 0u;
#pragma unroll
        for (int w = 0; w < (QREG ? HWMAX : 1); ++w) qw[w] = 0u;
        if (tvalid) {
            const uint32_t *pk = a.e_pk + (size_t)p * HW;
#pragma unroll
            for (int w = 0; w < HWMAX; ++w)
                if ((uint32_t)w < HW) tw[w] = pk[w];
        }
        if (QREG && lane < qn) {
            const uint32_t *pk = a.e_pk + (size_t)(tile.q0 + lane) * HW;
#pragma unroll
            for (int w = 0; w < (QREG ? HWMAX : 1); ++w)
                if ((uint32_t)w < HW) qw[w] = pk[w];
        }
        const uint32_t y = te.row, m = te.len_chk & 0xFFFFu, cy = te.cp;
        const bool chk_y = (te.len_chk >> 16) & 1u;
        int32_t core_y = 0, q_core = 0;
        if (FUSE) {
            if (tvalid) core_y = a.core[y];
            if (lane < qn) q_core = a.core[qe.row];
        }
        int32_t acc_y = 0;
        if (!FUSE) qacc[lane] = 0;
        for (uint32_t qi = 0; qi < qn; ++qi) {
            const uint32_t x = rl(qe.row, qi);
            const uint32_t lc = rl(qe.len_chk, qi);
            const uint32_t lx = lc & 0xFFFFu;
            const bool chk_x = (lc >> 16) & 1u;
            const uint32_t qt2 = rl(qe.t2, qi);
            bool valid = tvalid && (!diag || lane > qi) && lx == m;
            int32_t core_x = 0;
            if (FUSE) {
                core_x = (int32_t)rl((uint32_t)q_core, qi);
                valid = valid && (core_x | core_y);
            }
            if (valid) valid = gate_pair(a, chk_x, chk_y, rl(qe.sig, qi), te.sig, x, y, tile.bucket);
            const unsigned long long vm = __ballot(valid);
            if (!vm) continue;
            gated += __popcll(vm);
            gletters += (unsigned long long)__popcll(vm) * lx;
            constexpr uint32_t PER = 32 / BITS;  // letters per word
            const uint32_t nw = (lx + PER - 1) / PER;
            const uint32_t *xp = a.e_pk + (size_t)(tile.q0 + qi) * HW;
            uint32_t ham = 0;
#pragma unroll
            for (int w = 0; w < HWMAX; ++w) {
                if ((uint32_t)w < nw) {
                    const uint32_t xw = QREG ? rl(qw[QREG ? w : 0], qi) : xp[w];
                    ham += BITS == 2 ? idb_mismatch2(xw, tw[w])
                                     : (BITS == 5 ? idb_mismatch5(xw, tw[w]) : idb_mismatch8(xw, tw[w]));
                }
            }
            // count pass: the threshold is compared only after a mismatch
            // (PSDBSCAN.hcu:623-636), fuse pass after every letter (:709-712);
            // the per-length thresholds already encode both, and an empty row
            // passes either way
            const int32_t t = t_of(qt2, FUSE);
            const bool hit = valid && ((int32_t)ham <= t || lx == 0);
            const unsigned long long hm = __ballot(hit);
            if (!hm) continue;
            if (!FUSE) {
                hits += __popcll(hm);
                if (hit) {
                    acc_y += (int32_t)rl(qe.cp, qi);
                    atomicAdd(qacc + qi, (int32_t)cy);
                }
                const int32_t tf = t_of(qt2, true);
                stage_edges(a, est, hit && ((int32_t)ham <= tf || lx == 0), x, y, lane);
            } else if (hit) {
                if (core_x && core_y) uf_union(a.par, x, y);
                else if (core_x) atomicMin(a.bmin + y, x);
                else atomicMin(a.bmin + x, y);
            }
        }
        if (!FUSE) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            const int32_t qa = qacc[lane];
            if (lane < qn && qa) atomicAdd(a.cnt + qe.row, qa);
            if (acc_y) atomicAdd(a.cnt + y, acc_y);
        }
    }
    if (!FUSE) finish_edges(a, est, lane);
    if (lane == 0 && a.work) {
        atomicAdd(work_copy(a) + (FUSE ? 2 : 0), gated);
        if (!FUSE) {
            atomicAdd(work_copy(a) + 1, hits);
            atomicAdd(work_copy(a) + 3, gletters);
        }
    }
}

// The count pass over Hamming groups with the exact lower-gene words
// (canonical buckets, V and J words fitting in four).  A Hamming group is a
// bucket's run of one length, so every pair of a tile has one length and one
// threshold, both wave-uniform.  Queries run in windows of up to 32: each
// lane tests its target against every query of the window with no branch,
// the query's packed words broadcast from LDS, and a pair within the
// threshold sets bit r - 1 - k of the lane's mask (query w0 + k, r queries
// in the window); the exact gene test clears the bits of pairs outside their
// canonical bucket, and the bits then become count updates and staged
// neighbour pairs.  The gated pairs of a lane follow from the tile's shape:
// every query (the queries below the lane on a diagonal tile), less the gene
// test's rejects.  Replaces, for these runs, the query-at-a-time loop of
// k_pairs_hamming (per query: lane reads of the query's fields and words, the
// gate call and two ballots).
template <int HWMAX, int BITS>
__global__ __launch_bounds__(256) void k_count_ham(idb_pair_args a, uint32_t wpb)
{
    const uint32_t lane = lane_id();
    const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t total_waves = gridDim.x * wpb;
    const uint32_t wave0 = (a.no_xcd_order ? blockIdx.x : xcd_block(blockIdx.x, gridDim.x)) * wpb + wib;
    const uint32_t HW = a.HW;
    constexpr uint32_t QS = (HWMAX + 3) & ~3;  // words per query slot in LDS (whole uint4s)
    __shared__ uint2 stage[4][IDB_TILE];
    __shared__ int32_t qacc_s[4][IDB_TILE];
    __shared__ uint2 qinfo_s[4][IDB_TILE];              // query row, copies
    __shared__ uint4 qw_s[4][IDB_TILE * QS / 4];        // query words, QS per query
    EdgeStage est{stage[wib], 0u, IDB_NONE, 0u, 0u};
    int32_t *qacc = qacc_s[wib];
    uint2 *qinfo = qinfo_s[wib];
    uint4 *qw = qw_s[wib];
    unsigned long long glane = 0, gllane = 0, hlane = 0;

    for (uint32_t ti = wave0; ti < a.n_tiles; ti = next_tile(ti, total_waves)) {
        const idb_tile tile = tile_rec(a, ti);
        const uint32_t qn = tile.shape & 0xFFu, tn = (tile.shape >> 8) & 0xFFu;
        const bool diag = (tile.shape >> 16) & 1u;
        const bool tvalid = lane < tn;
        const uint32_t p = tile.t0 + (tvalid ? lane : 0u);
        idb_entry te{}, qe{};
        if (tvalid) te = load_entry(a.ent + p);
        if (lane < qn) qe = load_entry(a.ent + tile.q0 + lane);
        uint32_t tw[HWMAX];
#pragma unroll
        for (int w = 0; w < HWMAX; ++w) tw[w] = 0u;
        if (tvalid) {
            const uint32_t *pk = a.e_pk + (size_t)p * HW;
#pragma unroll
            for (int w = 0; w < HWMAX; ++w)
                if ((uint32_t)w < HW) tw[w] = pk[w];
        }
        {
            // words past a row's end are zero in every packed row, so a
            // pair's sum runs over all HWMAX words with no length test
            uint32_t v[QS];
#pragma unroll
            for (int w = 0; w < (int)QS; ++w) v[w] = 0u;
            if (lane < qn) {
                const uint32_t *pk = a.e_pk + (size_t)(tile.q0 + lane) * HW;
#pragma unroll
                for (int w = 0; w < HWMAX; ++w)
                    if ((uint32_t)w < HW) v[w] = pk[w];
            }
#pragma unroll
            for (int w = 0; w < (int)QS / 4; ++w)
                qw[lane * (QS / 4) + w] = make_uint4(v[4 * w], v[4 * w + 1], v[4 * w + 2], v[4 * w + 3]);
        }
        uint4 tlow = make_uint4(0u, 0u, 0u, 0u), qlow = tlow;
        if (tvalid) tlow = a.e_low[p];
        if (lane < qn) qlow = a.e_low[tile.q0 + lane];
        qinfo[lane] = make_uint2(qe.row, qe.cp);
        qacc[lane] = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        const uint32_t y = te.row, cy = te.cp;
        const bool chk_yl = ((te.len_chk >> 16) & 1u) && (tlow.x | tlow.y | tlow.z | tlow.w) != 0u;
        const uint64_t chkq = __ballot(lane < qn && ((qe.len_chk >> 16) & 1u));
        // one length and one threshold per tile (lane 0 holds a target); an
        // empty row matches every other one
        const uint32_t L = __builtin_amdgcn_readfirstlane(te.len_chk & 0xFFFFu);
        const uint32_t t2 = __builtin_amdgcn_readfirstlane(te.t2);
        const int32_t tc = L ? t_of(t2, false) : 0, tf = L ? t_of(t2, true) : 0;
        // neighbour pairs need the fuse threshold, which is below the count
        // threshold only when it is negative (PSDBSCAN.hcu:623-636, 709-712)
        const bool edges_ok = tf >= tc;
        uint32_t gcnt = tvalid ? (diag ? min(lane, qn) : qn) : 0u;
        int32_t acc_y = 0;
        uint32_t nh = 0;
        for (uint32_t w0 = 0; w0 < qn; w0 += 32) {
            // r: the window's queries rounded up to whole groups of four
            // (slots past qn fail the test below)
            const uint32_t r = min(32u, (qn - w0 + 3u) & ~3u);
            uint32_t mb = 0;
            for (uint32_t g = 0; g < r; g += 4) {
                uint4 qv[4][QS / 4];
#pragma unroll
                for (int k = 0; k < 4; ++k)
#pragma unroll
                    for (int w = 0; w < (int)QS / 4; ++w) qv[k][w] = qw[(w0 + g + k) * (QS / 4) + w];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t qi = w0 + g + k;
                    uint32_t ham = 0;
#pragma unroll
                    for (int w = 0; w < HWMAX; ++w) {
                        const uint4 q4 = qv[k][w / 4];
                        const uint32_t xw = (w & 3) == 0 ? q4.x : ((w & 3) == 1 ? q4.y : ((w & 3) == 2 ? q4.z : q4.w));
                        ham += BITS == 2 ? idb_mismatch2(xw, tw[w])
                                         : (BITS == 5 ? idb_mismatch5(xw, tw[w]) : idb_mismatch8(xw, tw[w]));
                    }
                    const bool pass = tvalid && qi < qn && (int32_t)ham <= tc && (!diag || lane > qi);
                    mb = mb + mb + (pass ? 1u : 0u);
                }
            }
            // the exact gene test for the window's queries that need it
            for (uint64_t cm = (chkq >> w0) & (r == 32 ? 0xFFFFFFFFull : ((1ull << r) - 1ull)); cm;
                 cm &= cm - 1ull) {
                const uint32_t k = (uint32_t)__builtin_ctzll(cm), qi = w0 + k;
                const uint32_t common = (__builtin_amdgcn_readlane(qlow.x, qi) & tlow.x) |
                                        (__builtin_amdgcn_readlane(qlow.y, qi) & tlow.y) |
                                        (__builtin_amdgcn_readlane(qlow.z, qi) & tlow.z) |
                                        (__builtin_amdgcn_readlane(qlow.w, qi) & tlow.w);
                if (tvalid && (!diag || lane > qi) && chk_yl && common != 0u) {
                    --gcnt;
                    mb &= ~(1u << (r - 1u - k));
                }
            }
            // hits: counts of both rows, the pair staged for cluster growth
            for (;;) {
                const uint64_t any = __ballot(mb != 0u);
                if (!any) break;
                const bool has = mb != 0u;
                const uint32_t hb = 31u - (uint32_t)__builtin_clz(mb | 1u);
                uint32_t x = 0;
                if (has) {
                    const uint32_t qi = w0 + r - 1u - hb;
                    const uint2 qinf = qinfo[qi];
                    x = qinf.x;
                    acc_y += (int32_t)qinf.y;
                    atomicAdd(qacc + qi, (int32_t)cy);
                    mb &= ~(1u << hb);
                    ++nh;
                }
                stage_edges(a, est, has && edges_ok, x, y, lane);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        const int32_t qa = qacc[lane];
        if (lane < qn && qa) atomicAdd(a.cnt + qe.row, qa);
        if (acc_y) atomicAdd(a.cnt + y, acc_y);
        glane += gcnt;
        gllane += (unsigned long long)gcnt * L;
        hlane += nh;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // LDS reuse by the next tile
    }
    finish_edges(a, est, lane);
    const unsigned long long gated = wave_sum_u64(glane), gletters = wave_sum_u64(gllane), hits = wave_sum_u64(hlane);
    if (lane == 0 && a.work) {
        atomicAdd(work_copy(a) + 0, gated);
        atomicAdd(work_copy(a) + 1, hits);
        atomicAdd(work_copy(a) + 3, gletters);
    }
}

// Cluster growth over the recorded pairs, in entry space.  The union-find
// over rows touches two scattered parents per pair (rows are in input order),
// and its speed is set by how many scattered cache lines the waves have in
// flight.  A pair's two entries lie in one 64 x 64 tile, so the pairs of a
// wave touch a few lines of an entry-indexed forest instead.  Four steps,
// each returning at once when the pair list overflowed (the fuse pass over
// the tiles grows the clusters then):
//   1. pairs to entries: a pair of core rows becomes the pair of the entries
//      that stand for them (pe, with the core flag in bit 31) and links the larger entry
//      under the smaller with atomicMin, which never fails: each entry ends up
//      under its smallest core neighbour, an acyclic forest of true links, so
//      the exchanges of step 2 meet far fewer contenders.  A core row with a
//      non-core neighbour claims it (R10: the smallest core row wins).  Other
//      pairs are marked dead.
//   2. every entry pointed at its root, then the union over the entry pairs
//      (exchange on roots, roots are minimum entries); a row is represented by
//      one entry only, so rows with several entries need no extra links;
//   3. the smallest row of each component, collected at its root entry;
//   4. each core row's parent becomes that row, so over rows the roots are
//      the component minima, as the renumbering expects.
// Parents are read through the caches and never rewritten except by the
// linking exchange: a stale parent only points to an earlier ancestor
// (parents only decrease), which ends in a failed exchange that returns the
// current value, so every retry still makes progress.
__device__ __forceinline__ uint32_t ld_cached(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ uint32_t find_cached(const uint32_t *par, uint32_t x)
{
    uint32_t p = ld_cached(par + x);
    while (p != x) {
        x = p;
        p = ld_cached(par + x);
    }
    return x;
}

// Both roots at once: the two walks' loads are in flight together, so a
// union waits for the longer walk, not for the sum of both.
__device__ __forceinline__ void find_both(const uint32_t *par, uint32_t &a, uint32_t &b)
{
    for (;;) {
        const uint32_t pa = ld_cached(par + a), pb = ld_cached(par + b);
        if (pa == a && pb == b) return;
        a = pa;
        b = pb;
    }
}

__device__ void union_cached(uint32_t *par, uint32_t a, uint32_t b)
{
    find_both(par, a, b);
    while (a != b) {
        if (a < b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        uint32_t expected = a;
        if (__hip_atomic_compare_exchange_strong(par + a, &expected, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
        a = expected;
        find_both(par, a, b);
    }
}

__device__ __forceinline__ bool pairs_overflowed(const uint32_t *n_chunks, uint32_t max_chunks)
{
    return *n_chunks > max_chunks;
}

constexpr uint32_t PE_CORE = 0x80000000u;

__global__ void k_union_to_entries(uint2 *edges, const uint32_t *fill, const uint32_t *n_chunks, uint32_t max_chunks,
                                   uint32_t chunk_log2, const uint32_t *pe, uint32_t *par_e, uint32_t *bmin)
{
    const uint32_t nc = *n_chunks;
    if (nc > max_chunks) return;
    const unsigned long long n = (unsigned long long)nc << chunk_log2;
    const uint32_t cmask = (1u << chunk_log2) - 1u;
    for (unsigned long long e = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; e < n;
         e += (unsigned long long)gridDim.x * blockDim.x) {
        if (((uint32_t)e & cmask) >= fill[e >> chunk_log2]) continue;
        const uint2 p = edges[e];
        const uint32_t px = pe[p.x], py = pe[p.y];
        const bool cx = px != IDB_NONE && (px & PE_CORE) != 0u, cy = py != IDB_NONE && (py & PE_CORE) != 0u;
        uint2 q = make_uint2(IDB_NONE, IDB_NONE);
        if (cx && cy) {
            q = make_uint2(px & ~PE_CORE, py & ~PE_CORE);
            const uint32_t lo = q.x < q.y ? q.x : q.y, hi = q.x < q.y ? q.y : q.x;
            if (lo != hi) atomicMin(par_e + hi, lo);
        } else if (cx) {
            atomicMin(bmin + p.y, p.x);
        } else if (cy) {
            atomicMin(bmin + p.x, p.y);
        }
        edges[e] = q;
    }
}

__global__ void k_union_entries(const uint2 *edges, const uint32_t *fill, const uint32_t *n_chunks,
                                uint32_t max_chunks, uint32_t chunk_log2, uint32_t *par_e)
{
    const uint32_t nc = *n_chunks;
    if (nc > max_chunks) return;
    const unsigned long long n = (unsigned long long)nc << chunk_log2;
    const uint32_t cmask = (1u << chunk_log2) - 1u;
    // The grid-stride step is a multiple of 64 and chunks hold a multiple of
    // 64 pairs, so the lanes of a wave always hold consecutive slots of one
    // chunk.  Consecutive pairs mostly share their first entry x (one query's
    // hits); a run x-y1, x-y2, ... is joined as the chain x-y1, y1-y2, ...:
    // the same component, without every lane queueing its exchange on the
    // root of x.
    const uint32_t lane = __lane_id();
    for (unsigned long long base = (blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x) & ~63ull; base < n;
         base += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned long long e = base + lane;
        const bool live = ((uint32_t)e & cmask) < fill[e >> chunk_log2];
        uint2 p = make_uint2(IDB_NONE, IDB_NONE);
        if (live) p = edges[e];
        const bool cc = p.x != IDB_NONE;
        // previous lane's pair, read with the whole wave active
        const uint32_t px = (uint32_t)__shfl_up((int)p.x, 1, 64), py = (uint32_t)__shfl_up((int)p.y, 1, 64);
        const bool pcc = __shfl_up((int)cc, 1, 64) != 0 && lane > 0;
        if (cc) union_cached(par_e, (pcc && px == p.x) ? py : p.x, p.y);
    }
}

// Between the pre-link and the union: every entry points straight at its
// tree's root, so the union's finds mostly take one step.  A concurrent find
// reads the old parent or the root, both ancestors.
__global__ void k_union_compress(const uint32_t *n_chunks, uint32_t max_chunks, uint32_t *par_e, uint32_t n_entries)
{
    if (pairs_overflowed(n_chunks, max_chunks)) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_entries; i += gridDim.x * blockDim.x) {
        const uint32_t r = find_cached(par_e, i);
        if (r != i) __hip_atomic_store(par_e + i, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Depth histogram of the union-find forest as cluster growth left it, before
// the roots are collected (the reference's tree_depth_evaluation_kernel,
// cuda/PSDBSCAN.hcu:154-177, run after grow_clusters at :1322-1325): a
// duplicate row is one step from its representative, a border row one step
// from the core row that claimed it, and a core row as many steps as its
// entry is below the root of the entry forest.  The depths depend on the
// linking order, as the reference's do.
__global__ void k_union_depth(const uint32_t *n_chunks, uint32_t max_chunks, const uint32_t *rep, const uint32_t *pe,
                              const uint32_t *bmin, const uint32_t *par_e, uint32_t n_rows, uint32_t *hist)
{
    if (pairs_overflowed(n_chunks, max_chunks)) return;
    __shared__ uint32_t h[51];
    for (uint32_t k = threadIdx.x; k < 51; k += blockDim.x) h[k] = 0u;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_rows; i += gridDim.x * blockDim.x) {
        uint32_t d = 0, x = i;
        if (rep[x] != x) {
            x = rep[x];
            ++d;
        }
        uint32_t f = pe[x];
        if ((f == IDB_NONE || !(f & PE_CORE)) && bmin[x] != IDB_NONE) {
            x = bmin[x];
            ++d;
            f = pe[x];
        }
        if (f != IDB_NONE && (f & PE_CORE)) {
            uint32_t e = f & ~PE_CORE;
            for (uint32_t p = ld_cached(par_e + e); p != e && d < 50; p = ld_cached(par_e + e)) {
                e = p;
                ++d;
            }
        }
        atomicAdd(&h[d > 50 ? 50 : d], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < 51; k += blockDim.x)
        if (h[k]) atomicAdd(hist + k, h[k]);
}

__global__ void k_union_comp_min(const uint32_t *n_chunks, uint32_t max_chunks, const uint32_t *pe,
                                 const uint32_t *par_e, uint32_t *cmin_e, uint32_t n_rows)
{
    if (pairs_overflowed(n_chunks, max_chunks)) return;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n_rows; u += gridDim.x * blockDim.x) {
        const uint32_t f = pe[u];
        if (f != IDB_NONE && (f & PE_CORE)) atomicMin(cmin_e + find_cached(par_e, f & ~PE_CORE), u);
    }
}

__global__ void k_union_to_rows(const uint32_t *n_chunks, uint32_t max_chunks, const uint32_t *pe,
                                const uint32_t *par_e, const uint32_t *cmin_e, uint32_t n_rows, uint32_t *par)
{
    if (pairs_overflowed(n_chunks, max_chunks)) return;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n_rows; u += gridDim.x * blockDim.x) {
        const uint32_t f = pe[u];
        if (f != IDB_NONE && (f & PE_CORE)) par[u] = cmin_e[find_cached(par_e, f & ~PE_CORE)];
    }
}

// Blocks that fit on the device at once for this kernel and shape.
// Blocks of `kernel` resident on the whole device, remembered per (device,
// kernel, block size, LDS bytes): the occupancy query is a runtime call on
// every launch otherwise.
uint32_t resident_blocks(const void *kernel, uint32_t block, size_t shm)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        dev = 0;
    }
    static std::mutex m;
    static std::map<std::tuple<int, const void *, uint32_t, size_t>, uint32_t> known;
    const auto key = std::make_tuple(dev, kernel, block, shm);
    {
        std::lock_guard<std::mutex> g(m);
        const auto it = known.find(key);
        if (it != known.end()) return it->second;
    }
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, (int)block, shm) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    (void)hipGetLastError();
    const uint32_t r = (uint32_t)per_cu * (uint32_t)cus;
    std::lock_guard<std::mutex> g(m);
    known[key] = r;
    return r;
}

uint32_t grid_for(uint32_t n_tiles, uint32_t wpb, uint32_t resident, uint32_t mul = 1)
{
    uint32_t b = (n_tiles + wpb - 1) / wpb;
    const uint32_t cap = resident * (mul ? mul : 1u);
    if (b > cap) b = cap;
    return b ? b : 1u;
}

// Waves per block: the tuned value when set, within [1, most].
uint32_t pick_wpb(const idb_pair_args &a, uint32_t most)
{
    uint32_t w = a.tune_wpb ? a.tune_wpb : most;
    return std::max(1u, std::min(w, most));
}

template <int W, bool FUSE, int FILT>
hipError_t launch_myers_w(const idb_pair_args &a, dim3 block, size_t shm, uint32_t wpb, hipStream_t s)
{
    const void *k = (const void *)k_pairs_myers<W, FUSE, FILT>;
    if (shm > 65536) {
        const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
    }
    const dim3 grid(grid_for(a.n_tiles, wpb, resident_blocks(k, block.x, shm), a.tune_grid_mul));
    hipLaunchKernelGGL((k_pairs_myers<W, FUSE, FILT>), grid, block, shm, s, a, wpb);
    return hipGetLastError();
}

template <bool FUSE, int FILT>
hipError_t launch_myers_f(const idb_pair_args &a, uint32_t W, dim3 block, size_t shm, uint32_t wpb, hipStream_t s)
{
    switch (W) {
    case 1: return launch_myers_w<1, FUSE, FILT>(a, block, shm, wpb, s);
    case 2: return launch_myers_w<2, FUSE, FILT>(a, block, shm, wpb, s);
    case 3: return launch_myers_w<3, FUSE, FILT>(a, block, shm, wpb, s);
    case 4: return launch_myers_w<4, FUSE, FILT>(a, block, shm, wpb, s);
    case 5: return launch_myers_w<5, FUSE, FILT>(a, block, shm, wpb, s);
    case 6: return launch_myers_w<6, FUSE, FILT>(a, block, shm, wpb, s);
    case 7: return launch_myers_w<7, FUSE, FILT>(a, block, shm, wpb, s);
    case 8: return launch_myers_w<8, FUSE, FILT>(a, block, shm, wpb, s);
    default: return hipErrorInvalidValue;
    }
}

template <bool TCONST>
hipError_t launch_count1_t(const idb_pair_args &a, uint32_t wpb, hipStream_t s)
{
    const size_t shm = idb_count1_lds_per_wave(a.K) * wpb;
    const void *k = (const void *)k_count1<TCONST>;
    if (shm > 65536) {
        const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
    }
    const dim3 block(64 * wpb), grid(grid_for(a.n_tiles, wpb, resident_blocks(k, block.x, shm), a.tune_grid_mul));
    hipLaunchKernelGGL(k_count1<TCONST>, grid, block, shm, s, a, wpb);
    return hipGetLastError();
}

// The one-word count pass (k_count1); the caller has filled the launch shape.
hipError_t launch_count1(const idb_pair_args &a, hipStream_t s)
{
    const size_t per_wave = idb_count1_lds_per_wave(a.K);
    uint32_t wpb = (uint32_t)(std::min<size_t>(idb_lds_limit(), 65536) / per_wave);
    if (wpb > 4) wpb = 4;
    if (wpb == 0) return hipErrorInvalidValue;
    wpb = pick_wpb(a, wpb);
    return a.t_const >= 0 ? launch_count1_t<true>(a, wpb, s) : launch_count1_t<false>(a, wpb, s);
}

template <bool FUSE>
hipError_t launch_myers(const idb_pair_args &a_in, uint32_t W, hipStream_t s)
{
    if (a_in.n_tiles == 0) return hipSuccess;
    // Unless a shape was tuned: one wave per block, plain block order, and a
    // grid of 4x (one-word rows) or 8x (longer rows) the resident waves, so
    // that waves done early are replaced while heavy tiles run.  These are
    // the autotuner's picks on 200k protein rows (0.276 against 0.305 ms for
    // one resident round of 4-wave blocks in XCD order) and on 20M
    // nucleotide rows (1.10 against 1.31 s).
    idb_pair_args a = a_in;
    const bool count1 = W == 1 && a.band_filter && !FUSE && !a.single_bucket;
    if (a.tune_wpb == 0 && a.tune_grid_mul == 0) {
        // the one-word count pass (k_count1): 4-wave blocks in XCD order,
        // twice the resident grid (the autotuner's pick on 200k protein
        // rows: 0.230 against 0.263 ms for the shape above)
        a.tune_wpb = count1 ? 4u : 1u;
        a.tune_grid_mul = count1 ? 2u : (W == 1 ? 4u : 8u);
        a.no_xcd_order = count1 ? 0u : 1u;
    }
    const size_t per_wave = idb_myers_lds_per_wave(a.K, W) + (a.qgram_filter ? idb_qgram_lds_per_wave() : 0);
    const size_t limit = idb_lds_limit();
    uint32_t wpb = (uint32_t)(std::min<size_t>(limit, 65536) / per_wave);
    if (wpb > 4) wpb = 4;
    if (wpb == 0) wpb = per_wave <= limit ? 1u : 0u;
    if (wpb == 0) return hipErrorInvalidValue;
    wpb = pick_wpb(a, wpb);
    const dim3 block(64 * wpb);
    const size_t shm = per_wave * wpb;
    if (W == 1 && a.band_filter) {
        if (a.S > 32 || a.K > 32) return hipErrorInvalidValue;  // one-word rows, alphabets of at most 31 letters
        if (count1) return launch_count1(a, s);
        return launch_myers_w<1, FUSE, 1>(a, block, shm, wpb, s);
    }
    if (a.qgram_filter) return launch_myers_f<FUSE, 2>(a, W, block, shm, wpb, s);
    return launch_myers_f<FUSE, 0>(a, W, block, shm, wpb, s);
}

template <int HWMAX, int BITS, bool FUSE>
hipError_t launch_hamming_hw(const idb_pair_args &a_in, hipStream_t s)
{
    // Unless a shape was tuned: one wave per block, 4x the resident grid,
    // plain block order (the autotuner's pick on the 512-repertoire batch).
    idb_pair_args a = a_in;
    if (a.tune_wpb == 0 && a.tune_grid_mul == 0) {
        a.tune_wpb = 1;
        a.tune_grid_mul = 4;
        a.no_xcd_order = 1;
    }
    const uint32_t wpb = pick_wpb(a, 4);
    const void *k = (const void *)k_pairs_hamming<HWMAX, BITS, FUSE>;
    const dim3 block(64 * wpb), grid(grid_for(a.n_tiles, wpb, resident_blocks(k, 64 * wpb, 0), a.tune_grid_mul));
    hipLaunchKernelGGL((k_pairs_hamming<HWMAX, BITS, FUSE>), grid, block, 0, s, a, wpb);
    return hipGetLastError();
}

template <int BITS, bool FUSE>
hipError_t launch_hamming_bits(const idb_pair_args &a, hipStream_t s)
{
    if (a.HW <= 2) return launch_hamming_hw<2, BITS, FUSE>(a, s);
    if (a.HW <= 4) return launch_hamming_hw<4, BITS, FUSE>(a, s);
    if (a.HW <= 8) return launch_hamming_hw<8, BITS, FUSE>(a, s);
    if (a.HW <= 16) return launch_hamming_hw<16, BITS, FUSE>(a, s);
    if (a.HW <= 32) return launch_hamming_hw<32, BITS, FUSE>(a, s);
    if (a.HW <= 64) return launch_hamming_hw<64, BITS, FUSE>(a, s);
    return hipErrorInvalidValue;
}

// The branch-free count pass (k_count_ham) for canonical-bucket runs with
// the exact lower-gene words and rows of at most 8 packed words.
template <int HWMAX, int BITS>
hipError_t launch_count_ham_hw(const idb_pair_args &a_in, hipStream_t s)
{
    // Unless a shape was tuned: 4-wave blocks, 4x the resident grid, plain
    // order (the autotuner's pick on the 512-repertoire batch: 1.31 against
    // 2.07 ms for one-wave blocks)
    idb_pair_args a = a_in;
    if (a.tune_wpb == 0 && a.tune_grid_mul == 0) {
        a.tune_wpb = 4;
        a.tune_grid_mul = 4;
        a.no_xcd_order = 1;
    }
    const uint32_t wpb = pick_wpb(a, 4);
    const void *k = (const void *)k_count_ham<HWMAX, BITS>;
    const dim3 block(64 * wpb), grid(grid_for(a.n_tiles, wpb, resident_blocks(k, 64 * wpb, 0), a.tune_grid_mul));
    hipLaunchKernelGGL((k_count_ham<HWMAX, BITS>), grid, block, 0, s, a, wpb);
    return hipGetLastError();
}

template <int BITS>
hipError_t launch_count_ham(const idb_pair_args &a, hipStream_t s)
{
    // the pair sum runs over all HWMAX words: the row width exactly up to 6
    if (a.HW <= 1) return launch_count_ham_hw<1, BITS>(a, s);
    if (a.HW <= 2) return launch_count_ham_hw<2, BITS>(a, s);
    if (a.HW <= 3) return launch_count_ham_hw<3, BITS>(a, s);
    if (a.HW <= 4) return launch_count_ham_hw<4, BITS>(a, s);
    if (a.HW <= 5) return launch_count_ham_hw<5, BITS>(a, s);
    if (a.HW <= 6) return launch_count_ham_hw<6, BITS>(a, s);
    return launch_count_ham_hw<8, BITS>(a, s);
}

template <bool FUSE>
hipError_t launch_hamming(const idb_pair_args &a, hipStream_t s)
{
    if (a.n_tiles == 0) return hipSuccess;
    if (!FUSE && a.e_low && !a.single_bucket && a.HW <= 8) {
        if (a.bits == 2) return launch_count_ham<2>(a, s);
        if (a.bits == 5) return launch_count_ham<5>(a, s);
        return launch_count_ham<8>(a, s);
    }
    if (a.bits == 2) return launch_hamming_bits<2, FUSE>(a, s);
    if (a.bits == 5) return launch_hamming_bits<5, FUSE>(a, s);
    return launch_hamming_bits<8, FUSE>(a, s);
}

hipError_t launch_learned(const idb_pair_args &a_in, hipStream_t s)
{
    if (a_in.n_tiles == 0) return hipSuccess;
    // the exact one-word pass's default shape, unless one was tuned
    idb_pair_args a = a_in;
    if (a.tune_wpb == 0 && a.tune_grid_mul == 0) {
        a.tune_wpb = 1;
        a.tune_grid_mul = 4;
        a.no_xcd_order = 1;
    }
    // pattern words of the longest row (the encoder reads up to 160 letters)
    const uint32_t W = std::max(1u, (a.S + 31) / 32);
    if (W > 5) return hipErrorInvalidValue;
    const size_t per_wave = idb_myers_lds_per_wave(a.K, W) + (a.qgram_filter ? 2 * idb_qgram_lds_per_wave() : 0);
    uint32_t wpb = (uint32_t)(std::min<size_t>(idb_lds_limit(), 65536) / per_wave);
    if (wpb > 4) wpb = 4;
    if (wpb == 0) return hipErrorInvalidValue;
    wpb = pick_wpb(a, wpb);
    const size_t shm = per_wave * wpb;
    const void *k = W == 1   ? (const void *)k_pairs_learned<1>
                    : W == 2 ? (const void *)k_pairs_learned<2>
                    : W == 3 ? (const void *)k_pairs_learned<3>
                    : W == 4 ? (const void *)k_pairs_learned<4>
                             : (const void *)k_pairs_learned<5>;
    if (shm > 65536) {
        const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
    }
    const dim3 block(64 * wpb), grid(grid_for(a.n_tiles, wpb, resident_blocks(k, block.x, shm), a.tune_grid_mul));
    switch (W) {
    case 1: hipLaunchKernelGGL(k_pairs_learned<1>, grid, block, shm, s, a, wpb); break;
    case 2: hipLaunchKernelGGL(k_pairs_learned<2>, grid, block, shm, s, a, wpb); break;
    case 3: hipLaunchKernelGGL(k_pairs_learned<3>, grid, block, shm, s, a, wpb); break;
    case 4: hipLaunchKernelGGL(k_pairs_learned<4>, grid, block, shm, s, a, wpb); break;
    default: hipLaunchKernelGGL(k_pairs_learned<5>, grid, block, shm, s, a, wpb); break;
    }
    return hipGetLastError();
}

template <int W>
hipError_t launch_pair_dist_w(const uint8_t *seq, const uint64_t *off, const uint32_t *len, const uint2 *pairs,
                              unsigned long long n, uint32_t ham, uint32_t *dist, hipStream_t s)
{
    unsigned long long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_pair_dist<W>, dim3((uint32_t)blocks), dim3(256), 0, s, seq, off, len, pairs, n, ham, dist);
    return hipGetLastError();
}

}  // namespace

hipError_t idb_launch_count_learned(const idb_pair_args &a, hipStream_t s) { return launch_learned(a, s); }

hipError_t idb_launch_pair_distances(const uint8_t *seq, const uint64_t *off, const uint32_t *len, const uint2 *pairs,
                                     unsigned long long n, uint32_t W, uint32_t ham, uint32_t *dist, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    switch (ham ? 1u : W) {
    case 1: return launch_pair_dist_w<1>(seq, off, len, pairs, n, ham, dist, s);
    case 2: return launch_pair_dist_w<2>(seq, off, len, pairs, n, ham, dist, s);
    case 3: return launch_pair_dist_w<3>(seq, off, len, pairs, n, ham, dist, s);
    case 4: return launch_pair_dist_w<4>(seq, off, len, pairs, n, ham, dist, s);
    case 5: return launch_pair_dist_w<5>(seq, off, len, pairs, n, ham, dist, s);
    case 6: return launch_pair_dist_w<6>(seq, off, len, pairs, n, ham, dist, s);
    case 7: return launch_pair_dist_w<7>(seq, off, len, pairs, n, ham, dist, s);
    case 8: return launch_pair_dist_w<8>(seq, off, len, pairs, n, ham, dist, s);
    default: return hipErrorInvalidValue;
    }
}

size_t idb_lds_limit()
{
    // per device, asked once
    static std::atomic<int> known[64];
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        dev = 0;
    }
    if (dev >= 0 && dev < 64 && known[dev].load(std::memory_order_relaxed) > 0)
        return (size_t)known[dev].load(std::memory_order_relaxed);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess || v <= 0) {
        (void)hipGetLastError();
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) {
            (void)hipGetLastError();
            v = 65536;
        }
    }
    if (dev >= 0 && dev < 64) known[dev].store(v, std::memory_order_relaxed);
    return (size_t)v;
}

hipError_t idb_launch_count_myers(const idb_pair_args &a, uint32_t W, hipStream_t s) { return launch_myers<false>(a, W, s); }
hipError_t idb_launch_fuse_myers(const idb_pair_args &a, uint32_t W, hipStream_t s) { return launch_myers<true>(a, W, s); }

hipError_t idb_launch_count_hamming(const idb_pair_args &a, hipStream_t s) { return launch_hamming<false>(a, s); }
hipError_t idb_launch_fuse_hamming(const idb_pair_args &a, hipStream_t s) { return launch_hamming<true>(a, s); }

hipError_t idb_launch_union_edges(const idb_pair_args &a, const idb_union_args &u, hipStream_t s)
{
    if (!a.edges || a.max_chunks == 0) return hipSuccess;
    const uint32_t b1 = resident_blocks((const void *)k_union_to_entries, 256, 0);
    hipLaunchKernelGGL(k_union_to_entries, dim3(b1), dim3(256), 0, s, a.edges, a.chunk_fill, a.n_chunks, a.max_chunks,
                       a.chunk_log2, u.pe, u.par_e, u.bmin);
    const uint32_t eb = (u.n_entries + 255) / 256 < 2048u ? (u.n_entries + 255) / 256 : 2048u;
    hipLaunchKernelGGL(k_union_compress, dim3(eb ? eb : 1u), dim3(256), 0, s, a.n_chunks, a.max_chunks, u.par_e,
                       u.n_entries);
    const uint32_t b2 = resident_blocks((const void *)k_union_entries, 256, 0);
    hipLaunchKernelGGL(k_union_entries, dim3(b2), dim3(256), 0, s, a.edges, a.chunk_fill, a.n_chunks, a.max_chunks,
                       a.chunk_log2, u.par_e);
    const uint32_t rb = (u.n_rows + 255) / 256 < 2048u ? (u.n_rows + 255) / 256 : 2048u;
    if (u.depth_hist)
        hipLaunchKernelGGL(k_union_depth, dim3(rb ? rb : 1u), dim3(256), 0, s, a.n_chunks, a.max_chunks, u.rep, u.pe,
                           u.bmin, u.par_e, u.n_rows, u.depth_hist);
    hipLaunchKernelGGL(k_union_comp_min, dim3(rb ? rb : 1u), dim3(256), 0, s, a.n_chunks, a.max_chunks, u.pe, u.par_e,
                       u.cmin_e, u.n_rows);
    hipLaunchKernelGGL(k_union_to_rows, dim3(rb ? rb : 1u), dim3(256), 0, s, a.n_chunks, a.max_chunks, u.pe, u.par_e,
                       u.cmin_e, u.n_rows, u.par);
    return hipGetLastError();
}
