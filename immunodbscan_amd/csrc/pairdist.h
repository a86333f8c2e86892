// pairdist.h — per-pair distance primitives shared by the device kernels and
// the host-side unit test (compiled there without HIP).
//
// Levenshtein: Myers' bit-vector algorithm (G. Myers, "A fast bit-vector
// algorithm for approximate string matching based on dynamic programming",
// JACM 1999) in the edit-distance form given by H. Hyyrö ("Explaining and
// extending the bit-parallel approximate string matching algorithm of Myers",
// 2001): the pattern's vertical deltas are kept as bit vectors Pv/Mv, one text
// letter is consumed per step, and shifting a 1 into Ph encodes the first DP
// row D[0][j] = j.  Patterns longer than 32 letters use W words with the add
// carry and the Ph/Mh shift carries passed from word to word.
//
// Hamming over packed codes: XOR of two packed words, then one bit per
// mismatching field, then popcount.
#pragma once
#include <stdint.h>

#ifndef IDB_HD
#if defined(__HIPCC__)
#define IDB_HD __host__ __device__ __forceinline__
#else
#define IDB_HD static inline
#endif
#endif

// Myers state for a pattern of up to 32*W letters.
template <int W>
struct idb_myers {
    uint32_t pv[W], mv[W];
    int32_t score;
};

template <int W>
IDB_HD void idb_myers_init(idb_myers<W> &s, uint32_t m)
{
#pragma unroll
    for (int w = 0; w < W; ++w) {
        s.pv[w] = 0xFFFFFFFFu;
        s.mv[w] = 0u;
    }
    s.score = (int32_t)m;
}

// One text letter; eq[w] = positions of that letter in the pattern.
// top_w / top_b locate bit m-1 (the last pattern row).
template <int W>
IDB_HD void idb_myers_step(idb_myers<W> &s, const uint32_t *eq, uint32_t top_w, uint32_t top_b)
{
    uint32_t add_c = 0u, ph_c = 1u, mh_c = 0u;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const uint32_t e = eq[w], pv = s.pv[w], mv = s.mv[w];
        const uint32_t xv = e | mv;
        const uint32_t t1 = e & pv;
        const uint64_t sum64 = (uint64_t)t1 + pv + add_c;
        const uint32_t sum = (uint32_t)sum64;
        add_c = (uint32_t)(sum64 >> 32);
        const uint32_t xh = (sum ^ pv) | e;
        const uint32_t ph = mv | ~(xh | pv);
        const uint32_t mh = pv & xh;
        if ((uint32_t)w == top_w) s.score += (int32_t)((ph >> top_b) & 1u) - (int32_t)((mh >> top_b) & 1u);
        const uint32_t phs = (ph << 1) | ph_c;
        const uint32_t mhs = (mh << 1) | mh_c;
        ph_c = ph >> 31;
        mh_c = mh >> 31;
        s.pv[w] = mhs | ~(xv | phs);
        s.mv[w] = phs & xv;
    }
}

// Mismatching fields of two packed words: 2-bit codes (16 per word) or 8-bit
// codes (4 per word).
IDB_HD uint32_t idb_mismatch2(uint32_t a, uint32_t b)
{
    const uint32_t x = a ^ b;
    return (uint32_t)__builtin_popcount((x | (x >> 1)) & 0x55555555u);
}

// Six 5-bit fields (bits 0..29): adding 0b01111 to a field's low four bits
// carries into its top bit exactly when they are not all zero, and never
// into the next field (15 + 15 < 32); or-ing the field's own top bit makes
// the top bit "the field differs".
IDB_HD uint32_t idb_mismatch5(uint32_t a, uint32_t b)
{
    const uint32_t x = a ^ b;
    const uint32_t y = ((x & 0x1EF7BDEFu) + 0x1EF7BDEFu) | x;
    return (uint32_t)__builtin_popcount(y & 0x21084210u);
}

IDB_HD uint32_t idb_mismatch8(uint32_t a, uint32_t b)
{
    const uint32_t x = a ^ b;
    const uint32_t y = ((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x;
    return (uint32_t)__builtin_popcount(y & 0x80808080u);
}
