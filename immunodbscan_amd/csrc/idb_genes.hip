// R2 gene bitsets on the device: bit g of row i is set iff gene name g occurs
// as a contiguous substring of call string i (cuda/BitsetVector.inl:151-175).
// The reference launches one grid per gene (find_string_occurences,
// cuda/BitsetVector.inl:181-209); here one wave takes one call string and
// tests every gene against it:
//   * the string is staged in the wave's LDS window, 2048 bytes at a time,
//     with start positions advancing 1024 at a time (names longer than 1024
//     bytes never match, as in the reference whose launch of more than 1024
//     threads fails), so every start position sees its whole candidate;
//   * lane p holds the 4 bytes at start position p (v_alignbyte of two LDS
//     words); a name's first 4 bytes are compared with one masked XOR and
//     its last 4 bytes with one more (gene names share their first bytes:
//     every IGHV name starts "IGHV"), and the lanes that pass check the
//     middle of the name 4 bytes at a time;
//   * the gene headers (the first 512) and the name words (when they fit in
//     2048 words) are staged in LDS once per block and read by broadcast; the
//     row's bit words collect in LDS and leave with one store per word.
#include "idb_internal.h"

namespace {

constexpr uint32_t GSEG = 1024;              // start positions per window
constexpr uint32_t GWIN = 2 * GSEG;          // window bytes
constexpr uint32_t GWIN_WORDS = GWIN / 4 + 4;
constexpr uint32_t GWPB = 4;                 // waves per block
constexpr uint32_t GHDR_LDS = 512;           // gene headers kept in LDS
constexpr uint32_t GW_LDS = 2048;            // name words kept in LDS

__device__ __forceinline__ uint32_t win_word(const uint32_t *win, uint32_t b)
{
    return __builtin_amdgcn_alignbyte(win[(b >> 2) + 1], win[b >> 2], b & 3u);
}

// One gene name against the positions of one chunk (lane lp, absolute p):
// sets the gene's bit when some lane holds a match.  WL: the name words are
// in LDS (wl) instead of global memory (gw); the two stay in separate code
// so that neither read becomes a flat load.
template <bool WL>
__device__ __forceinline__ void test_gene(const idb_gene_hdr &h, uint32_t g, const uint32_t *win, uint32_t w,
                                          uint32_t lp, uint32_t p, bool pv, uint32_t L, const uint32_t *wl,
                                          const uint32_t *gw, uint32_t *bits, uint32_t lane)
{
    if (h.len > L) return;  // longer than the row, or a name that never matches
    const uint32_t mask = h.len >= 4 ? 0xFFFFFFFFu : ((1u << (8 * h.len)) - 1u);
    bool ok = pv && p + h.len <= L && ((w ^ h.w0) & mask) == 0u;
    unsigned long long m = __ballot(ok);
    if (m && h.len > 4) {
        // the last 4 bytes
        if (ok) ok = win_word(win, lp + h.len - 4) == h.w1;
        m = __ballot(ok);
    }
    // the middle of the name, 4 bytes at a time
    for (uint32_t k = 4; m && k + 4 < h.len; k += 4) {
        const uint32_t gk = WL ? wl[h.woff + (k >> 2)] : gw[h.woff + (k >> 2)];
        if (ok) ok = win_word(win, lp + k) == gk;
        m = __ballot(ok);
    }
    if (m && lane == 0) bits[g >> 5] |= 1u << (g & 31);
}

template <bool WL>
__global__ __launch_bounds__(256) void k_gene_bitsets(const uint8_t *calls, const uint64_t *off, const uint32_t *len,
                                                      uint32_t n, const idb_gene_hdr *__restrict__ hdr,
                                                      const uint32_t *__restrict__ gw, uint32_t n_words,
                                                      uint32_t n_genes, int32_t *out, uint32_t width)
{
    __shared__ uint32_t win_all[GWPB][GWIN_WORDS];
    __shared__ uint32_t bits_all[GWPB][IDB_GENE_MAX_WIDTH];
    __shared__ idb_gene_hdr hl[GHDR_LDS];
    __shared__ uint32_t wl_all[GW_LDS];
    const uint32_t nl = n_genes < GHDR_LDS ? n_genes : GHDR_LDS;
    for (uint32_t g = threadIdx.x; g < nl; g += blockDim.x) hl[g] = hdr[g];
    if (WL)
        for (uint32_t k = threadIdx.x; k < n_words; k += blockDim.x) wl_all[k] = gw[k];
    __syncthreads();
    const uint32_t lane = __lane_id();
    const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t *win = win_all[wib];
    uint32_t *bits = bits_all[wib];
    const uint32_t total = gridDim.x * GWPB;
    for (uint32_t row = blockIdx.x * GWPB + wib; row < n; row += total) {
        const uint32_t L = len[row];
        const uint8_t *src = calls + off[row];
        for (uint32_t k = lane; k < width; k += 64) bits[k] = 0u;
        for (uint32_t s0 = 0; s0 < L; s0 += GSEG) {
            // window bytes [s0, s0 + wl), zero past them
            const uint32_t wl = L - s0 < GWIN ? L - s0 : GWIN;
            const uint32_t nw = (wl + 3) / 4 + 2;
            for (uint32_t k = lane; k < nw; k += 64) {
                // four independent byte loads, each clamped into the row
                uint32_t v = 0u;
#pragma unroll
                for (uint32_t b = 0; b < 4; ++b) {
                    const uint32_t i = 4 * k + b;
                    const uint32_t c = src[s0 + (i < wl ? i : wl - 1)];
                    v |= (i < wl ? c : 0u) << (8 * b);
                }
                win[k] = v;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            const uint32_t pend = L - s0 < GSEG ? L - s0 : GSEG;  // start positions in this window
            for (uint32_t pc = 0; pc < pend; pc += 64) {
                const uint32_t lp = pc + lane;
                const bool pv = lp < pend;
                const uint32_t w = win_word(win, pv ? lp : 0u);
                const uint32_t p = s0 + lp;
                // headers from LDS, four at a time so that their reads overlap,
                // then (past the first GHDR_LDS) from global memory
                uint32_t g = 0;
                for (; g + 4 <= nl; g += 4) {
                    const idb_gene_hdr h0 = hl[g], h1 = hl[g + 1], h2 = hl[g + 2], h3 = hl[g + 3];
                    test_gene<WL>(h0, g, win, w, lp, p, pv, L, wl_all, gw, bits, lane);
                    test_gene<WL>(h1, g + 1, win, w, lp, p, pv, L, wl_all, gw, bits, lane);
                    test_gene<WL>(h2, g + 2, win, w, lp, p, pv, L, wl_all, gw, bits, lane);
                    test_gene<WL>(h3, g + 3, win, w, lp, p, pv, L, wl_all, gw, bits, lane);
                }
                for (; g < nl; ++g)
                    test_gene<WL>(hl[g], g, win, w, lp, p, pv, L, wl_all, gw, bits, lane);
                for (uint32_t g = nl; g < n_genes; ++g)
                    test_gene<WL>(hdr[g], g, win, w, lp, p, pv, L, wl_all, gw, bits, lane);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        for (uint32_t k = lane; k < width; k += 64) out[(size_t)row * width + k] = (int32_t)bits[k];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    }
}

}  // namespace

hipError_t idb_launch_gene_bitsets(const uint8_t *calls, const uint64_t *off, const uint32_t *len, uint32_t n,
                                   const idb_gene_hdr *hdr, const uint32_t *gw, uint32_t n_words, uint32_t n_genes,
                                   int32_t *out, uint32_t width, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    if (width > IDB_GENE_MAX_WIDTH) return hipErrorInvalidValue;
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_gene_bitsets<true>, 64 * GWPB, 0) !=
            hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    (void)hipGetLastError();
    uint64_t blocks = ((uint64_t)n + GWPB - 1) / GWPB;
    const uint64_t cap = (uint64_t)per_cu * (uint64_t)cus;
    if (blocks > cap) blocks = cap;
    if (n_words <= GW_LDS)
        hipLaunchKernelGGL(k_gene_bitsets<true>, dim3((uint32_t)blocks), dim3(64 * GWPB), 0, s, calls, off, len, n,
                           hdr, gw, n_words, n_genes, out, width);
    else
        hipLaunchKernelGGL(k_gene_bitsets<false>, dim3((uint32_t)blocks), dim3(64 * GWPB), 0, s, calls, off, len, n,
                           hdr, gw, n_words, n_genes, out, width);
    return hipGetLastError();
}
