// idb_learned.hip — the learned candidate filter's encoder on gfx950: a small
// 1-D CNN over CDR3 letters, trained and run by this file's own kernels.
//
// No counterpart in the reference (it evaluates every pair exactly).  The
// encoder maps a row to a unit vector z; the learned filter proposes the pairs
// of a tile whose dot product z_x . z_y reaches a threshold, and the exact
// Myers kernel verifies every proposal, so distances are never approximated.
//
//   tokens (<= P letters) -> embedding 32 x 32 -> 3 x [conv1d k=5 'same',
//   128 channels, GELU] -> mean over letters -> linear 128 -> 64 -> L2 norm.
//
// P, the positions per row, is fixed when the encoder is made: 32 for
// protein rows, 160 for nucleotide rows of up to 160 letters.
//
// Weights are bfloat16 copies of float32 master weights (Adam keeps the
// master copy); activations are bfloat16, all products accumulate in float32.
// The three convolutions, their data gradients and their weight gradients run
// on one MFMA GEMM kernel (v_mfma_f32_32x32x16_bf16) whose operand loaders
// gather im2col windows straight from the activations, so no im2col matrix is
// ever materialised.  Row m of an activation matrix is position m % P of
// sequence m / P; positions at or past the sequence's length are zero after
// every layer (each row is convolved with its own zero padding).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "idb_internal.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr uint32_t T = 5, NTOK = 32, C0 = 32, C = 128, D = 64;
constexpr float GELU_K = 0.7978845608028654f;  // sqrt(2 / pi)

// Flat parameter layout (float32 master and bfloat16 copy share it).
constexpr uint32_t O_EMB = 0, O_W1 = O_EMB + NTOK * C0, O_B1 = O_W1 + C * T * C0, O_W2 = O_B1 + C,
                   O_B2 = O_W2 + C * T * C, O_W3 = O_B2 + C, O_B3 = O_W3 + C * T * C, O_WP = O_B3 + C,
                   O_BP = O_WP + D * C, N_PARAMS = O_BP + D;
static_assert(N_PARAMS == IDB_ENC_PARAMS, "parameter count");

// Round to nearest even (the hardware conversion, one instruction).
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// tanh(u) = 1 - 2 / (exp(2u) + 1): one hardware exponential and one
// reciprocal (exp overflows to inf for large u and gives 1; underflows to 0
// for very negative u and gives -1).  Its error is far below the bfloat16
// rounding every activation goes through.
__device__ __forceinline__ float tanh_exp(float u) { return 1.f - __fdividef(2.f, __expf(2.f * u) + 1.f); }

// 0.5 x (1 + tanh(u)) = x / (1 + exp(-2u)), u = k (x + 0.044715 x^3): one
// hardware base-2 exponential and one reciprocal.  For very negative x the
// exponential overflows to inf and the result is -0; for large x it is x.
__device__ __forceinline__ float gelu(float x)
{
    constexpr float C1 = -2.f * GELU_K * 1.4426950408889634f, C3 = C1 * 0.044715f;  // -2u log2(e) = x (C1 + C3 x^2)
    const float e = x * __builtin_fmaf(C3, x * x, C1);
    return x * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(e));
}
__device__ __forceinline__ float gelu_grad(float x)
{
    const float th = tanh_exp(GELU_K * (x + 0.044715f * x * x * x));
    return 0.5f * (1.f + th) + 0.5f * x * (1.f - th * th) * GELU_K * (1.f + 3.f * 0.044715f * x * x);
}

// ---------------------------------------------------------------------------
// GEMM: out[m][n] = sum_k A[m][k] B[n][k] over a block tile of 128 x 128 with
// k chunks of 64, double-buffered in LDS, with global loads two chunks ahead
// of the chunk whose 16 MFMAs per wave are running.  Four waves own
// 64 x 64 each (2 x 2 MFMA tiles of 32 x 32).  LDS rows hold 64 k values plus
// 8 of padding (144 bytes), so the 16-byte fragment reads of 16 consecutive
// rows hit disjoint banks.  Loaders fetch 8 consecutive k of one row (ROW) or
// 8 consecutive rows at one k (COL, stored transposed) per 16-byte load.
constexpr uint32_t GM = 128, GN = 128, GK = 64, LDK = GK + 8;
constexpr uint32_t GCH = GM * GK / 8 / 256;  // 16-byte chunks per thread and operand

// k at or past kend (the end of this block's k range) reads as zero.
template <bool COL, class L>
__device__ __forceinline__ void fetch_tile(const L &ld, uint32_t r0, uint32_t k0, uint32_t kend, uint32_t tid,
                                           uint4 (&v)[GCH])
{
#pragma unroll
    for (uint32_t i = 0; i < GCH; ++i) {
        const uint32_t c = tid + i * 256;
        const uint32_t k = COL ? k0 + (c >> 4) : k0 + (c & 7) * 8;
        v[i] = make_uint4(0, 0, 0, 0);
        if (k < kend) v[i] = COL ? ld.fetch(r0 + (c & 15) * 8, k) : ld.fetch(r0 + (c >> 3), k);
    }
}

template <bool COL>
__device__ __forceinline__ void store_tile(uint16_t *lds, const uint4 (&v)[GCH], uint32_t tid)
{
#pragma unroll
    for (uint32_t i = 0; i < GCH; ++i) {
        const uint32_t c = tid + i * 256;
        if (!COL) {
            *(uint4 *)(lds + (c >> 3) * LDK + (c & 7) * 8) = v[i];
        } else {
            const uint32_t r = (c & 15) * 8, k = c >> 4;
            const uint32_t w[4] = {v[i].x, v[i].y, v[i].z, v[i].w};
#pragma unroll
            for (uint32_t e = 0; e < 8; ++e) lds[(r + e) * LDK + k] = (uint16_t)(w[e >> 1] >> (16 * (e & 1)));
        }
    }
}

template <class LA, class LB, class EP>
__global__ __launch_bounds__(256, 2) void k_gemm(LA la, LB lb, EP ep, uint32_t M, uint32_t N, uint32_t K, uint32_t ksplit)
{
    __shared__ __align__(16) uint16_t As[2][GM * LDK];
    __shared__ __align__(16) uint16_t Bs[2][GN * LDK];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t m0 = blockIdx.x * GM, n0 = blockIdx.y * GN;
    const uint32_t kb = blockIdx.z * ksplit, ke = min(K, kb + ksplit);
    const uint32_t wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
    if (kb >= ke) return;
    // Loads run two chunks ahead: chunk c + 1 sits in registers (ra, rb)
    // while chunk c + 2 is in flight (na, nb) and chunk c is consumed from LDS.
    uint4 ra[GCH], rb[GCH], na[GCH] = {}, nb[GCH] = {};
    fetch_tile<LA::COL>(la, m0, kb, ke, tid, ra);
    fetch_tile<LB::COL>(lb, n0, kb, ke, tid, rb);
    store_tile<LA::COL>(As[0], ra, tid);
    store_tile<LB::COL>(Bs[0], rb, tid);
    fetch_tile<LA::COL>(la, m0, kb + GK, ke, tid, ra);
    fetch_tile<LB::COL>(lb, n0, kb + GK, ke, tid, rb);
    __syncthreads();
    uint32_t buf = 0;
    for (uint32_t k0 = kb; k0 < ke; k0 += GK) {
        const bool more = k0 + GK < ke;
        if (k0 + 2 * GK < ke) {
            fetch_tile<LA::COL>(la, m0, k0 + 2 * GK, ke, tid, na);
            fetch_tile<LB::COL>(lb, n0, k0 + 2 * GK, ke, tid, nb);
        }
#pragma unroll
        for (uint32_t ks = 0; ks < GK / 16; ++ks) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *(const bf16x8 *)(As[buf] + (wm + i * 32 + (lane & 31)) * LDK + ks * 16 + (lane >> 5) * 8);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *(const bf16x8 *)(Bs[buf] + (wn + j * 32 + (lane & 31)) * LDK + ks * 16 + (lane >> 5) * 8);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            store_tile<LA::COL>(As[buf ^ 1], ra, tid);
            store_tile<LB::COL>(Bs[buf ^ 1], rb, tid);
#pragma unroll
            for (uint32_t i = 0; i < GCH; ++i) {
                ra[i] = na[i];
                rb[i] = nb[i];
            }
        }
        __syncthreads();
        buf ^= 1;
    }
    // result layout of the 32 x 32 MFMA: lane l holds column l % 32 and rows
    // 8 (r / 4) + 4 (l / 32) + r % 4 for r = 0 .. 15
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) ep.tile(acc[i][j], m0 + wm + i * 32, n0 + wn + j * 32 + (lane & 31), lane, M, N);
}

// Row of element r of a lane's 32 x 32 result tile, relative to the tile.
__device__ __forceinline__ uint32_t tile_row(int r, uint32_t lane) { return (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3); }

// im2col window of an activation matrix X[rows][cin]: element (m, t * cin +
// c) is X[PP s + q][c] with q = p + t - 2 (forward) or q = p - t + 2 (data
// gradient), zero outside the PP positions of sequence s = m / PP.
template <uint32_t PP>
struct LdConv {
    static constexpr bool COL = false;
    const uint16_t *X;
    uint32_t cin, M;  // cin a power of two
    int flip;
    __device__ uint4 fetch(uint32_t m, uint32_t k) const
    {
        if (m >= M) return make_uint4(0, 0, 0, 0);
        const uint32_t sh = 31u - __builtin_clz(cin);
        const int t = (int)(k >> sh), c = (int)(k & (cin - 1u)), p = (int)(m % PP);
        const int q = flip ? p - t + 2 : p + t - 2;
        if (q < 0 || q >= (int)PP) return make_uint4(0, 0, 0, 0);
        return *(const uint4 *)(X + ((size_t)(m - (uint32_t)p) + q) * cin + c);
    }
};

// Plain row-major matrix R[rows][ld], k contiguous.
struct LdRows {
    static constexpr bool COL = false;
    const uint16_t *R;
    uint32_t ld, rows;
    __device__ uint4 fetch(uint32_t r, uint32_t k) const
    {
        if (r >= rows) return make_uint4(0, 0, 0, 0);
        return *(const uint4 *)(R + (size_t)r * ld + k);
    }
};

// Weight gradient, operand A: rows are output channels o, k runs over
// activation rows m: dH[m][o] (8 consecutive o per load).
struct LdGradT {
    static constexpr bool COL = true;
    const uint16_t *dH;
    uint32_t M;
    __device__ uint4 fetch(uint32_t o, uint32_t m) const
    {
        if (m >= M) return make_uint4(0, 0, 0, 0);
        return *(const uint4 *)(dH + (size_t)m * C + o);
    }
};

// Weight gradient, operand B: rows are im2col columns k' = t * cin + c, k runs
// over activation rows m: X[PP s + p + t - 2][c] (8 consecutive c per load).
template <uint32_t PP>
struct LdConvT {
    static constexpr bool COL = true;
    const uint16_t *X;
    uint32_t cin, M;  // cin a power of two
    __device__ uint4 fetch(uint32_t kc, uint32_t m) const
    {
        if (m >= M || kc >= T * cin) return make_uint4(0, 0, 0, 0);
        const uint32_t sh = 31u - __builtin_clz(cin);
        const uint32_t p = m % PP;
        const int t = (int)(kc >> sh), c = (int)(kc & (cin - 1u)), q = (int)p + t - 2;
        if (q < 0 || q >= (int)PP) return make_uint4(0, 0, 0, 0);
        return *(const uint4 *)(X + ((size_t)(m - p) + q) * cin + c);
    }
};

// Epilogues get one 32 x 32 result tile per call: rows row0 .. row0 + 31
// (within one sequence: row0 and PP are multiples of 32) and this lane's
// column `col`.

// Forward epilogue: bias, saved pre-activation, GELU, length mask.
template <uint32_t PP>
struct EpConvFwd {
    const uint16_t *bias;
    const uint32_t *len;
    uint16_t *H, *A;
    __device__ void tile(const f32x16 &v, uint32_t row0, uint32_t col, uint32_t lane, uint32_t M, uint32_t N) const
    {
        if (col >= N || row0 >= M) return;
        const float b = bf2f(bias[col]);
        // rows row0 + pl + d(r), d(r) = 8 (r / 4) + r % 4, within one sequence
        const uint32_t pl = (lane >> 5) * 4;
        const int32_t lim = (int32_t)len[row0 / PP] - (int32_t)(row0 % PP + pl);  // rows d < lim are letters
        const size_t base = (size_t)(row0 + pl) * C + col;
        uint16_t *a = A + base;
        if (H) {
            uint16_t *h = H + base;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int32_t d = (r >> 2) * 8 + (r & 3);
                const float x = v[r] + b;
                h[d * C] = f2bf(x);
                a[d * C] = d < lim ? f2bf(gelu(x)) : (uint16_t)0;
            }
        } else {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int32_t d = (r >> 2) * 8 + (r & 3);
                a[d * C] = d < lim ? f2bf(gelu(v[r] + b)) : (uint16_t)0;
            }
        }
    }
};

// Data-gradient epilogue: masked, times GELU' of the previous layer's saved
// pre-activation (bfloat16 out), or the embedding gradient (float out).
template <uint32_t PP>
struct EpConvBwd {
    const uint16_t *Hprev;
    const uint32_t *len;
    uint32_t cin;
    uint16_t *dHprev;
    float *dX0;
    __device__ void tile(const f32x16 &v, uint32_t row0, uint32_t col, uint32_t lane, uint32_t M, uint32_t N) const
    {
        if (col >= N || row0 >= M) return;
        const uint32_t L = len[row0 / PP], p0 = row0 % PP;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const uint32_t p = tile_row(r, lane);
            const size_t at = (size_t)(row0 + p) * cin + col;
            const bool valid = p0 + p < L;
            if (dHprev) dHprev[at] = valid ? f2bf(v[r] * gelu_grad(bf2f(Hprev[at]))) : (uint16_t)0;
            else dX0[at] = valid ? v[r] : 0.f;
        }
    }
};

struct EpAtomic {
    float *G;
    uint32_t ld;
    __device__ void tile(const f32x16 &v, uint32_t row0, uint32_t col, uint32_t lane, uint32_t M, uint32_t N) const
    {
        if (col >= N) return;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const uint32_t row = row0 + tile_row(r, lane);
            if (row < M) atomicAdd(G + (size_t)row * ld + col, v[r]);
        }
    }
};

// ---------------------------------------------------------------------------
// VALU kernels

// Tokens of every row: vocabulary map, first P letters, token 0 past the end.
__global__ void k_tokenize(const uint8_t *seq, const uint64_t *off, const uint32_t *len, uint32_t n,
                           const uint8_t *vocab, uint32_t P, uint8_t *tok, uint32_t *tlen)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t L = min(len[i], P);
        const uint8_t *s = seq + off[i];
        for (uint32_t p = 0; p < P; ++p) tok[(size_t)i * P + p] = p < L ? vocab[s[p]] : (uint8_t)0;
        tlen[i] = L;
    }
}

// Batch assembly: sequence b of the batch is row rows[b] (IDB_NONE: empty).
__global__ void k_gather_rows(const uint8_t *tok, const uint32_t *tlen, const uint32_t *rows, uint32_t nb,
                              uint32_t P, uint8_t *btok, uint32_t *blen)
{
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        const uint32_t r = rows[b];
        const uint32_t L = r == IDB_NONE ? 0u : tlen[r];
        for (uint32_t p = 0; p < P; ++p) btok[(size_t)b * P + p] = r == IDB_NONE ? (uint8_t)0 : tok[(size_t)r * P + p];
        blen[b] = L;
    }
}

// Consecutive rows first .. first + nb - 1 (past n: empty).
__global__ void k_iota_rows(uint32_t first, uint32_t n, uint32_t nb, uint32_t *rows)
{
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x)
        rows[b] = first + b < n ? first + b : IDB_NONE;
}

// Pairs of a training batch: sequences 2 i and 2 i + 1 are the pair's rows.
__global__ void k_pair_rows(const uint32_t *pairs, const uint32_t *order, uint64_t first, uint32_t nb, uint64_t n,
                            uint32_t *rows, float *y, const float *labels)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) {
        const uint64_t k = first + i;
        if (k < n) {
            const uint32_t q = order ? order[k] : (uint32_t)k;
            rows[2 * i] = pairs[2 * (size_t)q];
            rows[2 * i + 1] = pairs[2 * (size_t)q + 1];
            y[i] = labels[q];
        } else {
            rows[2 * i] = rows[2 * i + 1] = IDB_NONE;
            y[i] = -1.f;  // padding pair: no loss
        }
    }
}

__global__ void k_embed(const uint8_t *btok, const uint32_t *blen, uint32_t nseq, uint32_t P, const uint16_t *emb,
                        uint16_t *X0)
{
    const uint32_t total = nseq * P * (C0 / 8);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const uint32_t m = i / (C0 / 8), c8 = (i % (C0 / 8)) * 8;
        const bool valid = (m % P) < blen[m / P];
        uint4 v = make_uint4(0, 0, 0, 0);
        if (valid) v = *(const uint4 *)(emb + (size_t)btok[m] * C0 + c8);
        *(uint4 *)(X0 + (size_t)m * C0 + c8) = v;
    }
}

// Head: mean over the sequence's letters, projection, L2 norm.  One wave per
// sequence: lanes hold channels l and l + 64, then output l.
__global__ void k_head_fwd(const uint16_t *A3, const uint32_t *blen, uint32_t nseq, uint32_t P, const uint16_t *wp_b,
                           const uint16_t *bp, float *pool, float *u, float *z, uint16_t *zb)
{
    __shared__ float sp[4][C];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t s = blockIdx.x * 4 + w; s < nseq; s += gridDim.x * 4) {
        const uint32_t L = blen[s];
        float a0 = 0.f, a1 = 0.f;
        for (uint32_t p = 0; p < L; ++p) {
            a0 += bf2f(A3[((size_t)s * P + p) * C + lane]);
            a1 += bf2f(A3[((size_t)s * P + p) * C + lane + 64]);
        }
        const float inv = L ? 1.f / (float)L : 0.f;
        a0 *= inv;
        a1 *= inv;
        sp[w][lane] = a0;
        sp[w][lane + 64] = a1;
        if (pool) {
            pool[(size_t)s * C + lane] = a0;
            pool[(size_t)s * C + lane + 64] = a1;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        float acc = bf2f(bp[lane]);
        for (uint32_t o = 0; o < C; ++o) acc += bf2f(wp_b[(size_t)lane * C + o]) * sp[w][o];
        float ss = acc * acc;
#pragma unroll
        for (int k = 32; k > 0; k >>= 1) ss += __shfl_xor(ss, k, 64);
        const float nrm = fmaxf(sqrtf(ss), 1e-6f);
        const float zz = acc / nrm;
        if (u) u[(size_t)s * D + lane] = acc;
        if (z) z[(size_t)s * D + lane] = zz;
        if (zb) zb[(size_t)s * D + lane] = f2bf(zz);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    }
}

// Contrastive loss of pair i (sequences 2i, 2i+1); one wave per pair.
__global__ void k_pair_loss(const float *z, const float *y, uint32_t npairs, float margin, float inv_b, float *dz,
                            float *loss)
{
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    float lsum = 0.f;
    for (uint32_t i = blockIdx.x * 4 + w; i < npairs; i += gridDim.x * 4) {
        const float yi = y[i];
        const float diff = z[(size_t)(2 * i) * D + lane] - z[(size_t)(2 * i + 1) * D + lane];
        float d2 = diff * diff;
#pragma unroll
        for (int k = 32; k > 0; k >>= 1) d2 += __shfl_xor(d2, k, 64);
        float g = 0.f;
        if (yi >= 0.f) {
            const float d = sqrtf(d2 + 1e-12f);
            const float hinge = fmaxf(0.f, margin - d);
            if (lane == 0) lsum += yi * d2 + (1.f - yi) * hinge * hinge;
            g = (2.f * yi - (1.f - yi) * 2.f * hinge / d) * inv_b * diff;
        }
        dz[(size_t)(2 * i) * D + lane] = g;
        dz[(size_t)(2 * i + 1) * D + lane] = -g;
    }
    if (lane == 0 && lsum != 0.f) atomicAdd(loss, lsum * inv_b);
}

// Head backward, one wave per sequence: du, dpool, and the last layer's
// pre-activation gradient dH3 = (dpool / L) * GELU'(H3) at the valid letters.
__global__ void k_head_bwd(const float *dz, const float *z, const float *u, const uint32_t *blen, uint32_t nseq,
                           uint32_t P, const uint16_t *wp_b, const uint16_t *H3, float *du_out, uint16_t *dH3)
{
    __shared__ float sdu[4][D];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t s = blockIdx.x * 4 + w; s < nseq; s += gridDim.x * 4) {
        const float zz = z[(size_t)s * D + lane], g = dz[(size_t)s * D + lane], uu = u[(size_t)s * D + lane];
        float zg = zz * g, ss = uu * uu;
#pragma unroll
        for (int k = 32; k > 0; k >>= 1) {
            zg += __shfl_xor(zg, k, 64);
            ss += __shfl_xor(ss, k, 64);
        }
        const float nrm = fmaxf(sqrtf(ss), 1e-6f);
        const float du = (g - zz * zg) / nrm;
        du_out[(size_t)s * D + lane] = du;
        sdu[w][lane] = du;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        const uint32_t L = blen[s];
        const float inv = L ? 1.f / (float)L : 0.f;
        float dp0 = 0.f, dp1 = 0.f;
        for (uint32_t j = 0; j < D; ++j) {
            dp0 += bf2f(wp_b[(size_t)j * C + lane]) * sdu[w][j];
            dp1 += bf2f(wp_b[(size_t)j * C + lane + 64]) * sdu[w][j];
        }
        dp0 *= inv;
        dp1 *= inv;
        for (uint32_t p = 0; p < P; ++p) {
            const size_t m = (size_t)s * P + p;
            const bool valid = p < L;
            dH3[m * C + lane] = valid ? f2bf(dp0 * gelu_grad(bf2f(H3[m * C + lane]))) : (uint16_t)0;
            dH3[m * C + lane + 64] = valid ? f2bf(dp1 * gelu_grad(bf2f(H3[m * C + lane + 64]))) : (uint16_t)0;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    }
}

// Projection gradients: dWp[j][o] = sum_s du[s][j] pool[s][o], dbp[j] = sum_s du[s][j].
__global__ void k_proj_grad(const float *du, const float *pool, uint32_t nseq, float *g_wp, float *g_bp)
{
    const uint32_t j = blockIdx.x, o = threadIdx.x;  // D blocks x C threads
    float acc = 0.f, accb = 0.f;
    for (uint32_t s = blockIdx.y; s < nseq; s += gridDim.y) {
        const float d = du[(size_t)s * D + j];
        acc += d * pool[(size_t)s * C + o];
        accb += d;
    }
    atomicAdd(g_wp + (size_t)j * C + o, acc);
    if (o == 0) atomicAdd(g_bp + j, accb);
}

// Bias gradient: column sums of dH[M][C].
__global__ void k_bias_grad(const uint16_t *dH, uint32_t M, float *gb)
{
    const uint32_t o = threadIdx.x;  // C threads
    float acc = 0.f;
    for (uint32_t m = blockIdx.x; m < M; m += gridDim.x) acc += bf2f(dH[(size_t)m * C + o]);
    atomicAdd(gb + o, acc);
}

// Embedding gradient: scatter-add of dX0 rows by token, through LDS.
__global__ void k_emb_grad(const float *dX0, const uint8_t *btok, const uint32_t *blen, uint32_t M, uint32_t P,
                           float *gemb)
{
    __shared__ float acc[NTOK * C0];
    for (uint32_t i = threadIdx.x; i < NTOK * C0; i += blockDim.x) acc[i] = 0.f;
    __syncthreads();
    const uint32_t c = threadIdx.x & (C0 - 1);
    for (uint32_t m = blockIdx.x * (blockDim.x / C0) + threadIdx.x / C0; m < M; m += gridDim.x * (blockDim.x / C0)) {
        if ((m % P) >= blen[m / P]) continue;
        atomicAdd(&acc[(uint32_t)btok[m] * C0 + c], dX0[(size_t)m * C0 + c]);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < NTOK * C0; i += blockDim.x)
        if (acc[i] != 0.f) atomicAdd(gemb + i, acc[i]);
}

// Data-gradient weights: WT[c][t * C + o] = W[o][t * cin + c] (bfloat16).
__global__ void k_transpose_w(const uint16_t *W, uint32_t cin, uint16_t *WT)
{
    const uint32_t total = C * T * cin;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const uint32_t o = i / (T * cin), k = i % (T * cin), t = k / cin, c = k % cin;
        WT[(size_t)c * (T * C) + t * C + o] = W[i];
    }
}

__global__ void k_to_bf16(const float *src, uint32_t n, uint16_t *dst)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = f2bf(src[i]);
}

__global__ void k_adam(float *th, const float *g, float *m, float *v, uint32_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2, uint16_t *thb)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float t = th[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        th[i] = t;
        thb[i] = f2bf(t);
    }
}

uint32_t nblocks(uint64_t n, uint32_t bs = 256)
{
    uint64_t b = (n + bs - 1) / bs;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 4096));
}

struct Fail {
    int code;
    std::string msg;
};

void hcheck(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw Fail{e == hipErrorOutOfMemory ? IDB_ERR_OOM : IDB_ERR_HIP,
                                    std::string(what) + ": " + hipGetErrorString(e)};
}

}  // namespace

// ---------------------------------------------------------------------------
// Encoder state: parameters, optimiser moments, and a workspace sized for the
// largest batch seen so far.
struct idb_encoder {
    int device = 0;
    uint32_t P = IDB_ENC_P;  // positions per row: 32 or 160
    hipStream_t stream = nullptr;
    uint8_t vocab[256];
    float *theta = nullptr, *grad = nullptr, *mom1 = nullptr, *mom2 = nullptr;
    uint16_t *theta_b = nullptr;
    uint16_t *wt[3] = {nullptr, nullptr, nullptr};  // transposed conv weights for data gradients
    uint32_t adam_step = 0;
    std::string err;
    // workspace
    uint32_t cap_seq = 0;
    bool cap_train = false;
    uint8_t *btok = nullptr;
    uint32_t *blen = nullptr, *brows = nullptr;
    uint16_t *X0 = nullptr, *A[3] = {nullptr, nullptr, nullptr}, *H[3] = {nullptr, nullptr, nullptr};
    uint16_t *dHa = nullptr, *dHb = nullptr;
    float *dX0 = nullptr, *pool = nullptr, *u = nullptr, *z = nullptr, *dz = nullptr, *du = nullptr, *y = nullptr;
    float *loss = nullptr;
    // tokens of the last dataset (kept apart from the batch workspace)
    uint8_t *tok = nullptr;
    uint32_t *tlen = nullptr;
    uint32_t tok_cap = 0;
    uint8_t *vocab_d = nullptr;
    std::vector<void *> fixed, ws;  // parameter and workspace allocations
};

namespace {

template <typename T>
T *dalloc(std::vector<void *> &owner, size_t count)
{
    void *p = nullptr;
    hcheck(hipMalloc(&p, std::max<size_t>(count * sizeof(T), 16)), "hipMalloc");
    owner.push_back(p);
    return (T *)p;
}

void free_workspace(idb_encoder *e)
{
    for (void *p : e->ws) (void)hipFree(p);
    e->ws.clear();
    e->cap_seq = 0;
}

void ensure_workspace(idb_encoder *e, uint32_t nseq, bool train)
{
    if (nseq <= e->cap_seq && (!train || e->cap_train)) return;
    free_workspace(e);
    const size_t M = (size_t)nseq * e->P;
    e->btok = dalloc<uint8_t>(e->ws, M);
    e->blen = dalloc<uint32_t>(e->ws, nseq);
    e->brows = dalloc<uint32_t>(e->ws, nseq);
    e->X0 = dalloc<uint16_t>(e->ws, M * C0);
    e->A[0] = dalloc<uint16_t>(e->ws, M * C);
    e->A[1] = dalloc<uint16_t>(e->ws, M * C);
    e->A[2] = train ? dalloc<uint16_t>(e->ws, M * C) : e->A[0];
    e->pool = dalloc<float>(e->ws, (size_t)nseq * C);
    e->u = dalloc<float>(e->ws, (size_t)nseq * D);
    e->z = dalloc<float>(e->ws, (size_t)nseq * D);
    if (train) {
        for (int l = 0; l < 3; ++l) e->H[l] = dalloc<uint16_t>(e->ws, M * C);
        e->dHa = dalloc<uint16_t>(e->ws, M * C);
        e->dHb = dalloc<uint16_t>(e->ws, M * C);
        e->dX0 = dalloc<float>(e->ws, M * C0);
        e->dz = dalloc<float>(e->ws, (size_t)nseq * D);
        e->du = dalloc<float>(e->ws, (size_t)nseq * D);
        e->y = dalloc<float>(e->ws, nseq / 2 + 1);
        e->loss = dalloc<float>(e->ws, 1);
    } else {
        for (int l = 0; l < 3; ++l) e->H[l] = nullptr;
    }
    e->cap_seq = nseq;
    e->cap_train = train;
}

void refresh_transposed(idb_encoder *e)
{
    const uint32_t cins[3] = {C0, C, C};
    const uint32_t offs[3] = {O_W1, O_W2, O_W3};
    for (int l = 0; l < 3; ++l)
        hipLaunchKernelGGL(k_transpose_w, dim3(nblocks((size_t)C * T * cins[l])), dim3(256), 0, e->stream,
                           e->theta_b + offs[l], cins[l], e->wt[l]);
    hcheck(hipGetLastError(), "transpose launch");
}

// Forward over the batch in e->btok / e->blen (nseq, a multiple of 4).
template <uint32_t PP>
void forward_p(idb_encoder *e, uint32_t nseq, bool train, uint16_t *zb_out)
{
    hipStream_t s = e->stream;
    const uint32_t M = nseq * PP;
    hipLaunchKernelGGL(k_embed, dim3(nblocks((size_t)M * C0 / 8)), dim3(256), 0, s, e->btok, e->blen, nseq, PP,
                       e->theta_b + O_EMB, e->X0);
    const uint16_t *in = e->X0;
    uint32_t cin = C0;
    const uint32_t offs[3] = {O_W1, O_W2, O_W3}, boff[3] = {O_B1, O_B2, O_B3};
    for (int l = 0; l < 3; ++l) {
        uint16_t *out = train ? e->A[l] : e->A[l & 1];
        LdConv<PP> la{in, cin, M, 0};
        LdRows lb{e->theta_b + offs[l], T * cin, C};
        EpConvFwd<PP> ep{e->theta_b + boff[l], e->blen, train ? e->H[l] : nullptr, out};
        hipLaunchKernelGGL((k_gemm<LdConv<PP>, LdRows, EpConvFwd<PP>>), dim3((M + GM - 1) / GM, 1, 1), dim3(256), 0, s,
                           la, lb, ep, M, C, T * cin, T * cin);
        in = out;
        cin = C;
    }
    hipLaunchKernelGGL(k_head_fwd, dim3(nblocks(nseq, 4)), dim3(256), 0, s, in, e->blen, nseq, PP, e->theta_b + O_WP,
                       e->theta_b + O_BP, train ? e->pool : nullptr, train ? e->u : nullptr, e->z, zb_out);
    hcheck(hipGetLastError(), "encoder forward launch");
}

void forward(idb_encoder *e, uint32_t nseq, bool train, uint16_t *zb_out)
{
    if (e->P == 160) forward_p<160>(e, nseq, train, zb_out);
    else forward_p<32>(e, nseq, train, zb_out);
}

// Backward of sum(dz . z) into e->grad (zeroed here).
template <uint32_t PP>
void backward_p(idb_encoder *e, uint32_t nseq)
{
    hipStream_t s = e->stream;
    const uint32_t M = nseq * PP;
    hcheck(hipMemsetAsync(e->grad, 0, sizeof(float) * N_PARAMS, s), "memset");
    hipLaunchKernelGGL(k_head_bwd, dim3(nblocks(nseq, 4)), dim3(256), 0, s, e->dz, e->z, e->u, e->blen, nseq, PP,
                       e->theta_b + O_WP, e->H[2], e->du, e->dHa);
    hipLaunchKernelGGL(k_proj_grad, dim3(D, 64), dim3(C), 0, s, e->du, e->pool, nseq, e->grad + O_WP, e->grad + O_BP);
    const uint32_t offs[3] = {O_W1, O_W2, O_W3}, boff[3] = {O_B1, O_B2, O_B3}, cins[3] = {C0, C, C};
    uint16_t *dH = e->dHa, *dHn = e->dHb;
    // split the weight-gradient reduction over rows into chunks of 4096 rows
    const uint32_t ks = 4096, nsplit = (M + ks - 1) / ks;
    for (int l = 2; l >= 0; --l) {
        const uint32_t cin = cins[l];
        const uint16_t *Aprev = l ? e->A[l - 1] : e->X0;
        LdGradT la{dH, M};
        LdConvT<PP> lb{Aprev, cin, M};
        EpAtomic ep{e->grad + offs[l], T * cin};
        hipLaunchKernelGGL((k_gemm<LdGradT, LdConvT<PP>, EpAtomic>), dim3(1, (T * cin + GN - 1) / GN, nsplit), dim3(256),
                           0, s, la, lb, ep, C, T * cin, M, ks);
        hipLaunchKernelGGL(k_bias_grad, dim3(2048), dim3(C), 0, s, dH, M, e->grad + boff[l]);
        LdConv<PP> la2{dH, C, M, 1};
        LdRows lb2{e->wt[l], T * C, cin};
        EpConvBwd<PP> ep2{l ? e->H[l - 1] : nullptr, e->blen, cin, l ? dHn : nullptr, l ? nullptr : e->dX0};
        hipLaunchKernelGGL((k_gemm<LdConv<PP>, LdRows, EpConvBwd<PP>>), dim3((M + GM - 1) / GM, (cin + GN - 1) / GN, 1),
                           dim3(256), 0, s, la2, lb2, ep2, M, cin, T * C, T * C);
        std::swap(dH, dHn);
    }
    hipLaunchKernelGGL(k_emb_grad, dim3(256), dim3(256), 0, s, e->dX0, e->btok, e->blen, M, PP, e->grad + O_EMB);
    hcheck(hipGetLastError(), "encoder backward launch");
}

void backward(idb_encoder *e, uint32_t nseq)
{
    if (e->P == 160) backward_p<160>(e, nseq);
    else backward_p<32>(e, nseq);
}

void tokenize(idb_encoder *e, const idb_dataset *d)
{
    const uint32_t n = d->n_rows;
    if (d->mem != IDB_MEM_DEVICE) throw Fail{IDB_ERR_ARG, "encoder inputs must be device memory here"};
    if (n > e->tok_cap) {
        (void)hipStreamSynchronize(e->stream);
        (void)hipFree(e->tok);
        (void)hipFree(e->tlen);
        e->tok = nullptr;
        e->tlen = nullptr;
        e->tok_cap = 0;
        hcheck(hipMalloc(&e->tok, (size_t)n * e->P), "hipMalloc");
        hcheck(hipMalloc(&e->tlen, sizeof(uint32_t) * n), "hipMalloc");
        e->tok_cap = n;
    }
    hipLaunchKernelGGL(k_tokenize, dim3(nblocks(n)), dim3(256), 0, e->stream, d->seq, d->off, d->len, n, e->vocab_d,
                       e->P, e->tok, e->tlen);
    hcheck(hipGetLastError(), "tokenize launch");
}

// Device copies of host inputs for the encoder entry points.
struct DevData {
    idb_dataset d{};
    std::vector<void *> bufs;
    ~DevData()
    {
        for (void *p : bufs) (void)hipFree(p);
    }
    template <typename T>
    const T *up(const T *h, size_t count, hipStream_t s)
    {
        void *p = nullptr;
        hcheck(hipMalloc(&p, std::max<size_t>(count * sizeof(T), 16)), "hipMalloc");
        bufs.push_back(p);
        if (count) hcheck(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s), "input copy");
        return (const T *)p;
    }
};

const idb_dataset *device_view(const idb_dataset *d, DevData &dd, hipStream_t s)
{
    if (d->mem == IDB_MEM_DEVICE) return d;
    dd.d = *d;
    dd.d.mem = IDB_MEM_DEVICE;
    dd.d.seq = dd.up(d->seq, d->seq_bytes, s);
    dd.d.off = dd.up(d->off, d->n_rows, s);
    dd.d.len = dd.up(d->len, d->n_rows, s);
    return &dd.d;
}

uint32_t round4(uint32_t n) { return (n + 3u) & ~3u; }

// One training step over pairs [first, first + nb) of the (ordered) pair list.
float step(idb_encoder *e, const uint32_t *pairs_d, const float *labels_d, const uint32_t *order_d, uint64_t first,
           uint32_t nb, uint64_t npairs, float margin, float inv_b, bool read_loss)
{
    hipStream_t s = e->stream;
    const uint32_t nseq = round4(2 * nb);
    hipLaunchKernelGGL(k_iota_rows, dim3(nblocks(nseq)), dim3(256), 0, s, 0u, 0u, nseq, e->brows);
    hipLaunchKernelGGL(k_pair_rows, dim3(nblocks(nb)), dim3(256), 0, s, pairs_d, order_d, first, nb, npairs, e->brows,
                       e->y, labels_d);
    hipLaunchKernelGGL(k_gather_rows, dim3(nblocks(nseq)), dim3(256), 0, s, e->tok, e->tlen, e->brows, nseq, e->P,
                       e->btok, e->blen);
    forward(e, nseq, true, nullptr);
    hcheck(hipMemsetAsync(e->loss, 0, sizeof(float), s), "memset");
    hcheck(hipMemsetAsync(e->dz, 0, sizeof(float) * (size_t)nseq * D, s), "memset");
    hipLaunchKernelGGL(k_pair_loss, dim3(nblocks(nb, 4)), dim3(256), 0, s, e->z, e->y, nb, margin, inv_b, e->dz, e->loss);
    backward(e, nseq);
    float l = 0.f;
    if (read_loss) {
        hcheck(hipMemcpyAsync(&l, e->loss, sizeof(float), hipMemcpyDeviceToHost, s), "loss");
        hcheck(hipStreamSynchronize(s), "sync");
    }
    return l;
}

template <class F>
int guard(idb_encoder *e, F &&f)
{
    try {
        if (e) hcheck(hipSetDevice(e->device), "hipSetDevice");
        f();
        return IDB_OK;
    } catch (const Fail &x) {
        if (e) e->err = x.msg;
        return x.code;
    } catch (const std::bad_alloc &) {
        if (e) e->err = "host allocation failed";
        return IDB_ERR_OOM;
    }
}

}  // namespace

// Embeddings (bfloat16, 64 per row) of rows [0, n) for the learned pair
// kernel; chunked so the workspace stays bounded.
hipError_t idb_encoder_embed_rows(idb_encoder *e, const uint8_t *seq, const uint64_t *off, const uint32_t *len,
                                  uint32_t n, uint16_t *zb, hipStream_t s)
{
    idb_dataset d{};
    d.n_rows = n;
    d.mem = IDB_MEM_DEVICE;
    d.seq = seq;
    d.off = off;
    d.len = len;
    const int rc = guard(e, [&] {
        hipStream_t keep = e->stream;
        e->stream = s;
        try {
            tokenize(e, &d);
            // 32768 rows of 32 positions per chunk: activations of 1M positions
            const uint32_t chunk = round4(32768u * 32u / e->P);
            ensure_workspace(e, std::min(round4(n), chunk), false);
            for (uint32_t first = 0; first < n; first += chunk) {
                const uint32_t nb = round4(std::min(chunk, n - first));
                hipLaunchKernelGGL(k_iota_rows, dim3(nblocks(nb)), dim3(256), 0, s, first, n, nb, e->brows);
                hipLaunchKernelGGL(k_gather_rows, dim3(nblocks(nb)), dim3(256), 0, s, e->tok, e->tlen, e->brows, nb,
                                   e->P, e->btok, e->blen);
                forward(e, nb, false, zb + (size_t)first * D);
            }
        } catch (...) {
            e->stream = keep;
            throw;
        }
        e->stream = keep;
    });
    return rc == IDB_OK ? hipSuccess : hipErrorUnknown;
}

const char *idb_encoder_error(const idb_encoder *e) { return e ? e->err.c_str() : "no encoder"; }
uint32_t idb_encoder_positions(const idb_encoder *e) { return e ? e->P : 0u; }
int idb_encoder_device(const idb_encoder *e) { return e ? e->device : -1; }

extern "C" {

uint32_t idb_encoder_param_count(void) { return N_PARAMS; }

int idb_encoder_create(int device, const uint8_t *vocab, uint64_t seed, idb_encoder **out)
{
    return idb_encoder_create_positions(device, vocab, IDB_ENC_P, seed, out);
}

int idb_encoder_create_positions(int device, const uint8_t *vocab, uint32_t positions, uint64_t seed,
                                 idb_encoder **out)
{
    if (!out || !vocab) return IDB_ERR_ARG;
    *out = nullptr;
    if (positions != 32 && positions != 160) return IDB_ERR_UNSUPPORTED;
    auto *e = new (std::nothrow) idb_encoder();
    if (!e) return IDB_ERR_OOM;
    e->device = device;
    e->P = positions;
    std::memcpy(e->vocab, vocab, 256);
    for (int b = 0; b < 256; ++b)
        if (e->vocab[b] >= NTOK) {
            delete e;
            return IDB_ERR_ARG;
        }
    const int rc = guard(e, [&] {
        hcheck(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), "stream");
        e->theta = dalloc<float>(e->fixed, N_PARAMS);
        e->grad = dalloc<float>(e->fixed, N_PARAMS);
        e->mom1 = dalloc<float>(e->fixed, N_PARAMS);
        e->mom2 = dalloc<float>(e->fixed, N_PARAMS);
        e->theta_b = dalloc<uint16_t>(e->fixed, N_PARAMS);
        e->vocab_d = dalloc<uint8_t>(e->fixed, 256);
        hcheck(hipMemcpyAsync(e->vocab_d, e->vocab, 256, hipMemcpyHostToDevice, e->stream), "vocab");
        uint16_t *wt = dalloc<uint16_t>(e->fixed, (size_t)C * T * (C0 + 2 * C));
        e->wt[0] = wt;
        e->wt[1] = wt + (size_t)C * T * C0;
        e->wt[2] = e->wt[1] + (size_t)C * T * C;
        // N(0, 1) embedding, U(+-1/sqrt(fan_in)) weights, zero biases
        std::vector<float> h(N_PARAMS, 0.f);
        std::mt19937_64 rng(seed);
        std::normal_distribution<float> nd(0.f, 1.f);
        for (uint32_t i = O_EMB; i < O_W1; ++i) h[i] = nd(rng);
        auto uni = [&](uint32_t o, uint32_t n, uint32_t fan) {
            std::uniform_real_distribution<float> ud(-1.f / std::sqrt((float)fan), 1.f / std::sqrt((float)fan));
            for (uint32_t i = 0; i < n; ++i) h[o + i] = ud(rng);
        };
        uni(O_W1, C * T * C0, T * C0);
        uni(O_W2, C * T * C, T * C);
        uni(O_W3, C * T * C, T * C);
        uni(O_WP, D * C, C);
        hcheck(hipMemcpyAsync(e->theta, h.data(), sizeof(float) * N_PARAMS, hipMemcpyHostToDevice, e->stream), "init");
        hcheck(hipMemsetAsync(e->mom1, 0, sizeof(float) * N_PARAMS, e->stream), "memset");
        hcheck(hipMemsetAsync(e->mom2, 0, sizeof(float) * N_PARAMS, e->stream), "memset");
        hipLaunchKernelGGL(k_to_bf16, dim3(nblocks(N_PARAMS)), dim3(256), 0, e->stream, e->theta, N_PARAMS, e->theta_b);
        refresh_transposed(e);
        hcheck(hipStreamSynchronize(e->stream), "sync");
    });
    if (rc != IDB_OK) {
        idb_encoder_destroy(e);
        return rc;
    }
    *out = e;
    return IDB_OK;
}

void idb_encoder_destroy(idb_encoder *e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (void *p : e->fixed) (void)hipFree(p);
    for (void *p : e->ws) (void)hipFree(p);
    (void)hipFree(e->tok);
    (void)hipFree(e->tlen);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char *idb_encoder_last_error(const idb_encoder *e) { return idb_encoder_error(e); }

int idb_encoder_get_params(idb_encoder *e, float *host)
{
    if (!e || !host) return IDB_ERR_ARG;
    return guard(e, [&] {
        hcheck(hipMemcpyAsync(host, e->theta, sizeof(float) * N_PARAMS, hipMemcpyDeviceToHost, e->stream), "copy");
        hcheck(hipStreamSynchronize(e->stream), "sync");
    });
}

int idb_encoder_set_params(idb_encoder *e, const float *host)
{
    if (!e || !host) return IDB_ERR_ARG;
    return guard(e, [&] {
        hcheck(hipMemcpyAsync(e->theta, host, sizeof(float) * N_PARAMS, hipMemcpyHostToDevice, e->stream), "copy");
        hipLaunchKernelGGL(k_to_bf16, dim3(nblocks(N_PARAMS)), dim3(256), 0, e->stream, e->theta, N_PARAMS, e->theta_b);
        refresh_transposed(e);
        hcheck(hipMemsetAsync(e->mom1, 0, sizeof(float) * N_PARAMS, e->stream), "memset");
        hcheck(hipMemsetAsync(e->mom2, 0, sizeof(float) * N_PARAMS, e->stream), "memset");
        e->adam_step = 0;
        hcheck(hipStreamSynchronize(e->stream), "sync");
    });
}

int idb_encoder_embed(idb_encoder *e, const idb_dataset *d, float *z_host)
{
    if (!e || !d || !z_host) return IDB_ERR_ARG;
    return guard(e, [&] {
        DevData dd;
        const idb_dataset *dv = device_view(d, dd, e->stream);
        tokenize(e, dv);
        const uint32_t n = d->n_rows, chunk = round4(32768u * 32u / e->P);
        ensure_workspace(e, std::min(round4(std::max(n, 1u)), chunk), false);
        std::vector<float> buf;
        for (uint32_t first = 0; first < n; first += chunk) {
            const uint32_t cnt = std::min(chunk, n - first), nb = round4(cnt);
            hipLaunchKernelGGL(k_iota_rows, dim3(nblocks(nb)), dim3(256), 0, e->stream, first, n, nb, e->brows);
            hipLaunchKernelGGL(k_gather_rows, dim3(nblocks(nb)), dim3(256), 0, e->stream, e->tok, e->tlen, e->brows, nb,
                               e->P, e->btok, e->blen);
            forward(e, nb, false, nullptr);
            hcheck(hipMemcpyAsync(z_host + (size_t)first * D, e->z, sizeof(float) * cnt * D, hipMemcpyDeviceToHost,
                                  e->stream), "copy");
        }
        hcheck(hipStreamSynchronize(e->stream), "sync");
    });
}

int idb_encoder_loss_grad(idb_encoder *e, const idb_dataset *d, const uint32_t *pairs, const float *labels,
                          uint32_t n_pairs, float margin, float *loss, float *grad_host)
{
    if (!e || !d || !pairs || !labels || !loss || !grad_host || n_pairs == 0) return IDB_ERR_ARG;
    return guard(e, [&] {
        for (uint32_t i = 0; i < 2 * n_pairs; ++i)
            if (pairs[i] >= d->n_rows) throw Fail{IDB_ERR_ARG, "pair row out of range"};
        DevData dd;
        const idb_dataset *dv = device_view(d, dd, e->stream);
        tokenize(e, dv);
        const uint32_t *pd = dd.up(pairs, 2 * (size_t)n_pairs, e->stream);
        const float *ld = dd.up(labels, n_pairs, e->stream);
        ensure_workspace(e, round4(2 * n_pairs), true);
        *loss = step(e, pd, ld, nullptr, 0, n_pairs, n_pairs, margin, 1.f / (float)n_pairs, true);
        hcheck(hipMemcpyAsync(grad_host, e->grad, sizeof(float) * N_PARAMS, hipMemcpyDeviceToHost, e->stream), "copy");
        hcheck(hipStreamSynchronize(e->stream), "sync");
    });
}

int idb_encoder_train(idb_encoder *e, const idb_dataset *d, const uint32_t *pairs, const float *labels,
                      uint64_t n_pairs, const idb_train_params *tp, idb_train_stats *ts)
{
    if (!e || !d || !pairs || !labels || !tp || n_pairs == 0 || tp->batch_pairs == 0 || n_pairs >= (1ull << 32))
        return IDB_ERR_ARG;
    return guard(e, [&] {
        for (uint64_t i = 0; i < 2 * n_pairs; ++i)
            if (pairs[i] >= d->n_rows) throw Fail{IDB_ERR_ARG, "pair row out of range"};
        DevData dd;
        const idb_dataset *dv = device_view(d, dd, e->stream);
        tokenize(e, dv);
        const uint32_t *pd = dd.up(pairs, 2 * (size_t)n_pairs, e->stream);
        const float *ld = dd.up(labels, n_pairs, e->stream);
        const uint32_t B = tp->batch_pairs;
        ensure_workspace(e, round4(2 * B), true);
        std::vector<uint32_t> order(n_pairs);
        std::iota(order.begin(), order.end(), 0u);
        std::mt19937_64 rng(tp->seed);
        uint32_t *od = (uint32_t *)dd.up(order.data(), n_pairs, e->stream);
        hipEvent_t t0, t1;
        hcheck(hipEventCreate(&t0), "event");
        hcheck(hipEventCreate(&t1), "event");
        hcheck(hipEventRecord(t0, e->stream), "event");
        const float b1 = 0.9f, b2 = 0.999f;
        uint32_t steps = 0;
        float first_loss = 0.f, last = 0.f;
        double epoch_sum = 0.0;
        uint32_t epoch_n = 0;
        for (uint32_t ep = 0; ep < tp->epochs; ++ep) {
            std::shuffle(order.begin(), order.end(), rng);
            hcheck(hipMemcpyAsync(od, order.data(), sizeof(uint32_t) * n_pairs, hipMemcpyHostToDevice, e->stream),
                   "order");
            // the host copy of `order` is rewritten next epoch: wait for the upload
            hcheck(hipStreamSynchronize(e->stream), "sync");
            epoch_sum = 0.0;
            epoch_n = 0;
            for (uint64_t first = 0; first < n_pairs; first += B) {
                const uint32_t nb = (uint32_t)std::min<uint64_t>(B, n_pairs - first);
                const bool rd = steps == 0 || (tp->log_every && steps % tp->log_every == 0) ||
                                (first + B >= n_pairs);
                const float l = step(e, pd, ld, od, first, nb, n_pairs, tp->margin, 1.f / (float)nb, rd);
                if (rd) {
                    if (steps == 0) first_loss = l;
                    last = l;
                    epoch_sum += l;
                    ++epoch_n;
                }
                ++e->adam_step;
                const float c1 = 1.f - std::pow(b1, (float)e->adam_step), c2 = 1.f - std::pow(b2, (float)e->adam_step);
                hipLaunchKernelGGL(k_adam, dim3(nblocks(N_PARAMS)), dim3(256), 0, e->stream, e->theta, e->grad, e->mom1,
                                   e->mom2, N_PARAMS, tp->lr, b1, b2, 1e-8f, c1, c2, e->theta_b);
                refresh_transposed(e);
                ++steps;
            }
        }
        hcheck(hipEventRecord(t1, e->stream), "event");
        hcheck(hipStreamSynchronize(e->stream), "sync");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, t0, t1);
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        if (ts) {
            ts->steps = steps;
            ts->first_loss = first_loss;
            ts->last_loss = last;
            ts->mean_sampled_loss_last_epoch = epoch_n ? (float)(epoch_sum / epoch_n) : 0.f;
            ts->train_ms = ms;
        }
    });
}

}  // extern "C"
