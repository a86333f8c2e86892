// idb_internal.h — device data layout and launch interfaces shared by the HIP
// translation units of libidb_hip.so.
//
// Layout in HBM for one run (all arrays device-resident, row = input row):
//   rows:     seq bytes, off[u64], len[u32], vbits[n*vw], jbits[n*jw]  (inputs)
//             key[u64]  duplicate hash, gslot[u32] the row's duplicate-table slot
//             rep[u32]  representative (min index of the duplicate class)
//             copies[u32], cnt[i32]  class size / exact neighbour count
//             par[u32]  union-find parent over rows, bmin[u32] border claim
//             pe[u32]   the entry that stands for the row, core flag in bit 31
//   duplicate table: P >= 2n slots of (smallest row, group size)
//   entries:  one per (unique row, V gene, J gene) bucket membership, sorted by
//             (repertoire, V gene, J gene, length, row); a bucket is a
//             contiguous run.  Per entry a 32-byte record (idb_entry, loaded
//             in one round trip) and the row's letters as dense 8-bit codes
//             (stride S bytes, S a multiple of 16) or, for Hamming mode,
//             packed words (stride HW words); for nucleotide rows a 3-gram
//             profile (64 saturating byte counts) per entry.
//   tiles:    (query block, target block) pairs of <= 64 entries each inside a
//             group (a bucket; in Hamming mode a bucket's run of one length)
//             whose length band can hold a neighbour pair.
//   edges:    neighbour pairs of the count pass in chunks of 2^chunk_log2
//             pairs; a wave fills one chunk at a time, chunk_fill[c] says how
//             many pairs chunk c holds.  Cluster growth rewrites them in place
//             as entry pairs.
//   entry forest: par_e[u32] union-find parent per entry, cmin_e[u32] the
//             smallest row of the component rooted at an entry.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "idb.h"

#define IDB_NONE 0xFFFFFFFFu
#define IDB_TILE 64               // entries per query / target block (one wave)
// Gene bitsets on the device: one header per gene name (its first 4 bytes
// and its last 4 bytes as little-endian words, zero-padded when shorter, its
// length, the offset of its 4-byte words in the packed name array).  A name
// that never matches (empty, or longer than 1024 bytes) has len = 0xFFFFFFFF.
struct idb_gene_hdr {
    uint32_t w0, w1, len, woff;
};
#define IDB_GENE_MAX_WIDTH 256u   // bit words per row (8,192 gene names)

// Entry order for runs of up to IDB_SORT_SMALL_MAX entries: stable radix
// sort of (u64 key, u32 value) pairs with the keys' low end_bit bits, two
// launches per digit pass (idb_sort.hip).  The sorted pairs end in (k1, v1)
// after an odd number of passes and back in (k0, v0) after an even one
// (*in_first); gcnt holds 256 counters per block.  hipErrorInvalidValue:
// sizes this sort does not take (nothing was launched).
#define IDB_SORT_SMALL_MAX (1u << 18)
// Single-launch exclusive scans and flag selection (idb_scan.hip).  The
// workspace (idb_scan_workspace_bytes for the largest n, zeroed when first
// allocated) holds a ticket counter and per-tile status words; launches on
// one workspace must be ordered (one stream), each with a generation the
// workspace has not seen since it was zeroed (1, 2, ...).  A look-back that
// did not end sets the 32-bit error word at idb_scan_error_offset().
size_t idb_scan_workspace_bytes(uint64_t max_n);
hipError_t idb_scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint64_t n, void *ws, uint32_t gen, hipStream_t s);
hipError_t idb_scan_exclusive_u64(const uint64_t *in, uint64_t *out, uint64_t n, void *ws, uint32_t gen, hipStream_t s);
hipError_t idb_select_flagged(const uint8_t *flags, uint32_t *out, uint32_t *count, uint64_t n, void *ws, uint32_t gen,
                              hipStream_t s);
uint32_t idb_scan_error_offset();
uint32_t idb_sort_small_blocks(uint32_t n);
hipError_t idb_sort_small(uint64_t *k0, uint64_t *k1, uint32_t *v0, uint32_t *v1, uint32_t n, uint32_t end_bit,
                          uint32_t *gcnt, hipStream_t s, bool *in_first);

#define IDB_STAGE_BYTES 32768u    // each half of a context's pinned staging area
#define IDB_XFER_MAX_BYTES (256ull << 20)  // most pinned staging a context keeps for host inputs and outputs
// The run's small device words (zeroed at the start of a run): [12] chunks
// handed out, [16, 67) the depth histogram, [100, 106) sizes for the host,
// then 8 copies of the row scan's 9 words (letter set, longest row) and 16
// partial counts each of core and noise rows (copies keep blocks from
// queueing their atomics on one word).
#define IDB_SMALL_ROWSCAN 128u
#define IDB_SMALL_CORES 200u
#define IDB_SMALL_NOISE 216u
#define IDB_SMALL_GENES 232u      // 8 copies of (highest V gene + 1, highest J gene + 1), row scan
#define IDB_SMALL_WORDS 256u
// The run's device counters: 8 64-bit statistics
#define IDB_WORK_U64 8u
// ... kept in IDB_WORK_COPIES copies (a kernel block adds into copy
// blockIdx & (copies - 1)), so that the waves' closing atomics do not queue
// on one word; readers sum the copies.
#define IDB_WORK_COPIES 16u
#define IDB_MAX_W 8               // Myers 32-bit words per pattern (<= 256 letters)
#define IDB_MAX_HW 64             // Hamming packed words per row
#define IDB_MAX_LEN 65535u
#define IDB_ENC_P 32              // letters per row the encoder reads
#define IDB_QGRAM_WORDS 16        // 3-gram profile: 64 byte counts per row

struct idb_tile {
    uint32_t q0;                  // first query entry
    uint32_t t0;                  // first target entry
    uint32_t bucket;              // bucket id (gene pair) within its repertoire
    uint32_t shape;               // qn | tn << 8 | diag << 16
};

// Everything the pair kernels need about one entry, in one 32-byte record.
struct __align__(16) idb_entry {
    uint32_t row;                 // input row (a unique representative)
    uint32_t cp;                  // copies of the row
    uint32_t len_chk;             // length | chk << 16 (chk: the row may share a lower gene,
                                  // or every row in single-bucket mode)
    uint32_t t2;                  // threshold at this length: count pass (int16, low half),
                                  // fuse pass (int16, high half)
    uint32_t len;                 // the length alone (the second half is the count pass's
                                  // query record as it stands)
    uint32_t pres1, pres2;        // letters present at least once (the absent-letter filter) / twice
    uint32_t sig;                 // residues mod 16 of the genes the canonical test looks at:
                                  // V in the low half, J in the high half
};

// Uniform parameters of a pair-kernel launch.
struct idb_pair_args {
    const idb_tile *tiles;
    uint32_t n_tiles;
    const idb_entry *ent;
    const uint8_t *e_codes;       // Myers: stride S bytes of codes
    const uint32_t *e_pk;         // Hamming: stride HW packed words
    uint32_t S, HW, K;
    uint32_t bits;                // Hamming code width: 2, 5 or 8
    const int32_t *vbits, *jbits;
    uint32_t vw, jw;
    uint32_t gj;                  // J genes per bucket row (32*jw)
    uint32_t single_bucket;       // 1: every pair is gated explicitly
    uint32_t band_filter;         // 1: absent-letter filter before Myers (alphabets <= 31 letters)
    uint32_t qgram_filter;        // 1: 3-gram profile filter before Myers (alphabets <= 4 letters)
    const uint32_t *e_qg;         // 3-gram profiles, IDB_QGRAM_WORDS per entry
    const uint4 *e_low;           // per entry: its genes below the bucket's (V words, then J words), or null
    int32_t *cnt;                 // count pass output
    const int32_t *core;          // fuse pass input (per row)
    uint32_t *par;                // fuse pass: union-find parents
    uint32_t *bmin;               // fuse pass: border claims
    unsigned long long *work;     // [0] gated pairs, [1] neighbour pairs, [2] fuse pairs, [3] gated letters,
                                  // [4] edges, [5] Myers pairs, [6] Myers letters
    uint2 *edges;                 // count pass: neighbour pairs (x, y) in chunks, may be null
    uint32_t *chunk_fill;         // pairs written into each chunk
    uint32_t *n_chunks;           // chunks handed out (counts past max_chunks on overflow)
    uint32_t max_chunks;          // capacity in chunks
    uint32_t chunk_log2;          // pairs per chunk = 1 << chunk_log2 (>= IDB_TILE)
    uint32_t fuse_if_overflow;    // fuse pass: run only when the pair list overflowed
    int32_t t_const;              // >= 0: every pair's count-pass threshold (absolute mode); -1 otherwise
    const uint16_t *zb;           // learned filter: bfloat16 embedding (64) per row
    float min_sim;                // learned filter: proposal threshold on z_x . z_y
    // launch shape (0 = default): waves per block, grid = resident blocks x
    // grid_mul, and plain block order instead of the XCD-grouped one
    uint32_t tune_wpb, tune_grid_mul, no_xcd_order;
};

// LDS bytes one wave of the Myers kernels needs: Peq columns, the edge
// staging slots, the Myers queue (2 * IDB_TILE 16-bit entries), the per-tile
// count accumulators of queries and targets, and the tile's query records.
__host__ __device__ inline size_t idb_myers_lds_per_wave(uint32_t K, uint32_t W)
{
    return ((size_t)K * W * 64 + 2 * IDB_TILE + IDB_TILE + 2 * IDB_TILE) * sizeof(uint32_t) +
           IDB_TILE * sizeof(idb_entry);
}
// Extra LDS per wave for the query profiles of the 3-gram filter.
__host__ __device__ inline size_t idb_qgram_lds_per_wave() { return (size_t)IDB_TILE * IDB_QGRAM_WORDS * sizeof(uint32_t); }
// The device's LDS limit per block (opt-in value where the device has one).
size_t idb_lds_limit();

// Kernel launchers (defined in idb_pairs.hip).
hipError_t idb_launch_count_myers(const idb_pair_args &a, uint32_t W, hipStream_t s);
hipError_t idb_launch_fuse_myers(const idb_pair_args &a, uint32_t W, hipStream_t s);
hipError_t idb_launch_count_hamming(const idb_pair_args &a, hipStream_t s);
hipError_t idb_launch_fuse_hamming(const idb_pair_args &a, hipStream_t s);
hipError_t idb_launch_count_learned(const idb_pair_args &a, hipStream_t s);
hipError_t idb_launch_pair_distances(const uint8_t *seq, const uint64_t *off, const uint32_t *len, const uint2 *pairs,
                                     unsigned long long n, uint32_t W, uint32_t ham, uint32_t *dist, hipStream_t s);
hipError_t idb_launch_gene_bitsets(const uint8_t *calls, const uint64_t *off, const uint32_t *len, uint32_t n,
                                   const idb_gene_hdr *hdr, const uint32_t *gw, uint32_t n_words, uint32_t n_genes,
                                   int32_t *out, uint32_t width, hipStream_t s);
// Cluster growth over the count pass's pairs (rewrites the pair list in place).
struct idb_union_args {
    const uint32_t *pe;           // per row: one of its entries | core flag << 31; IDB_NONE without entries
    uint32_t *par_e;              // per entry: union-find parent, its own index on entry
    uint32_t *cmin_e;             // per entry: smallest row of the component rooted there, IDB_NONE on entry
    uint32_t n_entries, n_rows;
    uint32_t *par;                // per row: set to the component minimum for core rows
    uint32_t *bmin;               // per row: border claims
    const uint32_t *rep;          // per row: representative (for the depth histogram)
    uint32_t *depth_hist;         // 51 bins of union-find depths, or null (no histogram)
};
hipError_t idb_launch_union_edges(const idb_pair_args &a, const idb_union_args &u, hipStream_t s);

// Encoder (idb_learned.hip): bfloat16 embeddings of rows [0, n) on stream s.
struct idb_encoder;
hipError_t idb_encoder_embed_rows(idb_encoder *e, const uint8_t *seq, const uint64_t *off, const uint32_t *len,
                                  uint32_t n, uint16_t *zb, hipStream_t s);
const char *idb_encoder_error(const idb_encoder *e);
int idb_encoder_device(const idb_encoder *e);
