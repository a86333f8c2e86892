/*
 * idb.h — C ABI of libidb_hip.so, the MI355X (gfx950) implementation of
 * immunodbscan's DBSCAN clone-collapsing path.
 *
 * The reference crosses from host C++ into its device code through one class,
 * InterlinkClass (cuda/Interlink.hpp:46-69, cuda/Interlink.cu:34-95).  Its
 * constructor turns gene calls into bitsets, packs the sequences, runs
 * DBSCAN::doScan (cuda/PSDBSCAN.hcu:1300-1356) and copies clone ids, core
 * flags, uniques and copies back.  The entry points below are what a binding
 * of that class needs: the gene-bitset builder (bitset_container,
 * cuda/BitsetVector.inl:27-45) and the clustering run.  A drop-in C++
 * InterlinkClass built on them is in immunodbscan_amd/csrc/Interlink.{hpp,cpp}.
 *
 * Conventions: every function returns IDB_OK (0) or a negative IDB_ERR_*
 * code; idb_last_error() then describes the failure.  Calls are synchronous.
 * One context per host thread.  Output buffers are owned by the caller.
 */
#ifndef IDB_H
#define IDB_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IDB_OK 0
#define IDB_ERR_ARG (-1)          /* bad argument or inconsistent sizes          */
#define IDB_ERR_HIP (-2)          /* a HIP runtime call failed                    */
#define IDB_ERR_OOM (-3)          /* device allocation failed or cap exceeded     */
#define IDB_ERR_UNSUPPORTED (-4)  /* input outside the supported envelope         */

#define IDB_MEM_HOST 0            /* arrays are host memory                       */
#define IDB_MEM_DEVICE 1          /* arrays are device memory on the ctx device   */

#define IDB_N_STAGES 8

typedef struct idb_ctx idb_ctx;

/* Rows after the reference's empty-cell removal (cpp/TableHolder.hpp:283-285),
 * in table order.  Row i's sequence is seq[off[i] .. off[i]+len[i]); its V and
 * J gene bitsets are v_width / j_width int32 words, gene g in word g>>5 at bit
 * g&31 (cuda/BitsetVector.inl:34,73-82). */
typedef struct {
    uint32_t n_rows;
    uint32_t mem;                 /* IDB_MEM_HOST or IDB_MEM_DEVICE */
    const uint8_t *seq;
    uint64_t seq_bytes;
    const uint64_t *off;
    const uint32_t *len;
    const int32_t *vbits;
    uint32_t v_width;
    const int32_t *jbits;
    uint32_t j_width;
} idb_dataset;

/* The four hot-path settings of cpp/ConfigHolder.hpp:181-199, plus an absolute
 * threshold mode (t = abs_k edits for every pair) that the reference lacks. */
typedef struct {
    float eps;                    /* maximum_relative_edit_distance; t = trunc(max(lx,ly)*eps) */
    int32_t abs_k;                /* >= 0 selects the absolute mode; -1 = relative             */
    uint32_t min_pts;             /* minimum_number_of_neighbors                               */
    uint32_t use_hamming;         /* use_hamming                                               */
    uint32_t tree_stats;          /* get_tree_depth_statistics                                 */
    uint32_t debug_hash_bits;     /* 0 = 32-bit duplicate hash; 1..64 keeps that many bits
                                     (few bits force collisions in tests)                     */
    uint32_t edge_capacity;       /* 0 = automatic; else the most neighbour pairs kept from the
                                     count pass (more fall back to a second pass over pairs)  */
    uint32_t flags;               /* IDB_FLAG_* below (0: plain launches)                      */
} idb_params;

/* flags: AUTOTUNE times the count kernel's launch shapes (waves per block,
 * grid size, XCD block order) on the first run of each kernel and shape class
 * in a context and keeps the fastest; GRAPH captures everything after the
 * sizing steps (count pass, core flags, cluster growth, renumbering, outputs)
 * into one HIP graph per context, updated in place on later runs, with the
 * union / fuse choice made on the device.  Neither changes any result. */
#define IDB_FLAG_AUTOTUNE 1u
#define IDB_FLAG_GRAPH 2u
/* FORCE_LARGE_PATHS takes, at any size, the code paths that otherwise only
 * large runs take: the library radix sort of the entry order (used above 2^18
 * entries) and 1,024-pair chunks of the neighbour-pair list (used for lists
 * of 2^24 pairs or more).  Results are unchanged; it lets tests check those
 * paths on inputs that the CPU oracle finishes in seconds. */
#define IDB_FLAG_FORCE_LARGE_PATHS 4u
/* Off-diagonal tiles of more than 32 query entries run as two tiles of 32
 * query rows each (the heavy tiles of dense buckets then spread over two
 * waves); results are unchanged. */
#define IDB_FLAG_SPLIT_TILES 8u

/* Per-row outputs (cuda/Interlink.hpp:48-49 and :59-67).  uniques holds the
 * ascending representative indices in its first n_uniques entries and 0 after.
 * neighbour_count (optional, may be NULL) receives the exact copies-weighted
 * neighbour count of every representative and 0 for duplicate rows. */
typedef struct {
    uint32_t mem;                 /* IDB_MEM_HOST or IDB_MEM_DEVICE */
    uint32_t *clone_id;
    int32_t *core;
    uint32_t *copies;
    uint32_t *uniques;
    uint32_t *neighbour_count;
} idb_outputs;

/* Counters the reference writes to its log (cuda/PSDBSCAN.hcu:1113,1198,1279,1290)
 * plus work counters and stage times.  Stage order matches the reference's
 * timer names: duplicates, duplicate-tree reduction, copies, core points,
 * grow clusters, cluster-tree reduction, renumber, whole run. */
typedef struct {
    uint32_t n_rows, n_uniques, n_cores, n_clusters, n_noise;
    uint32_t max_len, alphabet_size, n_buckets;
    uint32_t fuse_from_edges;     /* 1: clusters grown from the count pass's pair list       */
    uint32_t band_filter;         /* exact bound before Myers: 1 absent letters, 2 3-grams   */
    uint64_t bucket_entries, tiles;
    uint64_t gated_pairs;         /* pairs reaching a distance kernel (count pass)   */
    uint64_t neighbour_pairs;     /* unordered unique pairs with N(x,y)              */
    uint64_t fuse_pairs;          /* pairs evaluated by the fuse pass                */
    uint64_t gated_letters;       /* text letters of the count pass's gated pairs    */
    uint64_t myers_pairs;         /* gated pairs that reached Myers (after the band filter) */
    uint64_t myers_letters;       /* text letters those Myers runs consumed          */
    float count_kernel_ms;        /* count-pass kernel, HIP events on the library stream */
    float fuse_kernel_ms;         /* fuse-pass kernel                                */
    uint32_t tree_depth[51];      /* union-find depth histogram when tree_stats      */
    float stage_ms[IDB_N_STAGES]; /* with IDB_FLAG_GRAPH the graph's time is the core-point stage's */
    uint32_t tuned_wpb, tuned_grid_mul;  /* count-kernel launch shape: waves per block, grid / resident */
    uint32_t tuned_xcd_order;     /* bit 0: XCD-grouped block order */
    uint32_t tuned_trials;        /* shapes the autotuner timed                      */
    uint32_t graph_used;          /* 1: the tail ran as one graph launch             */
    float encoder_ms;             /* learned filter: embedding of every row (HIP events) */
} idb_stats;

/* Context on HIP device `device`: a stream and a reusable device arena.
 * Replaces the implicit device-0 setup of DBSCAN (cuda/PSDBSCAN.hcu:962-990). */
int idb_create(idb_ctx **ctx, int device);
void idb_destroy(idb_ctx *ctx);
const char *idb_last_error(const idb_ctx *ctx);

/* R2 gene bitsets on the host: bit g of row i is set iff gene name g occurs as
 * a substring of call string i.  Replaces bitset_container's constructor and
 * its per-gene find_string_occurences launches (cuda/BitsetVector.inl:27-45,
 * 118-209).  Strings are given as concatenated bytes with offsets/lengths;
 * out has n * width words, width = ceil(n_genes / 32). */
int idb_gene_bitsets(uint32_t n, const uint8_t *calls, const uint64_t *call_off, const uint32_t *call_len,
                     uint32_t n_genes, const uint8_t *genes, const uint64_t *gene_off, const uint32_t *gene_len,
                     int32_t *out, uint32_t width);

/* R2 gene bitsets on the device, with the same result as idb_gene_bitsets:
 * one wave per call string tests every gene name.  Replaces
 * bitset_container's constructor and its per-gene find_string_occurences
 * launches (cuda/BitsetVector.inl:27-45, 118-209).  calls, call_off,
 * call_len and out are in host memory (mem = IDB_MEM_HOST) or in device
 * memory on the ctx device (IDB_MEM_DEVICE); the gene names are host arrays.
 * width >= ceil(n_genes / 32) and at most 256 (IDB_ERR_UNSUPPORTED above). */
int idb_gene_bitsets_gpu(idb_ctx *ctx, uint32_t mem, uint32_t n, const uint8_t *calls, const uint64_t *call_off,
                         const uint32_t *call_len, uint32_t n_genes, const uint8_t *genes, const uint64_t *gene_off,
                         const uint32_t *gene_len, int32_t *out, uint32_t width);

/* One repertoire: duplicates, copies, core points, clusters and renumbering
 * (DBSCAN::doScan, cuda/PSDBSCAN.hcu:1300-1356, as driven by
 * InterlinkClass::InterlinkClass, cuda/Interlink.cu:34-95).  stats may be NULL. */
int idb_dbscan(idb_ctx *ctx, const idb_dataset *data, const idb_params *params, idb_outputs *out,
               idb_stats *stats);

/* Allocation-free variant of idb_dbscan, with the same arguments and the
 * same results.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (serving loops, harnesses that forbid
 * hipMalloc).  Every device buffer of the run (input and output staging,
 * duplicate table, bucket entries, tiles, neighbour-pair list, scan space)
 * comes from one static device array of IDB_NOALLOC_SCRATCH_BYTES per GPU,
 * which the library holds until the process ends; the call itself makes no
 * hipMalloc, hipFree or other allocating call.  Limits: at most
 * IDB_NOALLOC_MAX_ROWS rows (IDB_ERR_UNSUPPORTED above), and inputs whose
 * entries and tiles do not fit in the scratch fail with IDB_ERR_UNSUPPORTED;
 * the neighbour-pair list takes what is left (a shorter list only sends
 * cluster growth to the second pass over the tiles).  params->flags must be
 * 0 (autotuning and graph capture allocate).  Every call on one GPU shares
 * that GPU's scratch array: calls that overlap on one GPU (two contexts on
 * two host threads) race and must be serialised by the caller. */
#define IDB_NOALLOC_MAX_ROWS 65536u
#define IDB_NOALLOC_SCRATCH_BYTES (64ull << 20)
int idb_dbscan_noalloc(idb_ctx *ctx, const idb_dataset *data, const idb_params *params, idb_outputs *out,
                       idb_stats *stats);

/* Many repertoires in one pass: data holds the rows of all repertoires back to
 * back, repertoire r being rows rep_start[r] .. rep_start[r+1]-1.  Results
 * equal n_reps separate idb_dbscan calls: clone ids restart at 1, and uniques
 * holds repertoire-local indices at each repertoire's row offset, with the
 * per-repertoire count in n_uniques[r].  No reference counterpart: the
 * reference runs one table per process (cpp/psdbscan.cpp:25-106). */
int idb_dbscan_batch(idb_ctx *ctx, const idb_dataset *data, const uint32_t *rep_start, uint32_t n_reps,
                     const idb_params *params, idb_outputs *out, uint32_t *n_uniques, idb_stats *stats);

/* ------------------------------------------------------------------------
 * Exact pair evaluation.  No reference counterpart as entry points; they
 * expose the reference's own neighbour predicate (R6/R7) for labelling and
 * measurement.
 *
 * idb_pair_distances: the unit-cost Levenshtein (or, with use_hamming, the
 * Hamming distance, UINT32_MAX for unequal lengths) of each row pair
 * pairs[2i], pairs[2i+1], computed by the device's Myers kernel.
 * idb_neighbour_pairs: the unordered unique-row pairs (x < y) with N(x, y)
 * found by the exact count pass, at most cap of them (*n_pairs gets the total;
 * IDB_ERR_OOM when it exceeds cap). */
int idb_pair_distances(idb_ctx *ctx, const idb_dataset *data, const uint32_t *pairs, uint64_t n_pairs,
                       uint32_t use_hamming, uint32_t *dist);
int idb_neighbour_pairs(idb_ctx *ctx, const idb_dataset *data, const idb_params *params, uint32_t *pairs,
                        uint64_t cap, uint64_t *n_pairs);

/* ------------------------------------------------------------------------
 * Learned candidate filter.  No reference counterpart.  A small 1-D CNN
 * (token embedding 32 -> 3 x conv1d(k=5, 128 channels, GELU) -> mean over the
 * row's letters -> linear 128 -> 64 -> L2 norm; bfloat16 weights, float32
 * accumulation) maps each row (its first P letters, P = 32 or 160) to a unit
 * vector.  The
 * learned clustering run proposes, inside the same V/J bucket tiles as the
 * exact run, the pairs whose embedding dot product reaches min_similarity,
 * and verifies each proposal with the exact Myers kernel: distances are never
 * approximated, only pairs the encoder fails to propose can be missed. */
#define IDB_ENC_PARAMS 193984u    /* float parameters of the encoder            */

typedef struct idb_encoder idb_encoder;

typedef struct {
    uint32_t epochs;
    uint32_t batch_pairs;         /* pairs per optimiser step                     */
    float lr;                     /* Adam learning rate (beta 0.9 / 0.999)        */
    float margin;                 /* contrastive margin on |z_a - z_b|            */
    uint64_t seed;                /* pair order shuffling                         */
    uint32_t log_every;           /* read the loss back every this many steps (0: first/last only) */
    uint32_t reserved;
} idb_train_params;

typedef struct {
    uint32_t steps;
    float first_loss, last_loss;
    float mean_sampled_loss_last_epoch;
    float train_ms;
} idb_train_stats;

/* vocab maps every byte value to a token id below 32.  Parameters are
 * initialised from seed (N(0,1) embedding, U(+-1/sqrt(fan_in)) weights). */
int idb_encoder_create(int device, const uint8_t *vocab, uint64_t seed, idb_encoder **enc);
/* The same with P positions per row: 32 (protein rows) or 160 (nucleotide
 * rows of up to 160 letters); the parameters do not depend on P.  A learned
 * run needs rows of at most P letters. */
int idb_encoder_create_positions(int device, const uint8_t *vocab, uint32_t positions, uint64_t seed,
                                 idb_encoder **enc);
uint32_t idb_encoder_positions(const idb_encoder *enc);
void idb_encoder_destroy(idb_encoder *enc);
const char *idb_encoder_last_error(const idb_encoder *enc);
uint32_t idb_encoder_param_count(void);
/* Flat float32 parameters in the order emb[32][32], w1[128][5*32], b1[128],
 * w2[128][5*128], b2, w3[128][5*128], b3, wp[64][128], bp[64]; convolution
 * weights are out-channel major with k = tap * c_in + c_in_index. */
int idb_encoder_get_params(idb_encoder *enc, float *params);
int idb_encoder_set_params(idb_encoder *enc, const float *params);
/* z (n_rows x 64 floats) of every row of data. */
int idb_encoder_embed(idb_encoder *enc, const idb_dataset *data, float *z);
/* Contrastive loss over the pairs (label 1: neighbours, 0: not) and its
 * gradient with respect to the flat parameters; no update. */
int idb_encoder_loss_grad(idb_encoder *enc, const idb_dataset *data, const uint32_t *pairs, const float *labels,
                          uint32_t n_pairs, float margin, float *loss, float *grad);
/* Adam training over shuffled batches of the labelled pairs, on the device. */
int idb_encoder_train(idb_encoder *enc, const idb_dataset *data, const uint32_t *pairs, const float *labels,
                      uint64_t n_pairs, const idb_train_params *train, idb_train_stats *stats);

/* idb_dbscan with the learned candidate filter (Levenshtein mode, rows of at
 * most 32 letters).  The encoder must live on the context's device. */
int idb_dbscan_learned(idb_ctx *ctx, idb_encoder *enc, float min_similarity, const idb_dataset *data,
                       const idb_params *params, idb_outputs *out, idb_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
