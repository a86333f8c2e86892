/*
 * oracle.h — CPU restatement of immunodbscan's DBSCAN clone-collapsing
 * semantics.  TEST INFRASTRUCTURE ONLY: this library is the checker for the
 * HIP path (tests, smoke check and the CPU-baseline leg of the benchmark).
 * No product code links, loads or calls it.
 *
 * Two implementations of the same contract live in oracle.c:
 *   orc_ref_dbscan   literal restatement of DBSCAN::doScan
 *                    (cuda/PSDBSCAN.hcu:1300-1356): all pairs, the
 *                    reference's gate order, Wagner-Fischer / Hamming,
 *                    sequential union-find.  It is the timed CPU baseline.
 *   orc_fast_dbscan  bucketed restatement (canonical V/J buckets, length
 *                    band, banded DP with cut-off) used for parity at sizes
 *                    the all-pairs version cannot finish; tests pin it
 *                    bit-equal to orc_ref_dbscan.
 */
#ifndef IDB_ORACLE_H
#define IDB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t n;               /* rows (R1: rows kept after empty-cell removal) */
    const uint8_t *seq;       /* concatenated sequence bytes                   */
    const uint64_t *off;      /* byte offset of row i in seq                   */
    const uint32_t *len;      /* length of row i                               */
    const int32_t *vbits;     /* n * vw words, row-major (R2 layout)           */
    uint32_t vw;
    const int32_t *jbits;     /* n * jw words                                  */
    uint32_t jw;
} orc_data;

typedef struct {
    float eps;                /* maximum_relative_edit_distance                */
    int32_t abs_k;            /* >= 0: absolute threshold t = abs_k            */
    uint32_t min_pts;         /* minimum_number_of_neighbors                   */
    int32_t use_hamming;      /* use_hamming                                   */
    int32_t nthreads;         /* OpenMP threads (<= 0: default)                */
} orc_params;

typedef struct {
    uint32_t *clone_id;       /* n */
    int32_t *core;            /* n */
    uint32_t *copies;         /* n */
    uint32_t *uniques;        /* n, first n_uniques valid, rest 0 */
    uint32_t *counts;         /* n, exact copies-weighted counts (0 for non-reps); may be NULL */
    uint32_t n_uniques;
    uint32_t n_cores;
    uint32_t n_clusters;
    uint32_t n_noise;
    uint64_t n_edges;         /* unordered unique pairs with N(x,y) */
} orc_result;

/* R5 threshold t(lx,ly) with the device's saturating float->int rule. */
int32_t orc_threshold(uint32_t lx, uint32_t ly, float eps, int32_t abs_k);
/* Unit-cost distances over raw bytes (R7). */
uint32_t orc_levenshtein(const uint8_t *a, uint32_t la, const uint8_t *b, uint32_t lb);
uint32_t orc_hamming(const uint8_t *a, const uint8_t *b, uint32_t l);
/* Banded decision lev(a,b) <= t with cut-off (used by the bucketed oracle). */
int32_t orc_lev_within(const uint8_t *a, uint32_t la, const uint8_t *b, uint32_t lb, int32_t t);
/* R2 gene bitsets: bit g of row i set iff gene g occurs in call i. */
void orc_gene_bitsets(uint32_t n, const uint8_t *calls, const uint64_t *call_off, const uint32_t *call_len,
                      uint32_t n_genes, const uint8_t *genes, const uint64_t *gene_off,
                      const uint32_t *gene_len, int32_t *out, uint32_t width);
/* R8 neighbour predicates as the count pass and the fuse pass evaluate them. */
int32_t orc_neighbour_count_pass(const orc_data *d, const orc_params *p, uint32_t x, uint32_t y);
int32_t orc_neighbour_fuse_pass(const orc_data *d, const orc_params *p, uint32_t x, uint32_t y);

int32_t orc_ref_dbscan(const orc_data *d, const orc_params *p, orc_result *r);
/* Facts for the sampled rows xs[0 .. n_x) of an input too large to cluster
 * on the CPU: rep[s] = smallest row equal to xs[s] (R3/R4), copies[s] = its
 * class size, count[s] = |class| + rows outside the class that are count-pass
 * neighbours (the exact R8 count when xs[s] is a representative), and those
 * neighbour rows in nbr[nbr_off[s] .. nbr_off[s+1]) (nbr_off has n_x + 1
 * entries; rows past nbr_cap are not written, nbr_off[n_x] is the total).
 * Returns -1 for a sample row out of range. */
int32_t orc_sample_rows(const orc_data *d, const orc_params *p, const uint32_t *xs, uint32_t n_x, uint32_t *rep,
                        uint32_t *copies, uint32_t *count, uint64_t *nbr_off, uint32_t *nbr, uint64_t nbr_cap);
int32_t orc_fast_dbscan(const orc_data *d, const orc_params *p, orc_result *r);

#ifdef __cplusplus
}
#endif
#endif
