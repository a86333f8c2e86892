/*
 * oracle.c — CPU restatement of immunodbscan's DBSCAN clone-collapsing path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.
 * Nothing in the product (immunodbscan_amd/, include/) links or calls it.
 *
 * Semantics followed (all citations relative to the reference repository):
 *   R2  gene bitsets        cuda/BitsetVector.inl:27-45, 118-209
 *   R3  exact duplicates    cuda/PSDBSCAN.hcu:525-573, cuda/BitsetVector.inl:340-398
 *   R4  reps and copies     cuda/PSDBSCAN.hcu:76-107, 129-149, 1074-1085
 *   R5  threshold           cuda/SpecialFunctors.hcu:245-274 (used at PSDBSCAN.hcu:613,695,783,876)
 *   R6  V/J gate            cuda/BitsetVector.inl:214-237, 270-334
 *   R7  neighbour predicate cuda/PSDBSCAN.hcu:585-657 (Hamming), 751-833 (Levenshtein),
 *                           distances at :246-290 and :410-428
 *   R8  counts, core flags  cuda/PSDBSCAN.hcu:1097-1184
 *   R9  core components     cuda/PSDBSCAN.hcu:76-107, 664-742, 840-931
 *   R10 border points       cuda/PSDBSCAN.hcu:113-122, 915-924 (canonical rule: the
 *                           minimum-index core neighbour, which is what a sequential
 *                           run of the fuse loop in index order yields)
 *   R11 duplicates inherit  cuda/PSDBSCAN.hcu:1344-1347
 *   R12 labels              cuda/PSDBSCAN.hcu:1263-1294
 *
 * Parity pins (see tests/test_oracle.py): Wagner-Fischer against BFS over edit
 * operations and a simulation of the reference's anti-diagonal indexing;
 * threshold against the float32 probe table; bitsets against substring search
 * and the shipped gene lists; clustering against scikit-learn's DBSCAN, scipy's
 * connected components, hand-derived tiny cases and invariants.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_NONE 0xFFFFFFFFu

/* ------------------------------------------------------------------------ */
/* R5: the reference evaluates truncf(max(float(lx), float(ly)) * eps) on the  */
/* device and converts to int (SpecialFunctors.hcu:257-261).  The device      */
/* conversion saturates and maps NaN to 0; this helper states that rule.      */
static int32_t cvt_rzi_sat(float f)
{
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}

int32_t orc_threshold(uint32_t lx, uint32_t ly, float eps, int32_t abs_k)
{
    if (abs_k >= 0) return abs_k;
    float fx = (float)lx, fy = (float)ly;
    volatile float m = fx > fy ? fx : fy;
    volatile float prod = m * eps;
    return cvt_rzi_sat(truncf(prod));
}

/* ------------------------------------------------------------------------ */
/* R7 distances over raw bytes.                                               */
uint32_t orc_levenshtein(const uint8_t *a, uint32_t la, const uint8_t *b, uint32_t lb)
{
    /* Wagner-Fischer, two rows.  Equal to the reference's anti-diagonal DP
     * (PSDBSCAN.hcu:246-290), which fills the same recurrence in diagonal order. */
    uint32_t *row = (uint32_t *)malloc(sizeof(uint32_t) * (lb + 1));
    for (uint32_t j = 0; j <= lb; ++j) row[j] = j;
    for (uint32_t i = 1; i <= la; ++i) {
        uint32_t diag = row[0];
        row[0] = i;
        for (uint32_t j = 1; j <= lb; ++j) {
            uint32_t up = row[j];
            uint32_t best = diag + (a[i - 1] != b[j - 1]);
            if (up + 1 < best) best = up + 1;
            if (row[j - 1] + 1 < best) best = row[j - 1] + 1;
            row[j] = best;
            diag = up;
        }
    }
    uint32_t r = row[lb];
    free(row);
    return r;
}

uint32_t orc_hamming(const uint8_t *a, const uint8_t *b, uint32_t l)
{
    uint32_t h = 0;
    for (uint32_t i = 0; i < l; ++i) h += a[i] != b[i];
    return h;
}

int32_t orc_lev_within(const uint8_t *a, uint32_t la, const uint8_t *b, uint32_t lb, int32_t t)
{
    /* Ukkonen band |i-j| <= t with values capped at t+1 and row cut-off. */
    if (t < 0) return 0;
    int64_t d = (int64_t)la - (int64_t)lb;
    if (d < 0) d = -d;
    if (d > t) return 0;
    if ((uint64_t)t >= (uint64_t)la + lb) return 1;
    const uint32_t cap = (uint32_t)t + 1;
    uint32_t *prev = (uint32_t *)malloc(sizeof(uint32_t) * (lb + 2));
    uint32_t *cur = (uint32_t *)malloc(sizeof(uint32_t) * (lb + 2));
    for (uint32_t j = 0; j <= lb; ++j) prev[j] = j < cap ? j : cap;
    int ok = 1;
    for (uint32_t i = 1; i <= la && ok; ++i) {
        int64_t jlo = (int64_t)i - t, jhi = (int64_t)i + t;
        if (jlo < 0) jlo = 0;
        if (jhi > lb) jhi = lb;
        for (uint32_t j = 0; j <= lb; ++j) cur[j] = cap;
        uint32_t rowmin = cap;
        for (int64_t jj = jlo; jj <= jhi; ++jj) {
            uint32_t j = (uint32_t)jj, v;
            if (j == 0) {
                v = i < cap ? i : cap;
            } else {
                v = prev[j - 1] + (a[i - 1] != b[j - 1]);
                if (prev[j] + 1 < v) v = prev[j] + 1;
                if (cur[j - 1] + 1 < v) v = cur[j - 1] + 1;
                if (v > cap) v = cap;
            }
            cur[j] = v;
            if (v < rowmin) rowmin = v;
        }
        if (rowmin > (uint32_t)t) ok = 0;
        uint32_t *tmp = prev; prev = cur; cur = tmp;
    }
    int32_t res = ok && prev[lb] <= (uint32_t)t;
    free(prev);
    free(cur);
    return res;
}

/* ------------------------------------------------------------------------ */
/* R2: bit g of row i is set iff gene name g occurs as a contiguous substring  */
/* of call i (BitsetVector.inl:151-175).  The reference launches one block of  */
/* len(gene) threads per gene (:204); a length of 0 or more than 1024 makes    */
/* that launch fail and leaves the bit clear, which is reproduced here.        */
void orc_gene_bitsets(uint32_t n, const uint8_t *calls, const uint64_t *call_off, const uint32_t *call_len,
                      uint32_t n_genes, const uint8_t *genes, const uint64_t *gene_off,
                      const uint32_t *gene_len, int32_t *out, uint32_t width)
{
    memset(out, 0, sizeof(int32_t) * (size_t)n * width);
    for (uint32_t g = 0; g < n_genes; ++g) {
        uint32_t gl = gene_len[g];
        if (gl == 0 || gl > 1024) continue;
        const uint8_t *gs = genes + gene_off[g];
        for (uint32_t i = 0; i < n; ++i) {
            const uint8_t *cs = calls + call_off[i];
            uint32_t cl = call_len[i];
            int found = 0;
            for (uint32_t p = 0; p + gl <= cl && !found; ++p)
                found = memcmp(cs + p, gs, gl) == 0;
            if (found) out[(size_t)i * width + (g >> 5)] |= (int32_t)(1u << (g & 31));
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Pair predicates.                                                           */
static inline const uint8_t *row_seq(const orc_data *d, uint32_t i) { return d->seq + d->off[i]; }

static int gate_vj(const orc_data *d, uint32_t x, uint32_t y)
{
    /* R6: OR over words of (a & b) must be non-zero for V and for J. */
    int32_t cv = 0, cj = 0;
    for (uint32_t w = 0; w < d->vw; ++w) cv |= d->vbits[(size_t)x * d->vw + w] & d->vbits[(size_t)y * d->vw + w];
    for (uint32_t w = 0; w < d->jw; ++w) cj |= d->jbits[(size_t)x * d->jw + w] & d->jbits[(size_t)y * d->jw + w];
    return cv != 0 && cj != 0;
}

static int is_duplicate(const orc_data *d, uint32_t q, uint32_t s)
{
    /* R3: equal length, equal V words, equal J words, equal bytes
     * (PSDBSCAN.hcu:542-565).  The 32-bit hash there is only a prefilter. */
    if (d->len[q] != d->len[s]) return 0;
    if (memcmp(d->vbits + (size_t)q * d->vw, d->vbits + (size_t)s * d->vw, sizeof(int32_t) * d->vw)) return 0;
    if (memcmp(d->jbits + (size_t)q * d->jw, d->jbits + (size_t)s * d->jw, sizeof(int32_t) * d->jw)) return 0;
    return memcmp(row_seq(d, q), row_seq(d, s), d->len[q]) == 0;
}

int32_t orc_neighbour_count_pass(const orc_data *d, const orc_params *p, uint32_t x, uint32_t y)
{
    uint32_t lx = d->len[x], ly = d->len[y];
    int32_t t = orc_threshold(lx, ly, p->eps, p->abs_k);
    if (p->use_hamming) {
        /* half_hamming_core_points, PSDBSCAN.hcu:609-651: the running count
         * is compared with t only after a mismatch. */
        if (lx != ly || !gate_vj(d, x, y)) return 0;
        const uint8_t *a = row_seq(d, x), *b = row_seq(d, y);
        int64_t ham = 0;
        int result = 1;
        for (uint32_t pos = 0; pos < lx && result; ++pos)
            if (a[pos] != b[pos]) { ++ham; result = ham <= t; }
        return result;
    }
    /* half_core_points, PSDBSCAN.hcu:779-811. */
    int64_t dl = (int64_t)lx - (int64_t)ly;
    if (dl < 0) dl = -dl;
    if (!(gate_vj(d, x, y) && dl <= t)) return 0;
    const uint8_t *a = row_seq(d, x), *b = row_seq(d, y);
    if (lx == ly && (int64_t)orc_hamming(a, b, lx) <= t) return 1;
    return (int64_t)orc_levenshtein(a, lx, b, ly) <= t;
}

int32_t orc_neighbour_fuse_pass(const orc_data *d, const orc_params *p, uint32_t x, uint32_t y)
{
    uint32_t lx = d->len[x], ly = d->len[y];
    if (p->use_hamming) {
        /* fuse_hamming_clusters, PSDBSCAN.hcu:689-719: the running count is
         * compared with t after every position. */
        int32_t t = orc_threshold(lx, ly, p->eps, p->abs_k);
        if (lx != ly || !gate_vj(d, x, y)) return 0;
        const uint8_t *a = row_seq(d, x), *b = row_seq(d, y);
        int64_t ham = 0;
        int result = 1;
        for (uint32_t pos = 0; pos < lx && result; ++pos) {
            if (a[pos] != b[pos]) ++ham;
            result = ham <= t;
        }
        return result;
    }
    /* fuse_clusters, PSDBSCAN.hcu:870-906: same test as the count pass. */
    return orc_neighbour_count_pass(d, p, x, y);
}

/* ------------------------------------------------------------------------ */
/* Union-find linking the larger root under the smaller one, so that the root */
/* of a component is its minimum index (the invariant roots[v] <= v that the  */
/* reference's exchange-based union keeps, PSDBSCAN.hcu:76-107).              */
static uint32_t uf_find(uint32_t *par, uint32_t x)
{
    while (par[x] != x) {
        par[x] = par[par[x]];
        x = par[x];
    }
    return x;
}

static void uf_union(uint32_t *par, uint32_t a, uint32_t b)
{
    a = uf_find(par, a);
    b = uf_find(par, b);
    if (a == b) return;
    if (a < b) par[b] = a; else par[a] = b;
}

typedef struct { uint32_t a, b; } orc_edge;
typedef struct { orc_edge *e; size_t n, cap; } orc_edges;

static void edges_push(orc_edges *v, uint32_t a, uint32_t b)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 1024;
        v->e = (orc_edge *)realloc(v->e, sizeof(orc_edge) * v->cap);
    }
    v->e[v->n].a = a;
    v->e[v->n].b = b;
    v->n++;
}

static int nthreads_of(const orc_params *p)
{
#ifdef _OPENMP
    return p->nthreads > 0 ? p->nthreads : omp_get_max_threads();
#else
    (void)p;
    return 1;
#endif
}

static void atomic_min_u32(uint32_t *addr, uint32_t v)
{
    uint32_t cur = __atomic_load_n(addr, __ATOMIC_RELAXED);
    while (v < cur && !__atomic_compare_exchange_n(addr, &cur, v, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
}

/* R12 (re_number_clusters, PSDBSCAN.hcu:1263-1294): R = distinct roots of   */
/* core rows; label = 1 + rank in ascending order, 0 when not in R.           */
static void renumber(uint32_t n, const uint32_t *root, const int32_t *core, orc_result *r)
{
    uint32_t *rank = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i)
        if (core[i]) rank[root[i]] = 1;
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (rank[i]) rank[i] = ++k;
    uint32_t noise = 0;
    for (uint32_t i = 0; i < n; ++i) {
        r->clone_id[i] = rank[root[i]];
        noise += r->clone_id[i] == 0;
    }
    r->n_clusters = k;
    r->n_noise = noise;
    free(rank);
}

static void emit_common(uint32_t n, const uint32_t *copies, const int32_t *cnt, const orc_params *p,
                        int32_t *core, orc_result *r)
{
    uint32_t nu = 0, nc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        /* core flag: is_core_point_vector >= minPts with int operands (:1180) */
        core[i] = cnt[i] >= (int32_t)p->min_pts;
        nc += core[i] != 0;
        r->core[i] = core[i];
        r->copies[i] = copies[i];
        if (r->counts) r->counts[i] = (uint32_t)cnt[i];
        if (copies[i]) r->uniques[nu++] = i;
    }
    for (uint32_t i = nu; i < n; ++i) r->uniques[i] = 0;
    r->n_uniques = nu;
    r->n_cores = nc;
}

/* ------------------------------------------------------------------------ */
/* Literal restatement of DBSCAN::doScan (PSDBSCAN.hcu:1300-1356).            */
int32_t orc_ref_dbscan(const orc_data *d, const orc_params *p, orc_result *r)
{
    const uint32_t n = d->n;
    r->n_edges = 0;
    if (n == 0) {
        r->n_uniques = r->n_cores = r->n_clusters = r->n_noise = 0;
        return 0;
    }
    const int nt = nthreads_of(p);
    uint32_t *roots = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *copies = (uint32_t *)calloc(n, sizeof(uint32_t));
    int32_t *cnt = (int32_t *)calloc(n, sizeof(int32_t));
    int32_t *core = (int32_t *)calloc(n, sizeof(int32_t));
    uint32_t *bmin = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; ++i) { roots[i] = i; bmin[i] = ORC_NONE; }

    /* identify_duplicates (:1049-1072): every pair q < s. */
    orc_edges *tl = (orc_edges *)calloc((size_t)nt, sizeof(orc_edges));
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t q = 0; q < (int64_t)n; ++q) {
#ifdef _OPENMP
        orc_edges *mine = &tl[omp_get_thread_num()];
#else
        orc_edges *mine = &tl[0];
#endif
        for (uint32_t s = (uint32_t)q + 1; s < n; ++s)
            if (is_duplicate(d, (uint32_t)q, s)) edges_push(mine, s, (uint32_t)q);
    }
    for (int t = 0; t < nt; ++t) {
        for (size_t k = 0; k < tl[t].n; ++k) uf_union(roots, tl[t].e[k].a, tl[t].e[k].b);
        tl[t].n = 0;
    }
    /* reduce_trees_to_roots (:1007-1017) and prepare_copies_vector (:1074-1085). */
    for (uint32_t i = 0; i < n; ++i) roots[i] = uf_find(roots, i);
    for (uint32_t i = 0; i < n; ++i) copies[roots[i]]++;

    /* identify_core_points (:1097-1184): counts start at copies; uniques are
     * the indices with non-zero copies in ascending order. */
    uint32_t *uniq = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t nu = 0;
    for (uint32_t i = 0; i < n; ++i) {
        cnt[i] = (int32_t)copies[i];
        if (copies[i]) uniq[nu++] = i;
    }
    uint64_t edges_total = 0;
#pragma omp parallel for schedule(dynamic, 8) num_threads(nt) reduction(+ : edges_total)
    for (int64_t a = 0; a < (int64_t)nu; ++a) {
        uint32_t x = uniq[a];
        for (uint32_t b = (uint32_t)a + 1; b < nu; ++b) {
            uint32_t y = uniq[b];
            if (orc_neighbour_count_pass(d, p, x, y)) {
                __atomic_fetch_add(&cnt[x], (int32_t)copies[y], __ATOMIC_RELAXED);
                __atomic_fetch_add(&cnt[y], (int32_t)copies[x], __ATOMIC_RELAXED);
                edges_total++;
            }
        }
    }
    r->n_edges = edges_total;
    for (uint32_t i = 0; i < n; ++i) core[i] = cnt[i] >= (int32_t)p->min_pts;

    /* grow_clusters (:1189-1260): x over core rows, y over uniques, pairs with
     * x < y or y not core.  Core y: union.  Non-core y: the first core
     * neighbour in index order claims it. */
    uint32_t *cores = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t nc = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (core[i]) cores[nc++] = i;
#pragma omp parallel for schedule(dynamic, 8) num_threads(nt)
    for (int64_t a = 0; a < (int64_t)nc; ++a) {
#ifdef _OPENMP
        orc_edges *mine = &tl[omp_get_thread_num()];
#else
        orc_edges *mine = &tl[0];
#endif
        uint32_t x = cores[a];
        for (uint32_t b = 0; b < nu; ++b) {
            uint32_t y = uniq[b];
            if (!(x < y || !core[y])) continue;
            if (!orc_neighbour_fuse_pass(d, p, x, y)) continue;
            if (core[y]) edges_push(mine, y, x);
            else atomic_min_u32(&bmin[y], x);
        }
    }
    for (int t = 0; t < nt; ++t) {
        for (size_t k = 0; k < tl[t].n; ++k) uf_union(roots, tl[t].e[k].a, tl[t].e[k].b);
        free(tl[t].e);
    }
    free(tl);
    for (uint32_t b = 0; b < nu; ++b) {
        uint32_t y = uniq[b];
        if (!core[y] && bmin[y] != ORC_NONE && roots[y] == y) roots[y] = bmin[y];
    }
    /* reduce_trees_to_roots (:1344-1347), then re_number_clusters. */
    for (uint32_t i = 0; i < n; ++i) roots[i] = uf_find(roots, i);
    emit_common(n, copies, cnt, p, core, r);
    renumber(n, roots, core, r);

    free(roots); free(copies); free(cnt); free(core); free(bmin); free(uniq); free(cores);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Bucketed restatement with identical results.                               */
static const orc_data *g_sort_data;

static int cmp_rows(const void *pa, const void *pb)
{
    const orc_data *d = g_sort_data;
    uint32_t a = *(const uint32_t *)pa, b = *(const uint32_t *)pb;
    if (d->len[a] != d->len[b]) return d->len[a] < d->len[b] ? -1 : 1;
    int c = memcmp(row_seq(d, a), row_seq(d, b), d->len[a]);
    if (c) return c;
    c = memcmp(d->vbits + (size_t)a * d->vw, d->vbits + (size_t)b * d->vw, sizeof(int32_t) * d->vw);
    if (c) return c;
    c = memcmp(d->jbits + (size_t)a * d->jw, d->jbits + (size_t)b * d->jw, sizeof(int32_t) * d->jw);
    if (c) return c;
    return a < b ? -1 : (a > b);
}

typedef struct { uint64_t key; uint32_t row; } orc_entry;

static int cmp_entry(const void *pa, const void *pb)
{
    const orc_entry *a = (const orc_entry *)pa, *b = (const orc_entry *)pb;
    if (a->key != b->key) return a->key < b->key ? -1 : 1;
    return a->row < b->row ? -1 : (a->row > b->row);
}

static int lowest_common_bit(const int32_t *a, const int32_t *b, uint32_t w)
{
    for (uint32_t k = 0; k < w; ++k) {
        uint32_t m = (uint32_t)(a[k] & b[k]);
        if (m) return (int)(k * 32 + __builtin_ctz(m));
    }
    return -1;
}

/* The count-pass predicate decided with a banded DP; the fuse pass differs only */
/* in Hamming mode with a negative threshold.                                 */
static int fast_pred(const orc_data *d, const orc_params *p, uint32_t x, uint32_t y, int *fuse_ok)
{
    uint32_t lx = d->len[x], ly = d->len[y];
    int32_t t = orc_threshold(lx, ly, p->eps, p->abs_k);
    const uint8_t *a = row_seq(d, x), *b = row_seq(d, y);
    if (p->use_hamming) {
        if (lx != ly) return 0;
        int32_t te = t < 0 ? 0 : t;
        uint32_t h = 0;
        for (uint32_t k = 0; k < lx && (int64_t)h <= te; ++k) h += a[k] != b[k];
        int ok = (int64_t)h <= te;
        *fuse_ok = ok && (t >= 0 || lx == 0);
        return ok;
    }
    *fuse_ok = 1;
    return orc_lev_within(a, lx, b, ly, t);
}

int32_t orc_fast_dbscan(const orc_data *d, const orc_params *p, orc_result *r)
{
    const uint32_t n = d->n;
    r->n_edges = 0;
    if (n == 0) {
        r->n_uniques = r->n_cores = r->n_clusters = r->n_noise = 0;
        return 0;
    }
    const int nt = nthreads_of(p);
    uint32_t *rep = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *copies = (uint32_t *)calloc(n, sizeof(uint32_t));
    int32_t *cnt = (int32_t *)calloc(n, sizeof(int32_t));
    int32_t *core = (int32_t *)calloc(n, sizeof(int32_t));
    for (uint32_t i = 0; i < n; ++i) order[i] = i;

    /* R3/R4 by sorting rows on (length, bytes, V words, J words, index). */
    g_sort_data = d;
    qsort(order, n, sizeof(uint32_t), cmp_rows);
    for (uint32_t k = 0; k < n;) {
        uint32_t e = k + 1;
        while (e < n && is_duplicate(d, order[k], order[e])) ++e;
        for (uint32_t m = k; m < e; ++m) rep[order[m]] = order[k];
        copies[order[k]] = e - k;
        k = e;
    }
    for (uint32_t i = 0; i < n; ++i) cnt[i] = (int32_t)copies[i];

    /* Canonical buckets: a unique row joins bucket (v, j) for every V gene v
     * and J gene j it carries; a pair is evaluated only in the bucket of its
     * lowest common V gene and lowest common J gene. */
    const uint64_t gj = 32ull * d->jw;
    size_t ne = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (!copies[i]) continue;
        uint32_t pv = 0, pj = 0;
        for (uint32_t w = 0; w < d->vw; ++w) pv += __builtin_popcount((uint32_t)d->vbits[(size_t)i * d->vw + w]);
        for (uint32_t w = 0; w < d->jw; ++w) pj += __builtin_popcount((uint32_t)d->jbits[(size_t)i * d->jw + w]);
        ne += (size_t)pv * pj;
    }
    orc_entry *ent = (orc_entry *)malloc(sizeof(orc_entry) * (ne ? ne : 1));
    size_t k = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (!copies[i]) continue;
        for (uint32_t v = 0; v < 32 * d->vw; ++v) {
            if (!((uint32_t)d->vbits[(size_t)i * d->vw + (v >> 5)] >> (v & 31) & 1u)) continue;
            for (uint32_t j = 0; j < 32 * d->jw; ++j) {
                if (!((uint32_t)d->jbits[(size_t)i * d->jw + (j >> 5)] >> (j & 31) & 1u)) continue;
                ent[k].key = (((uint64_t)v * gj + j) << 32) | d->len[i];
                ent[k].row = i;
                ++k;
            }
        }
    }
    qsort(ent, ne, sizeof(orc_entry), cmp_entry);

    /* Length band: in a bucket sorted by length, the targets of entry a are
     * the following entries whose length satisfies L - t(L) <= len(a) (Lev)
     * or L == len(a) (Hamming).  The scan stops at the first longer entry
     * that fails only when L - t(L) is non-decreasing, which holds for eps in
     * [0, 1] and for the absolute mode; otherwise it scans the bucket. */
    uint32_t maxlen = 0;
    for (uint32_t i = 0; i < n; ++i) if (d->len[i] > maxlen) maxlen = d->len[i];
    int monotone = 1;
    for (uint32_t L = 1; L <= maxlen && monotone; ++L) {
        int64_t a = (int64_t)L - orc_threshold(L, L, p->eps, p->abs_k);
        int64_t b = (int64_t)(L - 1) - orc_threshold(L - 1, L - 1, p->eps, p->abs_k);
        if (a < b) monotone = 0;
    }
    orc_edges *tl = (orc_edges *)calloc((size_t)nt, sizeof(orc_edges));
    uint64_t edges_total = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt) reduction(+ : edges_total)
    for (int64_t a = 0; a < (int64_t)ne; ++a) {
#ifdef _OPENMP
        orc_edges *mine = &tl[omp_get_thread_num()];
#else
        orc_edges *mine = &tl[0];
#endif
        const uint64_t bkt = ent[a].key >> 32;
        const uint32_t x = ent[a].row, lx = d->len[x];
        const int vb = (int)(bkt / gj), jb = (int)(bkt % gj);
        for (size_t b = (size_t)a + 1; b < ne && (ent[b].key >> 32) == bkt; ++b) {
            const uint32_t y = ent[b].row, ly = d->len[y];
            if (p->use_hamming && ly != lx) break;
            if (!p->use_hamming) {
                int64_t tl_ = orc_threshold(ly, ly, p->eps, p->abs_k);
                if ((int64_t)ly - tl_ > (int64_t)lx) {
                    if (monotone) break;
                    continue;
                }
            }
            if (lowest_common_bit(d->vbits + (size_t)x * d->vw, d->vbits + (size_t)y * d->vw, d->vw) != vb) continue;
            if (lowest_common_bit(d->jbits + (size_t)x * d->jw, d->jbits + (size_t)y * d->jw, d->jw) != jb) continue;
            int fuse_ok = 0;
            if (!fast_pred(d, p, x, y, &fuse_ok)) continue;
            __atomic_fetch_add(&cnt[x], (int32_t)copies[y], __ATOMIC_RELAXED);
            __atomic_fetch_add(&cnt[y], (int32_t)copies[x], __ATOMIC_RELAXED);
            edges_total++;
            if (fuse_ok) edges_push(mine, x < y ? x : y, x < y ? y : x);
        }
    }
    r->n_edges = edges_total;
    for (uint32_t i = 0; i < n; ++i) core[i] = cnt[i] >= (int32_t)p->min_pts;

    /* Core components (union to min) and borders (min-index core neighbour). */
    uint32_t *par = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *bmin = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; ++i) { par[i] = i; bmin[i] = ORC_NONE; }
    for (int t = 0; t < nt; ++t) {
        for (size_t m = 0; m < tl[t].n; ++m) {
            uint32_t x = tl[t].e[m].a, y = tl[t].e[m].b;
            if (core[x] && core[y]) uf_union(par, x, y);
            else if (core[x]) { if (x < bmin[y]) bmin[y] = x; }
            else if (core[y]) { if (y < bmin[x]) bmin[x] = y; }
        }
        free(tl[t].e);
    }
    free(tl);
    uint32_t *root = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t u = rep[i];
        if (core[u]) root[i] = uf_find(par, u);
        else if (bmin[u] != ORC_NONE) root[i] = uf_find(par, bmin[u]);
        else root[i] = u;
    }
    emit_common(n, copies, cnt, p, core, r);
    /* With minPts == 0 duplicate rows are core rows too; their root is their
     * representative's root, so R is unchanged. */
    renumber(n, root, core, r);
    free(rep); free(order); free(copies); free(cnt); free(core); free(ent); free(par); free(bmin); free(root);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Per-row facts for a sample of a large input, without clustering it: R3/R4  */
/* (representative = smallest equal row, class size) and R8 (the exact        */
/* copies-weighted count, and the neighbour rows) for each sampled row.  All  */
/* rows are indexed once by (V gene, J gene, length); a sample then scans the */
/* length band of its own buckets, and a row z is taken only in the bucket of */
/* the lowest common V gene and J gene of x and z, so each pair counts once.  */
/* count(x) = |class(x)| + #{rows z outside class(x) with N(x, z)}, which is   */
/* copies[x] + sum over unique neighbours y of copies[y] for a representative */
/* x (every row of y's class has the same predicate as y).  Neighbours are    */
/* decided by the count pass's predicate with the banded DP (fast_pred; the   */
/* bucket and the length band already hold the gate and the length test).    */
static void radix_sort_entries(orc_entry *a, orc_entry *tmp, size_t n, unsigned bits)
{
    /* stable LSD passes of 16 bits over the key (rows stay in input order
     * within equal keys) */
    size_t *cnt = (size_t *)malloc(sizeof(size_t) * 65536);
    for (unsigned sh = 0; sh < bits; sh += 16) {
        memset(cnt, 0, sizeof(size_t) * 65536);
        for (size_t i = 0; i < n; ++i) cnt[(a[i].key >> sh) & 0xFFFFu]++;
        size_t s = 0;
        for (size_t b = 0; b < 65536; ++b) { size_t c = cnt[b]; cnt[b] = s; s += c; }
        for (size_t i = 0; i < n; ++i) tmp[cnt[(a[i].key >> sh) & 0xFFFFu]++] = a[i];
        memcpy(a, tmp, sizeof(orc_entry) * n);
    }
    free(cnt);
}

static size_t lower_key(const orc_entry *e, size_t n, uint64_t key)
{
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (e[mid].key < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

typedef struct { uint32_t *v; size_t n, cap; } orc_u32vec;

static void u32vec_push(orc_u32vec *v, uint32_t x)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 64;
        v->v = (uint32_t *)realloc(v->v, sizeof(uint32_t) * v->cap);
    }
    v->v[v->n++] = x;
}

int32_t orc_sample_rows(const orc_data *d, const orc_params *p, const uint32_t *xs, uint32_t n_x, uint32_t *rep,
                        uint32_t *copies, uint32_t *count, uint64_t *nbr_off, uint32_t *nbr, uint64_t nbr_cap)
{
    const uint32_t n = d->n;
    for (uint32_t k = 0; k < n_x; ++k)
        if (xs[k] >= n) return -1;
    const uint64_t gj = 32ull * d->jw;
    uint32_t maxlen = 0;
    for (uint32_t i = 0; i < n; ++i) if (d->len[i] > maxlen) maxlen = d->len[i];
    size_t ne = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t pv = 0, pj = 0;
        for (uint32_t w = 0; w < d->vw; ++w) pv += __builtin_popcount((uint32_t)d->vbits[(size_t)i * d->vw + w]);
        for (uint32_t w = 0; w < d->jw; ++w) pj += __builtin_popcount((uint32_t)d->jbits[(size_t)i * d->jw + w]);
        ne += (size_t)pv * pj;
    }
    orc_entry *ent = (orc_entry *)malloc(sizeof(orc_entry) * (ne ? ne : 1));
    orc_entry *tmp = (orc_entry *)malloc(sizeof(orc_entry) * (ne ? ne : 1));
    size_t k = 0;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t v = 0; v < 32 * d->vw; ++v) {
            if (!((uint32_t)d->vbits[(size_t)i * d->vw + (v >> 5)] >> (v & 31) & 1u)) continue;
            for (uint32_t j = 0; j < 32 * d->jw; ++j) {
                if (!((uint32_t)d->jbits[(size_t)i * d->jw + (j >> 5)] >> (j & 31) & 1u)) continue;
                ent[k].key = (((uint64_t)v * gj + j) << 16) | d->len[i];
                ent[k].row = i;
                ++k;
            }
        }
    unsigned bits = 16;
    while (bits < 64 && ((32ull * d->vw * gj) >> (bits - 16))) bits += 16;
    radix_sort_entries(ent, tmp, ne, bits);
    free(tmp);

    orc_u32vec *nb = (orc_u32vec *)calloc(n_x ? n_x : 1, sizeof(orc_u32vec));
    const int nt = nthreads_of(p);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t s = 0; s < (int64_t)n_x; ++s) {
        const uint32_t x = xs[s], lx = d->len[x];
        const int32_t *vx = d->vbits + (size_t)x * d->vw, *jx = d->jbits + (size_t)x * d->jw;
        const int vlow = lowest_common_bit(vx, vx, d->vw), jlow = lowest_common_bit(jx, jx, d->jw);
        uint32_t r = x, c = 0;
        /* R3 class of x: equal rows share every bucket of x (and rows with no
         * V or no J gene have no bucket: a linear scan finds theirs) */
        if (vlow < 0 || jlow < 0) {
            for (uint32_t z = 0; z < n; ++z)
                if (is_duplicate(d, x, z)) { ++c; if (z < r) r = z; }
        } else {
            const uint64_t key = (((uint64_t)vlow * gj + (uint64_t)jlow) << 16) | lx;
            for (size_t e = lower_key(ent, ne, key); e < ne && ent[e].key == key; ++e)
                if (is_duplicate(d, x, ent[e].row)) { ++c; if (ent[e].row < r) r = ent[e].row; }
        }
        rep[s] = r;
        copies[s] = c;
        uint32_t cntx = c;
        if (vlow >= 0 && jlow >= 0) {
            for (uint32_t v = (uint32_t)vlow; v < 32 * d->vw; ++v) {
                if (!((uint32_t)vx[v >> 5] >> (v & 31) & 1u)) continue;
                for (uint32_t j = (uint32_t)jlow; j < 32 * d->jw; ++j) {
                    if (!((uint32_t)jx[j >> 5] >> (j & 31) & 1u)) continue;
                    const uint64_t bkt = (uint64_t)v * gj + j;
                    for (uint32_t L = 0; L <= maxlen; ++L) {
                        /* the lengths a neighbour can have: equal (Hamming)
                         * or |lx - L| <= t(lx, L) (Levenshtein) */
                        if (p->use_hamming) {
                            if (L != lx) continue;
                        } else {
                            int64_t dl = (int64_t)lx - (int64_t)L;
                            if (dl < 0) dl = -dl;
                            if (dl > orc_threshold(lx, L, p->eps, p->abs_k)) continue;
                        }
                        const uint64_t key = (bkt << 16) | L;
                        for (size_t e = lower_key(ent, ne, key); e < ne && ent[e].key == key; ++e) {
                            const uint32_t z = ent[e].row;
                            if (z == x || is_duplicate(d, x, z)) continue;
                            const int32_t *vz = d->vbits + (size_t)z * d->vw, *jz = d->jbits + (size_t)z * d->jw;
                            if (lowest_common_bit(vx, vz, d->vw) != (int)v || lowest_common_bit(jx, jz, d->jw) != (int)j)
                                continue;
                            int fuse_ok = 0;  /* the count pass's predicate, banded */
                            if (!fast_pred(d, p, x, z, &fuse_ok)) continue;
                            ++cntx;
                            u32vec_push(&nb[s], z);
                        }
                    }
                }
            }
        }
        count[s] = cntx;
    }
    uint64_t o = 0;
    for (uint32_t s = 0; s < n_x; ++s) {
        nbr_off[s] = o;
        for (size_t m = 0; m < nb[s].n; ++m, ++o)
            if (o < nbr_cap) nbr[o] = nb[s].v[m];
        free(nb[s].v);
    }
    nbr_off[n_x] = o;
    free(nb);
    free(ent);
    return 0;
}
