// synth.h — seeded synthetic repertoire generator (host C ABI).
#ifndef IDB_SYNTH_H
#define IDB_SYNTH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { IDB_SYNTH_NT = 0, IDB_SYNTH_AA = 1 };
enum { IDB_SYNTH_LEN_UNIFORM = 0, IDB_SYNTH_LEN_NORMAL = 1 };
enum { IDB_SYNTH_SIZE_GEOMETRIC = 0, IDB_SYNTH_SIZE_LOGNORMAL = 1 };
enum { IDB_SYNTH_MUT_SUB = 0, IDB_SYNTH_MUT_EDITS = 1, IDB_SYNTH_MUT_SUB_INDEL = 2 };

typedef struct {
    uint64_t seed;
    uint32_t n_rows;      /* exact number of rows produced                       */
    uint32_t n_seeds;     /* seed clonotypes drawn first                          */
    uint32_t alphabet;    /* IDB_SYNTH_NT ("ACGT") or IDB_SYNTH_AA (20 letters)    */
    uint32_t len_mode;    /* uniform over lo..hi step, or rounded normal clipped   */
    uint32_t len_lo, len_hi, len_step;
    double len_mean, len_sd;
    uint32_t size_mode;   /* geometric(size_p) on {1,2,..} or lognormal(mu,sigma)  */
    double size_p, size_mu, size_sigma;
    uint32_t mut_mode;    /* per-letter substitution; k edits; sub + indel rates   */
    double sub_rate, indel_rate;
    uint32_t k_lo, k_hi;
    uint32_t n_v, n_j;    /* gene counts                                          */
    double v_zipf_s, j_zipf_s; /* 0 = uniform gene usage                           */
    double v_ambig, j_ambig;   /* probability of a second gene in a call          */
    uint32_t shuffle;
} idb_synth_params;

void idb_synth_default(idb_synth_params *p);
/* Returns the total sequence bytes (the needed capacity when seq is NULL or
 * seq_cap is too small; nothing else is written then), or -1 on bad params. */
int64_t idb_synth_generate(const idb_synth_params *p, uint8_t *seq, uint64_t seq_cap, uint64_t *off, uint32_t *len,
                           int32_t *vbits, uint32_t vw, int32_t *jbits, uint32_t jw, uint32_t *clone_of);
int32_t idb_synth_lognormal_sizes(uint64_t seed, uint32_t n, double median, double sigma, uint32_t lo, uint32_t hi,
                                  uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif
