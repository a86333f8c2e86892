// synth.cpp — seeded synthetic immune repertoires (CDR3 sequences with V/J gene
// bitsets) for tests and benchmarks.
//
// Every sampler is written here on top of splitmix64 so that a given seed gives
// the same bytes on every machine that runs this binary: the output of
// std::*_distribution differs between standard libraries, so it is not used.
// Floating-point steps use only +, *, log, exp, sqrt and cos in a fixed order.
//
// Model: a repertoire is a set of clonotypes.  Each clonotype has a seed CDR3
// (random length and letters), one V gene and one J gene.  Its members are
// mutated copies of the seed; each member's V (and J) call is the clonotype's
// gene, plus with a given probability one other gene (an ambiguous call).
// Rows are shuffled so that members of a clone are not adjacent.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "synth.h"

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in (0, 1]
    double unit() { return (double)((next() >> 11) + 1) * (1.0 / 9007199254740992.0); }
    // uniform integer in [0, n)
    uint32_t below(uint32_t n) { return n ? (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32) : 0; }
    bool bernoulli(double p) { return unit() <= p; }
    double normal()
    {
        double u1 = unit(), u2 = unit();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// Inverse-CDF sampler over {0, .., n-1} with P(k) proportional to (k+1)^-s.
struct Zipf {
    std::vector<double> cdf;
    Zipf(uint32_t n, double s)
    {
        cdf.resize(n);
        double acc = 0.0;
        for (uint32_t k = 0; k < n; ++k) {
            acc += s > 0.0 ? std::exp(-s * std::log((double)(k + 1))) : 1.0;
            cdf[k] = acc;
        }
        for (uint32_t k = 0; k < n; ++k) cdf[k] /= acc;
    }
    uint32_t draw(Rng &r) const
    {
        double u = r.unit();
        uint32_t lo = 0, hi = (uint32_t)cdf.size() - 1;
        while (lo < hi) {
            uint32_t mid = (lo + hi) / 2;
            if (cdf[mid] >= u) hi = mid; else lo = mid + 1;
        }
        return lo;
    }
};

const char kNt[] = "ACGT";
const char kAa[] = "ACDEFGHIKLMNPQRSTVWY";

struct Alpha {
    const char *sym;
    uint32_t k;
};

Alpha alphabet_of(uint32_t a) { return a == IDB_SYNTH_AA ? Alpha{kAa, 20} : Alpha{kNt, 4}; }

uint32_t draw_len(const idb_synth_params &p, Rng &r)
{
    if (p.len_mode == IDB_SYNTH_LEN_NORMAL) {
        double v = std::floor(p.len_mean + p.len_sd * r.normal() + 0.5);
        if (v < p.len_lo) v = p.len_lo;
        if (v > p.len_hi) v = p.len_hi;
        return (uint32_t)v;
    }
    uint32_t step = p.len_step ? p.len_step : 1;
    uint32_t choices = (p.len_hi - p.len_lo) / step + 1;
    return p.len_lo + step * r.below(choices);
}

char other_symbol(const Alpha &a, char c, Rng &r)
{
    uint32_t k = r.below(a.k - 1);
    char s = a.sym[k];
    return s == c ? a.sym[a.k - 1] : s;
}

void mutate(const idb_synth_params &p, const Alpha &a, std::string &s, Rng &r)
{
    if (p.mut_mode == IDB_SYNTH_MUT_EDITS) {
        uint32_t span = p.k_hi >= p.k_lo ? p.k_hi - p.k_lo + 1 : 1;
        uint32_t k = p.k_lo + r.below(span);
        for (uint32_t e = 0; e < k; ++e) {
            uint32_t kind = r.below(3);  // 0 substitution, 1 insertion, 2 deletion
            if (kind == 2 && s.size() <= 1) kind = 1;
            if (kind == 0 && s.empty()) kind = 1;
            if (kind == 0) {
                uint32_t pos = r.below((uint32_t)s.size());
                s[pos] = other_symbol(a, s[pos], r);
            } else if (kind == 1) {
                uint32_t pos = r.below((uint32_t)s.size() + 1);
                s.insert(s.begin() + pos, a.sym[r.below(a.k)]);
            } else {
                uint32_t pos = r.below((uint32_t)s.size());
                s.erase(s.begin() + pos);
            }
        }
        return;
    }
    std::string out;
    out.reserve(s.size() + 8);
    for (size_t i = 0; i < s.size(); ++i) {
        char c = s[i];
        if (p.sub_rate > 0.0 && r.bernoulli(p.sub_rate)) c = other_symbol(a, c, r);
        if (p.mut_mode == IDB_SYNTH_MUT_SUB_INDEL && p.indel_rate > 0.0 && r.bernoulli(p.indel_rate)) {
            if (r.below(2) == 0) {  // insertion before this position
                out.push_back(a.sym[r.below(a.k)]);
                out.push_back(c);
            }  // else deletion: drop c
        } else {
            out.push_back(c);
        }
    }
    if (out.empty()) out.push_back(a.sym[r.below(a.k)]);
    s.swap(out);
}

void set_bit(int32_t *words, uint32_t g) { words[g >> 5] |= (int32_t)(1u << (g & 31)); }

}  // namespace

extern "C" {

void idb_synth_default(idb_synth_params *p)
{
    std::memset(p, 0, sizeof(*p));
    p->seed = 1;
    p->n_rows = 2000;
    p->n_seeds = 200;
    p->alphabet = IDB_SYNTH_NT;
    p->len_mode = IDB_SYNTH_LEN_UNIFORM;
    p->len_lo = 30;
    p->len_hi = 75;
    p->len_step = 3;
    p->size_mode = IDB_SYNTH_SIZE_GEOMETRIC;
    p->size_p = 0.1;
    p->mut_mode = IDB_SYNTH_MUT_SUB;
    p->sub_rate = 0.02;
    p->n_v = 50;
    p->n_j = 6;
    p->v_ambig = 0.2;
    p->shuffle = 1;
}

int64_t idb_synth_generate(const idb_synth_params *p, uint8_t *seq, uint64_t seq_cap, uint64_t *off, uint32_t *len,
                           int32_t *vbits, uint32_t vw, int32_t *jbits, uint32_t jw, uint32_t *clone_of)
{
    if (!p || p->n_v == 0 || p->n_j == 0 || p->n_v > 32u * vw || p->n_j > 32u * jw) return -1;
    if (p->len_lo == 0 || p->len_hi < p->len_lo) return -1;
    const Alpha a = alphabet_of(p->alphabet);
    Rng rng(p->seed);
    const Zipf vz(p->n_v, p->v_zipf_s), jz(p->n_j, p->j_zipf_s);

    // Clone sizes: seeds 0..n_seeds-1 first; more seeds while rows are short;
    // the last clone is truncated to reach exactly n_rows.
    double scale = 1.0;
    if (p->size_mode == IDB_SYNTH_SIZE_LOGNORMAL && p->n_seeds)
        scale = (double)p->n_rows / ((double)p->n_seeds * std::exp(p->size_mu + 0.5 * p->size_sigma * p->size_sigma));
    std::vector<uint32_t> sizes;
    uint64_t total = 0;
    while (total < p->n_rows || sizes.size() < p->n_seeds) {
        uint32_t s;
        if (p->size_mode == IDB_SYNTH_SIZE_LOGNORMAL) {
            double v = std::floor(std::exp(p->size_mu + p->size_sigma * rng.normal()) * scale + 0.5);
            s = v < 1.0 ? 1u : (v > 4.0e9 ? 4000000000u : (uint32_t)v);
        } else {
            double q = std::log(rng.unit()) / std::log(1.0 - p->size_p);
            s = 1u + (uint32_t)std::floor(q);
        }
        if (total >= p->n_rows) s = 0;  // seeds beyond the row budget stay empty
        else if (total + s > p->n_rows) s = (uint32_t)(p->n_rows - total);
        sizes.push_back(s);
        total += s;
    }

    std::vector<uint32_t> perm(p->n_rows);
    for (uint32_t i = 0; i < p->n_rows; ++i) perm[i] = i;
    if (p->shuffle)
        for (uint32_t i = p->n_rows; i > 1; --i) {
            uint32_t j = rng.below(i);
            uint32_t t = perm[i - 1]; perm[i - 1] = perm[j]; perm[j] = t;
        }

    // Rows are produced in clone order into slot perm[k]; offsets follow row
    // order, so sequences are first collected and then laid out.
    std::vector<std::string> rows(p->n_rows);
    if (vbits) std::memset(vbits, 0, sizeof(int32_t) * (size_t)p->n_rows * vw);
    if (jbits) std::memset(jbits, 0, sizeof(int32_t) * (size_t)p->n_rows * jw);
    uint32_t k = 0;
    for (uint32_t c = 0; c < sizes.size(); ++c) {
        uint32_t L = draw_len(*p, rng);
        std::string seed_seq(L, 'A');
        for (uint32_t i = 0; i < L; ++i) seed_seq[i] = a.sym[rng.below(a.k)];
        uint32_t vg = vz.draw(rng), jg = jz.draw(rng);
        for (uint32_t m = 0; m < sizes[c]; ++m, ++k) {
            uint32_t row = perm[k];
            std::string s = seed_seq;
            mutate(*p, a, s, rng);
            rows[row] = s;
            if (clone_of) clone_of[row] = c;
            uint32_t v2 = vg, j2 = jg;
            if (p->v_ambig > 0.0 && rng.bernoulli(p->v_ambig) && p->n_v > 1)
                do v2 = vz.draw(rng); while (v2 == vg);
            if (p->j_ambig > 0.0 && rng.bernoulli(p->j_ambig) && p->n_j > 1)
                do j2 = jz.draw(rng); while (j2 == jg);
            if (vbits) { set_bit(vbits + (size_t)row * vw, vg); set_bit(vbits + (size_t)row * vw, v2); }
            if (jbits) { set_bit(jbits + (size_t)row * jw, jg); set_bit(jbits + (size_t)row * jw, j2); }
        }
    }
    uint64_t bytes = 0;
    for (uint32_t i = 0; i < p->n_rows; ++i) bytes += rows[i].size();
    if (!seq || bytes > seq_cap) return (int64_t)bytes;
    uint64_t pos = 0;
    for (uint32_t i = 0; i < p->n_rows; ++i) {
        if (off) off[i] = pos;
        if (len) len[i] = (uint32_t)rows[i].size();
        std::memcpy(seq + pos, rows[i].data(), rows[i].size());
        pos += rows[i].size();
    }
    return (int64_t)bytes;
}

int32_t idb_synth_lognormal_sizes(uint64_t seed, uint32_t n, double median, double sigma, uint32_t lo, uint32_t hi,
                                  uint32_t *out)
{
    if (!out || hi < lo) return -1;
    Rng rng(seed);
    const double mu = std::log(median);
    for (uint32_t i = 0; i < n; ++i) {
        double v = std::floor(std::exp(mu + sigma * rng.normal()) + 0.5);
        if (v < lo) v = lo;
        if (v > hi) v = hi;
        out[i] = (uint32_t)v;
    }
    return 0;
}

}  // extern "C"
