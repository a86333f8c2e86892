// Interlink.cpp — InterlinkClass over libidb_hip.so.
//
// Same steps as the reference constructor (cuda/Interlink.cu:34-95): V and J
// gene bitsets from substring search of the reference gene names on the
// device (R2, idb_gene_bitsets_gpu in place of bitset_container's per-gene
// launches, cuda/BitsetVector.inl:181-209), the sequence column packed as
// concatenated bytes with offsets and lengths, one clustering run, and host
// copies of clone ids, cores, uniques and copies.  The run's messages and
// stage timers go to the process's run log (RunLog.hpp) as the reference's
// DBSCAN writes them to its BenchmarkSingleton.  Device 0 is used, as in the
// reference (cuda/PSDBSCAN.hcu:976-977); one context per host thread is
// created on first use and kept.
#include "Interlink.hpp"

#include <stdexcept>
#include <vector>

#include "RunLog.hpp"
#include "idb.h"

namespace {

struct CtxHolder {
    idb_ctx *ctx = nullptr;
    ~CtxHolder()
    {
        if (ctx) idb_destroy(ctx);
    }
};

idb_ctx *thread_context()
{
    thread_local CtxHolder h;
    if (!h.ctx && idb_create(&h.ctx, 0) != IDB_OK) {
        h.ctx = nullptr;
        throw std::runtime_error("idb_create: no usable HIP device 0");
    }
    return h.ctx;
}

struct Packed {
    std::vector<uint8_t> bytes;
    std::vector<uint64_t> off;
    std::vector<uint32_t> len;
};

template <typename It>
Packed pack(It first, unsigned int n)
{
    Packed p;
    p.off.resize(n);
    p.len.resize(n);
    uint64_t pos = 0;
    It it = first;
    for (unsigned int i = 0; i < n; ++i, ++it) {
        const std::string &s = *it;
        p.off[i] = pos;
        p.len[i] = (uint32_t)s.size();
        p.bytes.insert(p.bytes.end(), s.begin(), s.end());
        pos += s.size();
    }
    if (p.bytes.empty()) p.bytes.push_back(0);
    return p;
}

std::vector<int32_t> bitsets(idb_ctx *ctx, NormalStringVectorIteratorType calls, unsigned int n,
                             NormalStringVectorIteratorType genes, unsigned int n_genes, unsigned int &width)
{
    width = (n_genes + 31) / 32;
    const Packed c = pack(calls, n), g = pack(genes, n_genes);
    std::vector<int32_t> out((size_t)n * (width ? width : 1), 0);
    if (width > 0 && idb_gene_bitsets_gpu(ctx, IDB_MEM_HOST, n, c.bytes.data(), c.off.data(), c.len.data(), n_genes,
                                          g.bytes.data(), g.off.data(), g.len.data(), out.data(), width) != IDB_OK)
        throw std::runtime_error(std::string("idb_gene_bitsets_gpu: ") + idb_last_error(ctx));
    out.resize((size_t)n * width);
    return out;
}

// The messages and stage timers of one run, in the order of the reference's
// DBSCAN constructor and doScan (cuda/PSDBSCAN.hcu:986-997, 1300-1354).
void log_run(const idb_params &p, const idb_stats &st, unsigned int n)
{
    static const char *const stage[IDB_N_STAGES] = {
        "Identify Duplicates", "Reduction of duplicate information tree structure", "Determine copies",
        "Identify core points", "Grow clusters", "Reduction of cluster information tree structure",
        "Re-number clusters and identify noise", "Clustering core process"};
    idb::RunLog &log = idb::RunLog::instance();
    log.message("DBSCAN settings: epsilon = " + std::to_string(p.eps));
    log.message("DBSCAN settings: minPts = " + std::to_string(p.min_pts));
    if (p.use_hamming) log.message("DBSCAN settings: using Hamming distance measure instead of Levenshtein distance");
    log.message("Total number of sequences = " + std::to_string(n));
    for (int k = 0; k < 3; ++k) log.timer(stage[k], st.stage_ms[k] * 1e-3);
    log.message("Number of uniques = " + std::to_string(st.n_uniques));
    log.timer(stage[3], st.stage_ms[3] * 1e-3);
    log.message("Number of core points = " + std::to_string(st.n_cores));
    log.timer(stage[4], st.stage_ms[4] * 1e-3);
    if (p.tree_stats) {
        log.message("Union find tree depth structure");
        for (unsigned int i = 0; i <= 50; ++i)
            log.message("Number of trees with depth " + std::to_string(i) + ": " + std::to_string(st.tree_depth[i]));
    }
    log.timer(stage[5], st.stage_ms[5] * 1e-3);
    log.message("Total number of clusters = " + std::to_string(st.n_clusters));
    log.message("Number of noise points = " + std::to_string(st.n_noise));
    // unsigned 32-bit arithmetic, as in the reference (:1291)
    const unsigned int perc = n ? (100u * st.n_noise) / n : 0u;
    log.message("Percentage of noise points = " + std::to_string(perc) + "%");
    log.timer(stage[6], st.stage_ms[6] * 1e-3);
    log.timer(stage[7], st.stage_ms[7] * 1e-3);
}

}  // namespace

InterlinkClass::InterlinkClass(NormalStringVectorZipIteratorType InputStrings, unsigned int _n_entries,
                               NormalStringVectorIteratorType V_genes_reference, unsigned int V_gene_variant_count,
                               NormalStringVectorIteratorType J_genes_reference, unsigned int J_gene_variant_count,
                               float _threshold, unsigned int _minPts, bool use_only_hamming, bool eval_trees)
    : clusters(_n_entries, 0), uniques(_n_entries, 0), copies(_n_entries, 0), cores(_n_entries, 0), n_uniques(0),
      v_bitset_width(0)
{
    const auto cols = InputStrings.get_iterator_tuple();
    idb_ctx *ctx = thread_context();
    unsigned int j_width = 0;
    std::vector<int32_t> v = bitsets(ctx, thrust::get<0>(cols), _n_entries, V_genes_reference, V_gene_variant_count,
                                     v_bitset_width);
    std::vector<int32_t> j = bitsets(ctx, thrust::get<1>(cols), _n_entries, J_genes_reference, J_gene_variant_count,
                                     j_width);
    vbitset.assign(v.begin(), v.end());
    const Packed seqs = pack(thrust::get<2>(cols), _n_entries);

    idb_dataset d{};
    d.n_rows = _n_entries;
    d.mem = IDB_MEM_HOST;
    d.seq = seqs.bytes.data();
    d.seq_bytes = seqs.bytes.size();
    d.off = seqs.off.data();
    d.len = seqs.len.data();
    d.vbits = v.empty() ? nullptr : v.data();
    d.v_width = v_bitset_width;
    d.jbits = j.empty() ? nullptr : j.data();
    d.j_width = j_width;

    idb_params p{};
    p.eps = _threshold;
    p.abs_k = -1;
    p.min_pts = _minPts;
    p.use_hamming = use_only_hamming ? 1u : 0u;
    p.tree_stats = eval_trees ? 1u : 0u;

    idb_outputs o{};
    o.mem = IDB_MEM_HOST;
    o.clone_id = clusters.data();
    o.core = cores.data();
    o.copies = copies.data();
    o.uniques = uniques.data();

    idb_stats st{};
    if (idb_dbscan(ctx, &d, &p, &o, &st) != IDB_OK) throw std::runtime_error(idb_last_error(ctx));
    n_uniques = st.n_uniques;
    log_run(p, st, _n_entries);
}

thrust::host_vector<unsigned int>::iterator InterlinkClass::get_clusters_iterator() { return clusters.begin(); }
thrust::host_vector<int>::iterator InterlinkClass::get_cores_iterator() { return cores.begin(); }
thrust::host_vector<unsigned int>::iterator InterlinkClass::get_uniques_iterator() { return uniques.begin(); }
thrust::host_vector<unsigned int>::iterator InterlinkClass::get_copies_iterator() { return copies.begin(); }
thrust::host_vector<int>::iterator InterlinkClass::v_gene_container_begin() { return vbitset.begin(); }
thrust::host_vector<int>::iterator InterlinkClass::v_gene_container_end() { return vbitset.end(); }
unsigned int InterlinkClass::get_n_uniques() { return n_uniques; }
