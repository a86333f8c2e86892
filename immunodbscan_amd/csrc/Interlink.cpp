// Interlink.cpp — InterlinkClass over libidb_hip.so.
//
// Same steps as the reference constructor (cuda/Interlink.cu:34-95): V and J
// gene bitsets from substring search of the reference gene names (R2), the
// sequence column packed as concatenated bytes with offsets and lengths, one
// clustering run, and host copies of clone ids, cores, uniques and copies.
// Device 0 is used, as in the reference (cuda/PSDBSCAN.hcu:976-977); one
// context per host thread is created on first use and kept.
#include "Interlink.hpp"

#include <stdexcept>
#include <vector>

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

std::vector<int32_t> bitsets(NormalStringVectorIteratorType calls, unsigned int n, NormalStringVectorIteratorType genes,
                             unsigned int n_genes, unsigned int &width)
{
    width = (n_genes + 31) / 32;
    const Packed c = pack(calls, n), g = pack(genes, n_genes);
    std::vector<int32_t> out((size_t)n * (width ? width : 1), 0);
    if (idb_gene_bitsets(n, c.bytes.data(), c.off.data(), c.len.data(), n_genes, g.bytes.data(), g.off.data(),
                         g.len.data(), out.data(), width ? width : 1) != IDB_OK)
        throw std::runtime_error("idb_gene_bitsets failed");
    out.resize((size_t)n * width);
    return out;
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
    unsigned int j_width = 0;
    std::vector<int32_t> v = bitsets(thrust::get<0>(cols), _n_entries, V_genes_reference, V_gene_variant_count,
                                     v_bitset_width);
    std::vector<int32_t> j = bitsets(thrust::get<1>(cols), _n_entries, J_genes_reference, J_gene_variant_count,
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
    idb_ctx *ctx = thread_context();
    if (idb_dbscan(ctx, &d, &p, &o, &st) != IDB_OK) throw std::runtime_error(idb_last_error(ctx));
    n_uniques = st.n_uniques;
}

thrust::host_vector<unsigned int>::iterator InterlinkClass::get_clusters_iterator() { return clusters.begin(); }
thrust::host_vector<int>::iterator InterlinkClass::get_cores_iterator() { return cores.begin(); }
thrust::host_vector<unsigned int>::iterator InterlinkClass::get_uniques_iterator() { return uniques.begin(); }
thrust::host_vector<unsigned int>::iterator InterlinkClass::get_copies_iterator() { return copies.begin(); }
thrust::host_vector<int>::iterator InterlinkClass::v_gene_container_begin() { return vbitset.begin(); }
thrust::host_vector<int>::iterator InterlinkClass::v_gene_container_end() { return vbitset.end(); }
unsigned int InterlinkClass::get_n_uniques() { return n_uniques; }
