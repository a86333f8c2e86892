// Interlink.hpp — drop-in replacement for the reference's host->device
// boundary class (cuda/Interlink.hpp:41-69).  The declarations below keep the
// reference's typedefs, constructor signature and getters, so its caller
// (cpp/TableHolder.hpp:337-347) compiles unchanged against rocThrust's
// thrust::host_vector; the implementation (Interlink.cpp) runs the clustering
// in libidb_hip.so through the C ABI of include/idb.h.
#pragma once
#include <memory>
#include <string>
#include <thrust/host_vector.h>
#include <thrust/iterator/permutation_iterator.h>
#include <thrust/iterator/zip_iterator.h>
#include <thrust/tuple.h>

typedef thrust::host_vector<std::string>::iterator NormalStringVectorIteratorType;
typedef thrust::tuple<NormalStringVectorIteratorType, NormalStringVectorIteratorType, NormalStringVectorIteratorType>
    NormalStringVectorIteratorTupleType;
typedef thrust::zip_iterator<NormalStringVectorIteratorTupleType> NormalStringVectorZipIteratorType;

class InterlinkClass
{
    thrust::host_vector<unsigned int> clusters, uniques, copies;
    thrust::host_vector<int> cores, vbitset;
    unsigned int n_uniques;
    unsigned int v_bitset_width;

  public:
    // InputStrings zips the V-call, J-call and sequence columns; _threshold is
    // maximum_relative_edit_distance; the call runs synchronously and throws
    // std::runtime_error on failure.
    InterlinkClass(NormalStringVectorZipIteratorType InputStrings, unsigned int _n_entries,
                   NormalStringVectorIteratorType V_genes_reference, unsigned int V_gene_variant_count,
                   NormalStringVectorIteratorType J_genes_reference, unsigned int J_gene_variant_count,
                   float _threshold, unsigned int _minPts, bool use_only_hamming, bool eval_trees);

    thrust::host_vector<unsigned int>::iterator get_clusters_iterator();
    thrust::host_vector<int>::iterator get_cores_iterator();
    thrust::host_vector<unsigned int>::iterator get_uniques_iterator();
    thrust::host_vector<unsigned int>::iterator get_copies_iterator();
    unsigned int get_n_uniques();

    thrust::host_vector<int>::iterator v_gene_container_begin();
    thrust::host_vector<int>::iterator v_gene_container_end();
    unsigned int get_v_bitset_width() { return v_bitset_width; }
};
