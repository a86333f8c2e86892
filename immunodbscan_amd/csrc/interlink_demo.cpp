// interlink_demo.cpp — drives InterlinkClass the way the reference's
// TableHolderClass::do_the_clustering does (cpp/TableHolder.hpp:275-348):
// three string columns and two gene lists in, clone_id / copies / cores out.
//
// usage: interlink_demo <rows.tsv> <V genes file> <J genes file> <eps> <minPts> <hamming 0|1>
//                       [<run log .tsv> [<tree stats 0|1>]]
//   rows.tsv   one row per line: V call <TAB> J call <TAB> sequence
//   gene files one gene name per line, empty lines skipped (cpp/GeneReader.hpp:65-76)
// Output: one line per row, "clone_id<TAB>copies<TAB>cores", then
// "# n_uniques <k>" and "# v_bitset_width <w>"; with a run-log path, the
// run's messages and stage timers in the reference's TSV format.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "Interlink.hpp"
#include "RunLog.hpp"

static thrust::host_vector<std::string> read_genes(const std::string &path)
{
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("Cannot read from file " + path);
    thrust::host_vector<std::string> g;
    std::string line;
    while (std::getline(f, line))
        if (!line.empty()) g.push_back(line);
    return g;
}

int main(int argc, char **argv)
{
    if (argc < 7 || argc > 9) {
        std::fprintf(stderr, "usage: %s rows.tsv Vgenes Jgenes eps minPts hamming [runlog.tsv [tree_stats]]\n", argv[0]);
        return 2;
    }
    try {
        thrust::host_vector<std::string> v, j, s;
        std::ifstream in(argv[1]);
        if (!in.is_open()) throw std::runtime_error(std::string("Cannot read from file ") + argv[1]);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string a, b, c;
            std::getline(ls, a, '\t');
            std::getline(ls, b, '\t');
            std::getline(ls, c, '\t');
            v.push_back(a);
            j.push_back(b);
            s.push_back(c);
        }
        thrust::host_vector<std::string> vg = read_genes(argv[2]), jg = read_genes(argv[3]);
        const float eps = std::strtof(argv[4], nullptr);
        const unsigned minpts = (unsigned)std::strtoul(argv[5], nullptr, 10);
        const bool ham = std::atoi(argv[6]) != 0;
        NormalStringVectorZipIteratorType zip =
            thrust::make_zip_iterator(thrust::make_tuple(v.begin(), j.begin(), s.begin()));
        const bool trees = argc > 8 && std::atoi(argv[8]) != 0;
        InterlinkClass il(zip, (unsigned)s.size(), vg.begin(), (unsigned)vg.size(), jg.begin(), (unsigned)jg.size(), eps,
                          minpts, ham, trees);
        auto cl = il.get_clusters_iterator();
        auto cp = il.get_copies_iterator();
        auto co = il.get_cores_iterator();
        for (size_t i = 0; i < s.size(); ++i) std::printf("%u\t%u\t%d\n", cl[i], cp[i], co[i]);
        std::printf("# n_uniques %u\n# v_bitset_width %u\n", il.get_n_uniques(), il.get_v_bitset_width());
        if (argc > 7 && !idb::RunLog::instance().write(argv[7]))
            throw std::runtime_error(std::string("Cannot write to file ") + argv[7]);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
