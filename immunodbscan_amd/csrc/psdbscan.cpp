// psdbscan.cpp — the reference's host program (cpp/psdbscan.cpp:25-106,
// cpp/TableHolder.hpp, cpp/ConfigHolder.hpp) over the drop-in InterlinkClass.
//
//   psdbscan [--input FILE ...] [--config_file FILE] [--verbose] [--help]
//
// Same flow, settings, files and log as the reference:
//   1. read the configuration (INFO format, default data/psdbscan.conf) and
//      every input CSV (default data/ExampleData.csv), keeping the columns
//      select_columns_to_read names, plus the file each row came from;
//   2. delete rows whose sequence, first-gene or second-gene cell is empty,
//      in that order (R1);
//   3. cluster: gene calls against the reference gene lists (a missing gene
//      column counts every row as carrying one gene, "V"; a missing sequence
//      column gives every row the sequence "Gene"), then InterlinkClass with
//      maximum_relative_edit_distance, minimum_number_of_neighbors,
//      use_hamming and get_tree_depth_statistics; clone_id, copies and cores
//      become columns;
//   4. optionally delete noise rows (clone_id 0), then write the columns of
//      select_columns_to_write, under their new names, to
//      <dir of first input>/<output_prefix><name><output_suffix><output_filetype>;
//   5. write the run log to the benchmark file named the same way.
// The reference starts its "Do clustering" and "Writing Output" timers twice
// and never stops them, so neither appears in its log; the same holds here.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "CsvTable.hpp"
#include "InfoConfig.hpp"
#include "Interlink.hpp"
#include "RunLog.hpp"

namespace {

// <dir of name><prefix><name without directory and extension><suffix><type>
// (cpp/ConfigHolder.hpp:89-113, including its handling of names without an
// extension and of dots in directory names).
std::string output_name(const idb::InfoNode &cfg, const std::string &name, const std::string &kind)
{
    const std::string path = name.substr(0, name.find_last_of('/') + 1);
    std::string stem = name.substr(0, name.find_last_of('.'));
    stem = stem.substr(stem.find_last_of('/') + 1, stem.length());
    return path + cfg.get_string(kind + "_prefix") + stem + cfg.get_string(kind + "_suffix") +
           cfg.get_string(kind + "_filetype");
}

// The (key, value) pairs of a node, in file order.
void pairs_of(const idb::InfoNode &n, std::vector<std::string> &keys, std::vector<std::string> &values)
{
    for (const auto &c : n.children) {
        keys.push_back(c.first);
        values.push_back(c.second.data);
    }
}

std::string column_setting(const idb::InfoNode &cfg, const std::string &path)
{
    const idb::InfoNode *n = cfg.find(path);
    return n ? n->data : std::string();
}

thrust::host_vector<std::string> read_genes(const std::string &path)
{
    // one name per line, empty lines skipped (cpp/GeneReader.hpp:63-77)
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("Cannot read from file " + path);
    thrust::host_vector<std::string> g;
    std::string line;
    while (std::getline(f, line))
        if (!line.empty()) g.push_back(line);
    return g;
}

int run(int argc, char **argv)
{
    idb::RunLog &log = idb::RunLog::instance();
    log.start("Overall");
    std::vector<std::string> inputs;
    std::string conf = "data/psdbscan.conf";
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--help") {
            std::printf("Allowed options:\n  --help                produce help message\n"
                        "  --input arg           set input file (csv format, settings may need to be adjusted in "
                        "config file)\n  --config_file arg     set config file (boost property_tree info file "
                        "data format)\n  --verbose             Print some status information during data "
                        "processing.\n");
            return 1;
        } else if (a == "--verbose") {
            log.be_verbose();
        } else if (a == "--config_file") {
            if (++i >= argc) throw std::runtime_error("the required argument for option '--config_file' is missing");
            conf = argv[i];
        } else if (a == "--input") {
            while (i + 1 < argc && std::strncmp(argv[i + 1], "--", 2) != 0) inputs.push_back(argv[++i]);
            if (inputs.empty()) throw std::runtime_error("the required argument for option '--input' is missing");
        } else {
            throw std::runtime_error("unrecognised option '" + a + "'");
        }
    }
    if (inputs.empty()) inputs.push_back("data/ExampleData.csv");
    std::printf("\n\n***************************************** PSDBSCAN "
                "************************************************\n");
    const idb::InfoNode cfg = idb::InfoParser::read_file(conf);
    for (const auto &f : inputs) log.message("Input file(s) " + f);
    log.message("Using configuration file " + conf);

    // 1. read
    std::printf("\nReading files ... \n");
    log.start("Reading Files");
    std::vector<std::string> keys, headers;
    pairs_of(cfg.at("select_columns_to_read"), keys, headers);
    idb::Table table(keys);
    const char sep = cfg.get_char("input_settings.separator");
    const char endline = cfg.get_char("input_settings.endline");
    const char quote = cfg.get_char("input_settings.quotation");
    for (const auto &f : inputs) table.read_file(f, headers, sep, endline, quote);
    log.stop("Reading Files");

    // 2. delete rows with empty cells in the clustering columns
    std::printf("\nStarting clustering process... \n");
    log.start("Do clustering");
    log.start("Deleting Empty Rows");
    const std::string vcol = column_setting(cfg, "gene_columns.first_gene.name");
    const std::string jcol = column_setting(cfg, "gene_columns.second_gene.name");
    const std::string scol = column_setting(cfg, "sequence_columns.name");
    for (const std::string &c : {scol, vcol, jcol}) {
        if (c.empty()) continue;
        const size_t idx = table.index_of(c);
        std::vector<int> flag(table.rows());
        for (size_t r = 0; r < flag.size(); ++r) flag[r] = table.cols[idx][r].empty();
        log.message("Deleting empty rows for column " + c);
        size_t gone = 0;
        for (int f : flag) gone += f != 0;
        log.message("Deleting " + std::to_string(gone) + " rows");
        table.delete_rows(flag);
    }
    log.stop("Deleting Empty Rows");

    // 3. cluster
    log.start("Transferring data to device");
    const size_t n = table.rows();
    thrust::host_vector<std::string> vcalls, jcalls, seqs, vnames, jnames;
    if (vcol.empty()) {
        vcalls.assign(n, "V");
        vnames.assign(1, "V");
    } else {
        const auto &c = table.column(vcol);
        vcalls.assign(c.begin(), c.end());
        vnames = read_genes(column_setting(cfg, "gene_columns.first_gene.ref"));
    }
    if (jcol.empty()) {
        jcalls.assign(n, "V");
        jnames.assign(1, "V");
    } else {
        const auto &c = table.column(jcol);
        jcalls.assign(c.begin(), c.end());
        jnames = read_genes(column_setting(cfg, "gene_columns.second_gene.ref"));
    }
    if (scol.empty()) {
        seqs.assign(n, "Gene");
    } else {
        const auto &c = table.column(scol);
        seqs.assign(c.begin(), c.end());
    }
    log.stop("Transferring data to device");
    NormalStringVectorZipIteratorType zip =
        thrust::make_zip_iterator(thrust::make_tuple(vcalls.begin(), jcalls.begin(), seqs.begin()));
    InterlinkClass il(zip, (unsigned)n, vnames.begin(), (unsigned)vnames.size(), jnames.begin(),
                      (unsigned)jnames.size(), cfg.get_float("maximum_relative_edit_distance"),
                      cfg.get_uint("minimum_number_of_neighbors"), cfg.get_bool("use_hamming", false),
                      cfg.get_bool("get_tree_depth_statistics", false));
    table.append_column("clone_id", il.get_clusters_iterator());
    table.append_column("copies", il.get_copies_iterator());
    table.append_column("cores", il.get_cores_iterator());
    log.start("Do clustering");

    // 4. write
    std::printf("\nWriting output ... \n");
    log.start("Writing Output");
    if (cfg.get_bool("remove_noise_from_output")) {
        log.message("Deleting noise points");
        std::vector<int> flag(table.rows());
        const auto &cid = table.column("clone_id");
        size_t gone = 0;
        for (size_t r = 0; r < flag.size(); ++r) gone += (flag[r] = cid[r] == "0");
        log.message("Deleting " + std::to_string(gone) + " rows");
        table.delete_rows(flag);
    }
    std::vector<std::string> wkeys, wheaders;
    pairs_of(cfg.at("select_columns_to_write"), wkeys, wheaders);
    for (const auto &k : wkeys)
        if (std::find(table.names.begin(), table.names.end(), k) == table.names.end()) table.append_empty_column(k);
    const std::string out = output_name(cfg, inputs[0], "output");
    log.message("Writing to file " + out);
    table.write(out, wkeys, wheaders, cfg.get_char("output_settings.separator"),
                cfg.get_char("output_settings.endline"), cfg.get_char("output_settings.quotation"));
    log.start("Writing Output");
    std::printf("\nFinished!\n\n");
    log.stop("Overall");
    std::printf("\n\n***************************************************************************************************\n");

    // 5. the run log
    const std::string bench = output_name(cfg, inputs[0], "benchmark_output");
    if (!log.write(bench)) throw std::runtime_error("Cannot write to file " + bench);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        return run(argc, argv);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
