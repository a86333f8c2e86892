// frontend_check.cpp — CPU driver for the front end's parsing and table code
// (InfoConfig.hpp, CsvTable.hpp), built with g++ by tests/test_frontend.py.
//
//   frontend_check conf FILE PATH...        value of each dotted PATH
//   frontend_check children FILE PATH       the (key, value) pairs under PATH
//   frontend_check split SEP QUOTE          fields of each stdin line
//   frontend_check table IN OUT SEP QUOTE OUTSEP OUTQUOTE DROP key=header...
//        read IN, delete rows whose DROP column is empty, write every key
//        under its header to OUT
// Characters are given as decimal codes.  Output strings are printed with
// every byte outside 0x21..0x7e as \xHH, fields separated by '|'.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "CsvTable.hpp"
#include "InfoConfig.hpp"

static std::string show(const std::string &s)
{
    std::string o;
    char buf[8];
    for (unsigned char c : s) {
        if (c > 0x20 && c < 0x7f && c != '|' && c != '\\') {
            o += (char)c;
        } else {
            std::snprintf(buf, sizeof buf, "\\x%02x", c);
            o += buf;
        }
    }
    return o;
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const std::string cmd = argv[1];
    try {
        if (cmd == "conf" && argc >= 3) {
            const idb::InfoNode root = idb::InfoParser::read_file(argv[2]);
            for (int i = 3; i < argc; ++i) {
                const idb::InfoNode *n = root.find(argv[i]);
                std::printf("%s|%s\n", argv[i], n ? show(n->data).c_str() : "<absent>");
            }
        } else if (cmd == "children" && argc == 4) {
            const idb::InfoNode root = idb::InfoParser::read_file(argv[2]);
            for (const auto &c : root.at(argv[3]).children)
                std::printf("%s|%s\n", show(c.first).c_str(), show(c.second.data).c_str());
        } else if (cmd == "split" && argc == 4) {
            const char sep = (char)std::atoi(argv[2]), quote = (char)std::atoi(argv[3]);
            std::string line;
            while (std::getline(std::cin, line)) {
                try {
                    const auto f = idb::split_fields(line, sep, quote);
                    std::string o = std::to_string(f.size());
                    for (const auto &x : f) o += "|" + show(x);
                    std::printf("%s\n", o.c_str());
                } catch (const std::exception &e) {
                    std::printf("error\n");
                }
            }
        } else if (cmd == "table" && argc >= 10) {
            std::vector<std::string> keys, heads;
            for (int i = 9; i < argc; ++i) {
                const std::string kv = argv[i];
                keys.push_back(kv.substr(0, kv.find('=')));
                heads.push_back(kv.substr(kv.find('=') + 1));
            }
            idb::Table t(keys);
            t.read_file(argv[2], heads, (char)std::atoi(argv[4]), '\n', (char)std::atoi(argv[5]));
            const size_t read = t.rows();
            const size_t gone = t.delete_where_equal(argv[8], "");
            std::vector<std::string> wk{"file_origin"}, wh{"origin"};
            wk.insert(wk.end(), keys.begin(), keys.end());
            wh.insert(wh.end(), heads.begin(), heads.end());
            t.write(argv[3], wk, wh, (char)std::atoi(argv[6]), '\n', (char)std::atoi(argv[7]));
            std::printf("read %zu deleted %zu\n", read, gone);
        } else {
            return 2;
        }
    } catch (const std::exception &e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
    return 0;
}
