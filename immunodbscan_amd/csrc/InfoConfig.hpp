// InfoConfig.hpp — the configuration file of the reference's host program.
//
// psdbscan reads its settings with boost::property_tree's INFO parser
// (cpp/ConfigHolder.hpp:56-59, data/psdbscan.conf).  Boost is not part of
// this build, so this header reads the subset of that format the program's
// files use, into an ordered tree:
//   * one entry per line: a key, optionally followed by a value;
//   * keys and values are bare words or double-quoted strings with the C
//     escapes \0 \a \b \f \n \r \t \v \" \' \\ ;
//   * "{" (on the key's line or the next one) opens the children of the
//     last entry, "}" closes them;
//   * ";" starts a comment outside quotes;
//   * #include "file" reads another file (path relative to the including
//     file) into the current level.
// Lookups follow dotted paths to the first child of each name, as
// ptree::get does; children keep file order, which the column selections
// rely on (cpp/ConfigHolder.hpp:120-170).
#pragma once
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace idb {

struct InfoNode {
    std::string data;
    std::vector<std::pair<std::string, InfoNode>> children;

    const InfoNode *child(const std::string &key) const
    {
        for (const auto &c : children)
            if (c.first == key) return &c.second;
        return nullptr;
    }

    // The node at a dotted path, or null.
    const InfoNode *find(const std::string &path) const
    {
        const InfoNode *n = this;
        size_t pos = 0;
        while (n) {
            const size_t dot = path.find('.', pos);
            n = n->child(path.substr(pos, dot == std::string::npos ? std::string::npos : dot - pos));
            if (dot == std::string::npos) break;
            pos = dot + 1;
        }
        return n;
    }

    const InfoNode &at(const std::string &path) const
    {
        const InfoNode *n = find(path);
        if (!n) throw std::runtime_error("No such node (" + path + ")");
        return *n;
    }

    std::string get_string(const std::string &path) const { return at(path).data; }

    std::string get_string(const std::string &path, const std::string &fallback) const
    {
        const InfoNode *n = find(path);
        return n ? n->data : fallback;
    }

    // A character setting: the first character of the value, whitespace
    // included ("\t"), and '\0' for "\0".
    char get_char(const std::string &path) const
    {
        const std::string &d = at(path).data;
        if (d.size() != 1) throw std::runtime_error("conversion of data to type \"char\" failed (" + path + ")");
        return d[0];
    }

    float get_float(const std::string &path) const
    {
        const std::string &d = at(path).data;
        char *end = nullptr;
        const float v = std::strtof(d.c_str(), &end);
        if (d.empty() || !end || *end) throw std::runtime_error("conversion of data to type \"float\" failed (" + path + ")");
        return v;
    }

    unsigned int get_uint(const std::string &path) const
    {
        const std::string &d = at(path).data;
        char *end = nullptr;
        const unsigned long v = std::strtoul(d.c_str(), &end, 10);
        if (d.empty() || !end || *end || d[0] == '-')
            throw std::runtime_error("conversion of data to type \"unsigned int\" failed (" + path + ")");
        return (unsigned int)v;
    }

    bool get_bool(const std::string &path) const
    {
        const std::string &d = at(path).data;
        if (d == "true" || d == "1") return true;
        if (d == "false" || d == "0") return false;
        throw std::runtime_error("conversion of data to type \"bool\" failed (" + path + ")");
    }

    bool get_bool(const std::string &path, bool fallback) const
    {
        return find(path) ? get_bool(path) : fallback;
    }
};

class InfoParser {
  public:
    static InfoNode read_file(const std::string &path)
    {
        InfoNode root;
        InfoParser p;
        p.read_into(path, root, 0);
        return root;
    }

  private:
    static std::string dir_of(const std::string &path)
    {
        const size_t s = path.find_last_of('/');
        return s == std::string::npos ? std::string() : path.substr(0, s + 1);
    }

    [[noreturn]] static void error(const std::string &file, unsigned line, const std::string &what)
    {
        throw std::runtime_error(file + "(" + std::to_string(line) + "): " + what);
    }

    // One token from line[i..]: a quoted string (escapes resolved) or a bare
    // word.  Returns false at the end of the line or at a comment.
    static bool token(const std::string &line, size_t &i, std::string &out, bool &is_brace, const std::string &file,
                      unsigned ln)
    {
        while (i < line.size() && (line[i] == ' ' || line[i] == '\t' || line[i] == '\r')) ++i;
        if (i >= line.size() || line[i] == ';') return false;
        out.clear();
        is_brace = false;
        if (line[i] == '{' || line[i] == '}') {
            out = line[i++];
            is_brace = true;
            return true;
        }
        if (line[i] == '"') {
            ++i;
            while (i < line.size() && line[i] != '"') {
                char c = line[i++];
                if (c == '\\') {
                    if (i >= line.size()) error(file, ln, "unterminated escape");
                    const char e = line[i++];
                    switch (e) {
                    case '0': c = '\0'; break;
                    case 'a': c = '\a'; break;
                    case 'b': c = '\b'; break;
                    case 'f': c = '\f'; break;
                    case 'n': c = '\n'; break;
                    case 'r': c = '\r'; break;
                    case 't': c = '\t'; break;
                    case 'v': c = '\v'; break;
                    case '"': c = '"'; break;
                    case '\'': c = '\''; break;
                    case '\\': c = '\\'; break;
                    default: error(file, ln, std::string("unknown escape sequence \\") + e);
                    }
                }
                out.push_back(c);
            }
            if (i >= line.size()) error(file, ln, "unterminated string");
            ++i;
            return true;
        }
        while (i < line.size() && line[i] != ' ' && line[i] != '\t' && line[i] != '\r' && line[i] != ';' &&
               line[i] != '{' && line[i] != '}')
            out.push_back(line[i++]);
        return true;
    }

    void read_into(const std::string &path, InfoNode &root, int depth)
    {
        if (depth > 16) throw std::runtime_error(path + ": #include nested too deeply");
        std::ifstream f(path);
        if (!f.is_open()) throw std::runtime_error(path + ": cannot open file");
        std::vector<InfoNode *> stack{&root};
        InfoNode *last = nullptr;  // the last entry of the current level
        std::string line, tok;
        unsigned ln = 0;
        while (std::getline(f, line)) {
            ++ln;
            size_t i = 0;
            bool brace = false;
            if (!token(line, i, tok, brace, path, ln)) continue;
            if (!brace && tok == "#include") {
                std::string file;
                if (!token(line, i, file, brace, path, ln) || brace) error(path, ln, "#include needs a file name");
                read_into(dir_of(path) + file, *stack.back(), depth + 1);
                last = nullptr;
                continue;
            }
            for (;;) {
                if (brace && tok == "{") {
                    if (!last) error(path, ln, "'{' without a key");
                    stack.push_back(last);
                    last = nullptr;
                } else if (brace) {  // "}"
                    if (stack.size() < 2) error(path, ln, "unmatched '}'");
                    stack.pop_back();
                    last = nullptr;
                } else {
                    InfoNode *parent = stack.back();
                    parent->children.emplace_back(tok, InfoNode{});
                    last = &parent->children.back().second;
                    std::string val;
                    bool vb = false;
                    const size_t save = i;
                    if (token(line, i, val, vb, path, ln)) {
                        if (vb) i = save;  // a brace after the key: handled next
                        else last->data = val;
                    }
                }
                if (!token(line, i, tok, brace, path, ln)) break;
            }
        }
        if (stack.size() != 1) throw std::runtime_error(path + ": unmatched '{'");
    }
};

}  // namespace idb
