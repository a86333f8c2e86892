// CsvTable.hpp — the table of the reference's host program: CSV input,
// the row deletions around the clustering, and CSV output.
//
// Follows cpp/csvreader.hpp, cpp/TableHolder.hpp and cpp/csvwriter.hpp:
//   * a line is split into fields by the separator; the quotation character
//     toggles quoting (it is dropped, and separators inside quotes are
//     kept); a backslash escapes a backslash, the quotation character or the
//     separator, and "\n" stands for a newline; any other escape is an
//     error.  An empty line has no fields, and a line that ends with a
//     separator has one more, empty, field (the rules of the escaped-list
//     tokenizer the reference uses, cpp/csvreader.hpp:47);
//   * the header line maps the configured header names to fields; a missing
//     one is an error; the reference reads lines up to end of file, so a
//     final line break yields one more, empty, row (:149-172);
//   * column 0 holds the file each row came from ("file_origin");
//   * output: the header names joined by the separator, then each row with
//     every field wrapped in the quotation character unless it is '\0'
//     (cpp/csvwriter.hpp:37-73).
#pragma once
#include <algorithm>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace idb {

inline std::vector<std::string> split_fields(const std::string &line, char sep, char quote)
{
    std::vector<std::string> out;
    if (line.empty()) return out;
    std::string tok;
    bool in_quote = false;
    for (size_t i = 0; i < line.size(); ++i) {
        const char c = line[i];
        if (c == '\\') {
            if (++i == line.size()) throw std::runtime_error("cannot end with escape");
            const char e = line[i];
            if (e == 'n') tok += '\n';
            else if (e == quote || e == sep || e == '\\') tok += e;
            else throw std::runtime_error("unknown escape sequence");
        } else if (c == sep && !in_quote) {
            out.push_back(tok);
            tok.clear();
        } else if (c == quote) {
            in_quote = !in_quote;
        } else {
            tok += c;
        }
    }
    out.push_back(tok);  // after a final separator: the extra empty field
    return out;
}

class Table {
  public:
    std::vector<std::string> names;                 // column names ("file_origin", then the configured keys)
    std::vector<std::vector<std::string>> cols;     // one vector of strings per column

    explicit Table(const std::vector<std::string> &keys)
    {
        names.push_back("file_origin");
        names.insert(names.end(), keys.begin(), keys.end());
        cols.resize(names.size());
    }

    size_t rows() const { return cols.empty() ? 0 : cols[0].size(); }

    size_t index_of(const std::string &name) const
    {
        const auto it = std::find(names.begin(), names.end(), name);
        if (it == names.end()) throw std::runtime_error("Column " + name + " not contained in TableHolder!");
        return (size_t)(it - names.begin());
    }

    std::vector<std::string> &column(const std::string &name) { return cols[index_of(name)]; }

    // Rows of one CSV file; headers[k] is the file's header name for column k + 1.
    void read_file(const std::string &file, const std::vector<std::string> &headers, char sep, char endline, char quote)
    {
        std::ifstream in(file);
        if (!in.is_open()) throw std::runtime_error("Cannot read from file " + file);
        std::string line;
        std::getline(in, line, endline);
        purify(line);
        const std::vector<std::string> head = split_fields(line, sep, quote);
        std::vector<int> where(head.size(), -1);  // field -> column
        for (size_t k = 0; k < headers.size(); ++k) {
            const auto pos = std::find(head.begin(), head.end(), headers[k]);
            if (pos == head.end()) throw std::runtime_error("Required column " + headers[k] + " not found in " + file);
            where[(size_t)(pos - head.begin())] = (int)k + 1;
        }
        while (!in.eof()) {
            std::getline(in, line, endline);
            purify(line);
            const std::vector<std::string> f = split_fields(line, sep, quote);
            for (auto &c : cols) c.emplace_back();
            cols[0].back() = file;
            for (size_t n = 0; n < f.size() && n < where.size(); ++n)
                if (where[n] >= 0) cols[(size_t)where[n]].back() = f[n];
        }
    }

    // Deletes the rows whose flag is set; returns how many went.
    size_t delete_rows(const std::vector<int> &flag)
    {
        size_t gone = 0;
        for (int f : flag) gone += f != 0;
        for (auto &c : cols) {
            size_t t = 0;
            for (size_t r = 0; r < c.size(); ++r) {
                if (flag[r]) continue;
                if (t != r) c[t] = std::move(c[r]);  // a self-move would empty the string
                ++t;
            }
            c.resize(t);
        }
        return gone;
    }

    size_t delete_where_equal(const std::string &name, const std::string &value)
    {
        const std::vector<std::string> &c = cols[index_of(name)];
        std::vector<int> flag(c.size());
        for (size_t r = 0; r < c.size(); ++r) flag[r] = c[r] == value;
        return delete_rows(flag);
    }

    template <typename It>
    void append_column(const std::string &name, It first)
    {
        std::vector<std::string> c(rows());
        for (size_t r = 0; r < c.size(); ++r, ++first) c[r] = std::to_string(*first);
        names.push_back(name);
        cols.push_back(std::move(c));
    }

    void append_empty_column(const std::string &name)
    {
        names.push_back(name);
        cols.emplace_back(rows());
    }

    // The selected columns (by name, in order) under the given header names.
    void write(const std::string &file, const std::vector<std::string> &select, const std::vector<std::string> &headers,
               char sep, char endline, char quote) const
    {
        std::ofstream out(file);
        if (!out.is_open()) throw std::runtime_error("Cannot write to file " + file);
        for (size_t k = 0; k < headers.size(); ++k) {
            out << headers[k];
            if (k + 1 < headers.size()) out << sep;
        }
        out << endline;
        std::vector<size_t> idx;
        for (const auto &s : select) idx.push_back(index_of(s));
        const size_t n = rows();
        for (size_t r = 0; r < n; ++r) {
            for (size_t k = 0; k < idx.size(); ++k) {
                if (quote != '\0') out << quote << cols[idx[k]][r] << quote;
                else out << cols[idx[k]][r];
                if (k + 1 < idx.size()) out << sep;
            }
            out << endline;
        }
    }

  private:
    static void purify(std::string &line)
    {
        if (!line.empty() && line.back() == '\r') line.pop_back();
    }
};

}  // namespace idb
