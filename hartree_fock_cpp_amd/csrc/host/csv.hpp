// csv.hpp -- minimal CSV reading and writing for basis files and results.
//
// Writing matches the reference's result files (written through rapidcsv,
// external/rapidcsv/rapidcsv.h:133-170): one header row, comma separators,
// doubles printed with std::setprecision(17) in the default float format,
// integers and strings as-is.
#pragma once
#include <fstream>
#include <iomanip>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

namespace hfxh {

struct CsvTable {
    std::vector<std::string> header;
    std::map<std::string, std::vector<std::string>> cols;
    const std::vector<std::string>& column(const std::string& name) const {
        auto it = cols.find(name);
        if (it == cols.end()) throw std::runtime_error("CSV column not found: " + name);
        return it->second;
    }
};

inline std::vector<std::string> csv_split(const std::string& line) {
    std::vector<std::string> out;
    std::string cur;
    for (char ch : line) {
        if (ch == ',') {
            out.push_back(cur);
            cur.clear();
        } else if (ch != '\r' && ch != '\n') {
            cur.push_back(ch);
        }
    }
    out.push_back(cur);
    return out;
}

inline CsvTable read_csv(const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("Could not open CSV file: " + path);
    CsvTable t;
    std::string line;
    if (!std::getline(f, line)) throw std::runtime_error("empty CSV file: " + path);
    t.header = csv_split(line);
    for (auto& h : t.header) t.cols[h];
    while (std::getline(f, line)) {
        if (line.empty() || line == "\r") continue;
        auto v = csv_split(line);
        for (size_t i = 0; i < t.header.size(); ++i) t.cols[t.header[i]].push_back(i < v.size() ? v[i] : "");
    }
    return t;
}

using CsvColumn = std::variant<std::vector<double>, std::vector<int>, std::vector<std::string>>;

inline std::string fmt_double(double v) {
    std::ostringstream o;
    o << std::setprecision(17) << v;
    return o.str();
}

inline void write_csv(const std::string& path, const std::vector<std::string>& names,
                      const std::vector<CsvColumn>& columns) {
    if (names.size() != columns.size()) throw std::runtime_error("number of columns is different from number of headers");
    std::ofstream f(path);
    if (!f.is_open()) throw std::runtime_error("Could not open CSV file for writing: " + path);
    size_t rows = 0;
    for (const auto& c : columns) rows = std::max(rows, std::visit([](const auto& v) { return v.size(); }, c));
    for (size_t i = 0; i < names.size(); ++i) f << (i ? "," : "") << names[i];
    f << "\n";
    for (size_t r = 0; r < rows; ++r) {
        for (size_t i = 0; i < columns.size(); ++i) {
            if (i) f << ",";
            std::visit(
                [&](const auto& v) {
                    using T = typename std::decay_t<decltype(v)>::value_type;
                    if (r >= v.size()) return;
                    if constexpr (std::is_same_v<T, double>) {
                        f << fmt_double(v[r]);
                    } else if constexpr (std::is_same_v<T, std::string>) {
                        // quoted when it holds a separator, space or line break (rapidcsv's auto-quote)
                        const std::string& c = v[r];
                        if (c.find_first_of(", \n") != std::string::npos) {
                            f << '"';
                            for (char ch : c) f << (ch == '"' ? std::string("\"\"") : std::string(1, ch));
                            f << '"';
                        } else {
                            f << c;
                        }
                    } else {
                        f << v[r];
                    }
                },
                columns[i]);
        }
        f << "\n";
    }
}

}  // namespace hfxh
