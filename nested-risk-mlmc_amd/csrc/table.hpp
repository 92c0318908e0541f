// Text table with auto-widened columns, laid out like the reference's level
// table (table_output.hpp:38-112, ASCII mode): a "| a | b |" row per line,
// headers centred, values right-aligned, and a "+---+---+" rule under the
// header and under the last row.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <functional>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

namespace nmlmc {

struct table_t {
    struct field_t {
        std::string header;
        size_t width;
        std::function<void(std::stringstream&, unsigned)> value;
    };
    explicit table_t(unsigned rows) : rows(rows) {}
    unsigned rows;
    std::vector<field_t> fields;

    void add(const std::string& header, int min_len, std::function<void(std::stringstream&, unsigned)> v) {
        fields.push_back(field_t{header, (size_t)min_len, v});
    }

    // Cell i of field f: row 0 is the header.
    static void cell(std::stringstream& o, const field_t& f, unsigned i) {
        if (i == 0) {
            const std::streamsize w = o.width();
            const std::string& h = f.header;
            if (w > (std::streamsize)h.size()) {
                const std::streamsize left = (w + (std::streamsize)h.size()) / 2;
                o.width(left);
                o << h;
                o.width(w - left);
                o << "";
            } else {
                o << h;
            }
        } else {
            f.value(o, i - 1);
        }
    }
};

inline std::stringstream& operator<<(std::stringstream& s, table_t& t) {
    const unsigned rows = t.rows + 1;
    for (auto& f : t.fields) {
        for (unsigned i = 0; i < rows; i++) {
            std::stringstream tmp;
            tmp.copyfmt(s);
            tmp << std::setw(0);
            table_t::cell(tmp, f, i);
            f.width = std::max(tmp.str().size(), f.width);
        }
    }
    for (unsigned i = 0; i < rows; i++) {
        s << "| ";
        for (const auto& f : t.fields) {
            s << std::setw(f.width);
            table_t::cell(s, f, i);
            s << " | ";
        }
        s << std::endl;
        if (i == 0 || i == rows - 1) {
            s << "+" << std::setfill('-');
            for (const auto& f : t.fields) s << std::setw(3 + f.width) << "+";
            s << std::setfill(' ') << std::endl;
        }
    }
    return s;
}

}  // namespace nmlmc
