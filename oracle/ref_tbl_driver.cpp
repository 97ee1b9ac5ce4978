// Test infrastructure only: a driver around the reference's own CSV reader
// (dogqc/include/csv.h), built into oracle/_ref/tbl_ref by oracle/Makefile
// when the reference sources are present.  It reads a .tbl file the way the
// reference's generated importer does (dbio.py:172-244): a CSVReader with one
// column more than the table (the terminal '|'), trim_chars<' '>,
// no_quote_escape<'|'>, throw_on_overflow, each column parsed by csv.h's
// parse() for the C type the importer declares (dbio.py:185-191;
// codegen.py:37-45), dates by the schema's fixed-position toDate
// (query/schema/tpch.py:9-20).
//
// Usage: tbl_ref <file.tbl> <types>, one type letter per column:
//   i INT (int), l INT64 (long long), f FLOAT (float), d DOUBLE parsed as
//   double, c CHAR (char), t DATE, s STRING.
// Output: one line per row, tab-separated: integers in decimal, float and
// double as the hex of their bit patterns, strings as hex of their bytes.
// A parse error prints "ERROR <line> <what>" and exits 1.
#include <array>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>

#include "csv.h"

namespace {

int dateOf(const char* c) {
    // the schema's toDate: digits at fixed positions of YYYY-MM-DD
    const int w[8] = {10000000, 1000000, 100000, 10000, 1000, 100, 10, 1};
    const int pos[8] = {0, 1, 2, 3, 5, 6, 8, 9};
    int d = 0;
    for (int k = 0; k < 8; ++k) d += (int)(c[pos[k]] - 48) * w[k];
    return d;
}

void emitColumn(char type, char* col, std::string& out) {
    char buf[32];
    switch (type) {
        case 'i': {
            int x;
            io::detail::parse<io::throw_on_overflow>(col, x);
            std::snprintf(buf, sizeof buf, "%d", x);
            out += buf;
            break;
        }
        case 'l': {
            long long x;
            io::detail::parse<io::throw_on_overflow>(col, x);
            std::snprintf(buf, sizeof buf, "%lld", x);
            out += buf;
            break;
        }
        case 'f': {
            float x;
            io::detail::parse<io::throw_on_overflow>(col, x);
            uint32_t b;
            std::memcpy(&b, &x, 4);
            std::snprintf(buf, sizeof buf, "%08" PRIx32, b);
            out += buf;
            break;
        }
        case 'd': {
            double x;
            io::detail::parse<io::throw_on_overflow>(col, x);
            uint64_t b;
            std::memcpy(&b, &x, 8);
            std::snprintf(buf, sizeof buf, "%016" PRIx64, b);
            out += buf;
            break;
        }
        case 'c': {
            char x;
            io::detail::parse<io::throw_on_overflow>(col, x);
            std::snprintf(buf, sizeof buf, "%d", (int)(signed char)x);
            out += buf;
            break;
        }
        case 't': {
            std::string x;
            io::detail::parse<io::throw_on_overflow>(col, x);
            std::snprintf(buf, sizeof buf, "%u", (unsigned)dateOf(x.c_str()));
            out += buf;
            break;
        }
        case 's': {
            std::string x;
            io::detail::parse<io::throw_on_overflow>(col, x);
            out += "s";
            for (unsigned char ch : x) {
                std::snprintf(buf, sizeof buf, "%02x", ch);
                out += buf;
            }
            break;
        }
    }
}

template <unsigned N, size_t... I>
bool readRow(io::CSVReader<N, io::trim_chars<' '>, io::no_quote_escape<'|'>>& r, std::array<char*, N>& cols,
             std::index_sequence<I...>) {
    return r.read_row(cols[I]...);
}

template <unsigned N>
int run(const char* path, const char* types) {
    io::CSVReader<N, io::trim_chars<' '>, io::no_quote_escape<'|'>> r(path);
    std::array<char*, N> cols;
    unsigned long line = 0;
    std::string row;
    try {
        for (;;) {
            ++line;
            if (!readRow<N>(r, cols, std::make_index_sequence<N>{})) break;
            row.clear();
            for (unsigned k = 0; k + 1 < N; ++k) {
                if (k) row += "\t";
                emitColumn(types[k], cols[k], row);
            }
            std::printf("%s\n", row.c_str());
        }
    } catch (const io::error::base& e) {
        std::printf("ERROR %lu %s\n", line, e.what());
        return 1;
    }
    return 0;
}

template <unsigned N>
int dispatch(unsigned n, const char* path, const char* types) {
    if constexpr (N > 18) {
        return 2;
    } else {
        return n == N ? run<N>(path, types) : dispatch<N + 1>(n, path, types);
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <file.tbl> <types>\n", argv[0]);
        return 2;
    }
    const unsigned ncols = (unsigned)std::strlen(argv[2]);
    try {
        return dispatch<2>(ncols + 1, argv[1], argv[2]);
    } catch (const std::exception& e) {
        std::printf("ERROR 0 %s\n", e.what());
        return 1;
    }
}
