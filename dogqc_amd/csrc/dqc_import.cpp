// libdqcimport.so: .tbl text tables into mmdb column files (C ABI in
// include/dogqc_amd_import.h).
//
// The input is memory-mapped and cut into chunks at line boundaries; worker
// threads take chunks from a shared counter.  Pass 1 counts each chunk's
// rows and STRING bytes and checks field counts; prefix sums over the chunks
// give every chunk its first row and its offset in each char heap; pass 2
// parses the fields and writes them straight into the memory-mapped output
// files.  Parsing restates csv.h's rules (see the header): fields split at
// '|' with no quoting, spaces trimmed, a '\r' before the '\n' dropped.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "dogqc_amd.h"
#include "dogqc_amd_import.h"

namespace {

constexpr int64_t CHUNK_BYTES = 1 << 20;
constexpr int64_t NO_ERROR = std::numeric_limits<int64_t>::max();

struct Input {
    int fd = -1;
    const char* data = nullptr;
    int64_t size = 0;
    ~Input() {
        if (data && size) munmap((void*)data, (size_t)size);
        if (fd >= 0) close(fd);
    }
    bool open(const char* path) {
        fd = ::open(path, O_RDONLY);
        if (fd < 0) return false;
        struct stat sb;
        if (fstat(fd, &sb) != 0) return false;
        size = (int64_t)sb.st_size;
        if (size == 0) return true;
        void* p = mmap(nullptr, (size_t)size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (p == MAP_FAILED) {
            size = 0;
            return false;
        }
        madvise(p, (size_t)size, MADV_SEQUENTIAL);
        data = (const char*)p;
        return true;
    }
};

// an mmdb file being written: header, payload, NUL
struct Output {
    std::string path;
    int fd = -1;
    char* map = nullptr;
    int64_t bytes = 0;
    char* payload() { return map + 8; }
    bool create(const std::string& p, int64_t payloadBytes) {
        path = p;
        fd = ::open(p.c_str(), O_RDWR | O_CREAT | O_TRUNC, 0644);
        if (fd < 0) return false;
        bytes = 8 + payloadBytes + 1;
        if (ftruncate(fd, (off_t)bytes) != 0) return false;
        void* m = mmap(nullptr, (size_t)bytes, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
        if (m == MAP_FAILED) return false;
        map = (char*)m;
        const uint64_t header = (uint64_t)bytes;
        std::memcpy(map, &header, 8);     // little-endian host
        map[bytes - 1] = 0;
        return true;
    }
    bool finish() {
        bool ok = true;
        if (map) ok = munmap(map, (size_t)bytes) == 0;
        map = nullptr;
        if (fd >= 0) ok = (close(fd) == 0) && ok;
        fd = -1;
        return ok;
    }
    void discard() {
        finish();
        if (!path.empty()) unlink(path.c_str());
    }
};

struct Field {
    const char* b;
    const char* e;
};

// chunk [begin, end) of whole lines
struct Chunk {
    int64_t begin, end;
    int64_t rows = 0;
    int64_t firstRow = 0;
    std::vector<int64_t> charBytes, charStart;   // per column (STRING only)
    int64_t errLine = NO_ERROR;                   // local 0-based line of the first error
    int32_t errColumn = -1, errKind = 0;
};

inline bool nextLine(const char* data, int64_t end, int64_t& pos, const char*& lb, const char*& le) {
    if (pos >= end) return false;
    lb = data + pos;
    const char* nl = (const char*)std::memchr(lb, '\n', (size_t)(end - pos));
    le = nl ? nl : data + end;
    pos = (le - data) + (nl ? 1 : 0);
    if (le > lb && le[-1] == '\r') --le;
    return true;
}

// split a line at '|', trimming spaces; false unless it has exactly n fields
inline bool splitFields(const char* b, const char* e, int32_t n, Field* f) {
    int32_t k = 0;
    const char* s = b;
    for (;;) {
        const char* p = (const char*)std::memchr(s, '|', (size_t)(e - s));
        const char* fe = p ? p : e;
        if (k == n) return false;
        const char* tb = s;
        const char* te = fe;
        while (tb < te && *tb == ' ') ++tb;
        while (te > tb && te[-1] == ' ') --te;
        f[k++] = Field{tb, te};
        if (!p) break;
        s = p + 1;
    }
    return k == n;
}

template <typename T>
int parseInt(const Field& f, T& out) {
    const char* c = f.b;
    T x = 0;
    if (c < f.e && *c == '-') {
        for (++c; c < f.e; ++c) {
            if (*c < '0' || *c > '9') return DQC_IMPORT_DIGIT;
            const T y = (T)(*c - '0');
            if (x < (std::numeric_limits<T>::min() + y) / 10) return DQC_IMPORT_OVERFLOW;
            x = (T)(10 * x - y);
        }
    } else {
        if (c < f.e && *c == '+') ++c;
        for (; c < f.e; ++c) {
            if (*c < '0' || *c > '9') return DQC_IMPORT_DIGIT;
            const T y = (T)(*c - '0');
            if (x > (std::numeric_limits<T>::max() - y) / 10) return DQC_IMPORT_OVERFLOW;
            x = (T)(10 * x + y);
        }
    }
    out = x;
    return 0;
}

// the exponent: a saturating signed integer that stops at saturation
int parseExponent(const char* c, const char* e, int& out) {
    int x = 0;
    if (c < e && *c == '-') {
        for (++c; c < e; ++c) {
            if (*c < '0' || *c > '9') return DQC_IMPORT_DIGIT;
            const int y = *c - '0';
            if (x < (std::numeric_limits<int>::min() + y) / 10) {
                out = std::numeric_limits<int>::min();
                return 0;
            }
            x = 10 * x - y;
        }
    } else {
        if (c < e && *c == '+') ++c;
        for (; c < e; ++c) {
            if (*c < '0' || *c > '9') return DQC_IMPORT_DIGIT;
            const int y = *c - '0';
            if (x > (std::numeric_limits<int>::max() - y) / 10) {
                out = std::numeric_limits<int>::max();
                return 0;
            }
            x = 10 * x + y;
        }
    }
    out = x;
    return 0;
}

// decimal text to T with every step rounded to T, as the reference's reader
// computes it (so FLOAT columns match it bit for bit)
template <typename T>
int parseDecimal(const Field& f, T& out) {
    const char* c = f.b;
    const char* e = f.e;
    bool neg = false;
    if (c < e && *c == '-') {
        neg = true;
        ++c;
    } else if (c < e && *c == '+') {
        ++c;
    }
    volatile T x = 0;     // volatile: keep each rounding step, no contraction
    while (c < e && *c >= '0' && *c <= '9') {
        x = x * (T)10;
        x = x + (T)(*c - '0');
        ++c;
    }
    if (c < e && (*c == '.' || *c == ',')) {
        ++c;
        volatile T pos = 1;
        while (c < e && *c >= '0' && *c <= '9') {
            pos = pos / (T)10;
            const T step = (T)(*c - '0') * pos;
            x = x + step;
            ++c;
        }
    }
    if (c < e && (*c == 'e' || *c == 'E')) {
        int ex = 0;
        if (int rc = parseExponent(c + 1, e, ex)) return rc;
        if (ex != 0) {
            volatile T base;
            int64_t k = ex;
            if (k < 0) {
                base = (T)0.1;
                k = -k;
            } else {
                base = (T)10;
            }
            while (k != 1) {
                if ((k & 1) == 0) {
                    base = base * base;
                    k >>= 1;
                } else {
                    x = x * base;
                    --k;
                }
            }
            x = x * base;
        }
    } else if (c < e) {
        return DQC_IMPORT_DIGIT;
    }
    out = neg ? -x : x;
    return 0;
}

int parseDate(const Field& f, uint32_t& out) {
    if (f.e - f.b < 10) return DQC_IMPORT_SHORT_DATE;
    static const int pos[8] = {0, 1, 2, 3, 5, 6, 8, 9};
    static const int w[8] = {10000000, 1000000, 100000, 10000, 1000, 100, 10, 1};
    int d = 0;
    for (int k = 0; k < 8; ++k) d += (int)((signed char)f.b[pos[k]] - 48) * w[k];
    out = (uint32_t)d;
    return 0;
}

struct Job {
    const char* data;
    int32_t ncols;
    const int32_t* types;
    std::vector<Chunk>* chunks;
    std::vector<Output>* outs;        // numeric: one per column; STRING: offsets at k, chars at ncols + k
    bool write;
};

// parse (and with job.write, store) the lines of one chunk; returns at the
// first bad field.  `limit` caps the lines looked at (validation before a
// field-count error).
void parseChunk(const Job& job, Chunk& ch, int64_t limit) {
    std::vector<Field> f(job.ncols + 1);
    std::vector<int64_t> heap(ch.charStart);
    int64_t pos = ch.begin, line = 0;
    const char *lb, *le;
    while (line < limit && nextLine(job.data, ch.end, pos, lb, le)) {
        if (!splitFields(lb, le, job.ncols + 1, f.data())) {
            ch.errLine = line;
            ch.errColumn = -1;
            ch.errKind = DQC_IMPORT_COLUMNS;
            return;
        }
        const int64_t row = ch.firstRow + line;
        for (int32_t k = 0; k < job.ncols; ++k) {
            int rc = 0;
            char* dst = job.write ? (*job.outs)[k].payload() : nullptr;
            switch (job.types[k]) {
                case DQC_COL_INT: {
                    int32_t v;
                    rc = parseInt(f[k], v);
                    if (!rc && dst) std::memcpy(dst + row * 4, &v, 4);
                    break;
                }
                case DQC_COL_INT64: {
                    long long v;
                    rc = parseInt(f[k], v);
                    if (!rc && dst) std::memcpy(dst + row * 8, &v, 8);
                    break;
                }
                case DQC_COL_FLOAT: {
                    float v;
                    rc = parseDecimal(f[k], v);
                    if (!rc && dst) std::memcpy(dst + row * 4, &v, 4);
                    break;
                }
                case DQC_COL_DOUBLE: {
                    double v;
                    rc = parseDecimal(f[k], v);
                    if (!rc && dst) std::memcpy(dst + row * 8, &v, 8);
                    break;
                }
                case DQC_COL_CHAR:
                    if (f[k].e - f[k].b != 1)
                        rc = DQC_IMPORT_SINGLE_CHAR;
                    else if (dst)
                        dst[row] = *f[k].b;
                    break;
                case DQC_COL_DATE: {
                    uint32_t v;
                    rc = parseDate(f[k], v);
                    if (!rc && dst) std::memcpy(dst + row * 4, &v, 4);
                    break;
                }
                default: {   // STRING
                    const int64_t len = f[k].e - f[k].b;
                    if (dst) {
                        std::memcpy((*job.outs)[job.ncols + k].payload() + heap[k], f[k].b, (size_t)len);
                        const uint64_t end = (uint64_t)(heap[k] + len);
                        std::memcpy(dst + (row + 1) * 8, &end, 8);
                    }
                    heap[k] += len;
                    break;
                }
            }
            if (rc) {
                ch.errLine = line;
                ch.errColumn = k;
                ch.errKind = rc;
                return;
            }
        }
        ++line;
    }
}

// pass 1: rows, STRING bytes, the first field-count error
void countChunk(const Job& job, Chunk& ch) {
    std::vector<Field> f(job.ncols + 1);
    int64_t pos = ch.begin;
    const char *lb, *le;
    while (nextLine(job.data, ch.end, pos, lb, le)) {
        if (!splitFields(lb, le, job.ncols + 1, f.data())) {
            if (ch.errLine == NO_ERROR) {
                ch.errLine = ch.rows;
                ch.errKind = DQC_IMPORT_COLUMNS;
                ch.errColumn = -1;
            }
        } else {
            for (int32_t k = 0; k < job.ncols; ++k)
                if (job.types[k] == DQC_COL_STRING) ch.charBytes[k] += f[k].e - f[k].b;
        }
        ++ch.rows;
    }
}

template <typename F>
void parallelFor(size_t n, int threads, F fn) {
    std::atomic<size_t> next{0};
    auto worker = [&]() {
        for (size_t i; (i = next.fetch_add(1)) < n;) fn(i);
    };
    const int t = (int)std::min<size_t>((size_t)std::max(1, threads), n);
    std::vector<std::thread> pool;
    for (int i = 1; i < t; ++i) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
}

std::vector<Chunk> cutChunks(const char* data, int64_t size, int32_t ncols) {
    std::vector<Chunk> chunks;
    int64_t b = 0;
    while (b < size) {
        int64_t e = std::min(size, b + CHUNK_BYTES);
        if (e < size) {
            const char* nl = (const char*)std::memchr(data + e, '\n', (size_t)(size - e));
            e = nl ? (nl - data) + 1 : size;
        }
        Chunk c;
        c.begin = b;
        c.end = e;
        c.charBytes.assign(ncols, 0);
        c.charStart.assign(ncols, 0);
        chunks.push_back(std::move(c));
        b = e;
    }
    return chunks;
}

int threadCount(int32_t threads) {
    if (threads > 0) return std::min(threads, 256);
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h, 256u));
}

const char* kindName(int k) {
    switch (k) {
        case DQC_IMPORT_COLUMNS: return "wrong number of fields";
        case DQC_IMPORT_DIGIT: return "invalid digit";
        case DQC_IMPORT_OVERFLOW: return "integer out of range";
        case DQC_IMPORT_SINGLE_CHAR: return "not a single character";
        case DQC_IMPORT_SHORT_DATE: return "date shorter than YYYY-MM-DD";
        default: return "error";
    }
}

}  // namespace

extern "C" {

int dqc_tbl_rows(const char* tbl_path, int64_t* rows) {
    if (!tbl_path || !rows) return DQC_EINVAL;
    Input in;
    if (!in.open(tbl_path)) return DQC_EIO;
    std::vector<Chunk> chunks = cutChunks(in.data, in.size, 0);
    std::vector<int64_t> counts(chunks.size(), 0);
    parallelFor(chunks.size(), threadCount(0), [&](size_t i) {
        int64_t pos = chunks[i].begin, n = 0;
        const char *lb, *le;
        while (nextLine(in.data, chunks[i].end, pos, lb, le)) ++n;
        counts[i] = n;
    });
    int64_t total = 0;
    for (int64_t c : counts) total += c;
    *rows = total;
    return DQC_OK;
}

int dqc_import_tbl(const char* tbl_path, int32_t ncols, const int32_t* types, const char* const* column_paths,
                   int32_t threads, dqc_import_report* report) {
    if (!report) return DQC_EINVAL;
    std::memset(report, 0, sizeof *report);
    report->error_column = -1;
    if (!tbl_path || ncols <= 0 || ncols > DQC_IMPORT_MAX_COLUMNS || !types || !column_paths) {
        std::snprintf(report->message, sizeof report->message, "invalid arguments");
        return DQC_EINVAL;
    }
    for (int32_t k = 0; k < ncols; ++k) {
        if (types[k] < DQC_COL_INT || types[k] > DQC_COL_STRING || !column_paths[k]) {
            std::snprintf(report->message, sizeof report->message, "invalid type or path of column %d", (int)k);
            return DQC_EINVAL;
        }
    }
    Input in;
    if (!in.open(tbl_path)) {
        report->error_kind = DQC_IMPORT_IO;
        std::snprintf(report->message, sizeof report->message, "cannot open %s: %s", tbl_path, std::strerror(errno));
        return DQC_EIO;
    }
    const int nthreads = threadCount(threads);
    std::vector<Chunk> chunks = cutChunks(in.data, in.size, ncols);
    std::vector<Output> outs(2 * (size_t)ncols);
    Job job{in.data, ncols, types, &chunks, &outs, false};

    parallelFor(chunks.size(), nthreads, [&](size_t i) { countChunk(job, chunks[i]); });
    int64_t rows = 0;
    std::vector<int64_t> heap(ncols, 0);
    for (Chunk& c : chunks) {
        c.firstRow = rows;
        rows += c.rows;
        for (int32_t k = 0; k < ncols; ++k) {
            c.charStart[k] = heap[k];
            heap[k] += c.charBytes[k];
        }
    }

    // the first chunk with a field-count error bounds where a first error can be
    size_t firstBad = chunks.size();
    for (size_t i = 0; i < chunks.size(); ++i)
        if (chunks[i].errLine != NO_ERROR) {
            firstBad = i;
            break;
        }
    auto fail = [&](const Chunk& c) {
        for (Output& o : outs) o.discard();
        report->rows = 0;
        report->error_line = c.firstRow + c.errLine + 1;
        report->error_column = c.errColumn;
        report->error_kind = c.errKind;
        std::snprintf(report->message, sizeof report->message, "%s line %lld field %d: %s", tbl_path,
                      (long long)report->error_line, (int)c.errColumn, kindName(c.errKind));
        return DQC_EIO;
    };
    if (firstBad < chunks.size()) {
        // validate the lines before the field-count error, in order
        std::vector<Chunk> check(chunks.begin(), chunks.begin() + (std::ptrdiff_t)firstBad + 1);
        const int64_t limit = chunks[firstBad].errLine;
        for (Chunk& c : check) c.errLine = NO_ERROR;
        parallelFor(check.size(), nthreads, [&](size_t i) {
            parseChunk(job, check[i], i == firstBad ? limit : NO_ERROR);
        });
        for (const Chunk& c : check)
            if (c.errLine != NO_ERROR) return fail(c);
        return fail(chunks[firstBad]);
    }

    for (int32_t k = 0; k < ncols; ++k) {
        const std::string base(column_paths[k]);
        bool ok;
        switch (types[k]) {
            case DQC_COL_INT:
            case DQC_COL_FLOAT:
            case DQC_COL_DATE: ok = outs[k].create(base, rows * 4); break;
            case DQC_COL_INT64:
            case DQC_COL_DOUBLE: ok = outs[k].create(base, rows * 8); break;
            case DQC_COL_CHAR: ok = outs[k].create(base, rows); break;
            default:
                ok = outs[k].create(base + "_offset", (rows + 1) * 8) &&
                     outs[ncols + k].create(base + "_char", heap[k]);
                if (ok) std::memset(outs[k].payload(), 0, 8);
                report->char_bytes[k] = heap[k];
                break;
        }
        if (!ok) {
            const int err = errno;
            for (Output& o : outs) o.discard();
            report->error_kind = DQC_IMPORT_IO;
            std::snprintf(report->message, sizeof report->message, "cannot write column file of %s: %s",
                          column_paths[k], std::strerror(err));
            return DQC_EIO;
        }
    }
    job.write = true;
    parallelFor(chunks.size(), nthreads, [&](size_t i) { parseChunk(job, chunks[i], NO_ERROR); });
    for (const Chunk& c : chunks)
        if (c.errLine != NO_ERROR) return fail(c);
    bool ok = true;
    for (Output& o : outs) ok = o.finish() && ok;
    if (!ok) {
        for (Output& o : outs) o.discard();
        report->error_kind = DQC_IMPORT_IO;
        std::snprintf(report->message, sizeof report->message, "writing the column files failed");
        return DQC_EIO;
    }
    report->rows = rows;
    return DQC_OK;
}

}  // extern "C"
