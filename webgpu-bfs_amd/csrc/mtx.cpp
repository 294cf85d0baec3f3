// Matrix Market (coordinate) reader: text -> canonical CSR.
//
// Behaviour follows crates/core/src/mtx.rs:124-287 (MtxReader::parse) and its
// error type crates/core/src/error.rs:45-81: the first non-empty line is the
// banner `%%MatrixMarket matrix <format> <field> <symmetry>` (case-insensitive,
// only `coordinate` accepted), then blank and `%` lines are skipped up to the
// `rows cols nnz` line, then every non-blank, non-comment line is a 1-based
// `row col [value...]` entry whose values are ignored.  Self loops are
// dropped, files marked symmetric / skew-symmetric / hermitian are mirrored,
// |V| = max(rows, cols), and the edge list becomes the canonical CSR of
// CsrGraph::from_edge_list (graph.rs:51-80: sorted rows, duplicates removed).
// The entry count in the size line is not enforced (the reference does not
// enforce it either).  The first error in file order wins; its line number
// counts every line of the file from 1.
//
// The entry lines are parsed in parallel: the body is cut into chunks at line
// boundaries, a first pass counts each chunk's lines (so every chunk knows
// the file line number it starts at) and a second pass parses the chunks
// into per-thread pair lists.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"

namespace {

using msbfs::fail;

// Error detail of the last failed parse on this thread (msbfs_mtx_last_error).
thread_local msbfs_mtx_error g_mtx_err{};

int mtx_fail(int kind, uint64_t line, uint64_t row, uint64_t col, uint64_t max_rows,
             uint64_t max_cols, const std::string &msg) {
    g_mtx_err = msbfs_mtx_error{};
    g_mtx_err.kind = kind;
    g_mtx_err.line_number = line;
    g_mtx_err.row = row;
    g_mtx_err.col = col;
    g_mtx_err.max_rows = max_rows;
    g_mtx_err.max_cols = max_cols;
    return fail(MSBFS_ERR_PARSE, "%s", msg.c_str());
}

inline bool is_space(char c) {
    return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f';
}

// One line of the buffer, without its terminating '\n'.
struct Span {
    const char *b, *e;
    std::string str() const { return std::string(b, e); }
};

Span trim(Span s) {
    while (s.b < s.e && is_space(*s.b)) ++s.b;
    while (s.e > s.b && is_space(s.e[-1])) --s.e;
    return s;
}

std::vector<Span> tokens(Span s) {
    std::vector<Span> t;
    const char *p = s.b;
    while (p < s.e) {
        while (p < s.e && is_space(*p)) ++p;
        const char *q = p;
        while (q < s.e && !is_space(*q)) ++q;
        if (q > p) t.push_back({p, q});
        p = q;
    }
    return t;
}

std::string lower(Span s) {
    std::string r = s.str();
    for (char &c : r)
        if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
    return r;
}

// Unsigned decimal as Rust's `str::parse::<usize>()` reads it: an optional
// '+', then one or more digits, no overflow.  `why` gets the text of Rust's
// ParseIntError on failure.
bool parse_u64(Span s, uint64_t &v, const char **why) {
    const char *p = s.b;
    if (p < s.e && *p == '+') ++p;
    if (p == s.e) {
        *why = s.b == s.e ? "cannot parse integer from empty string" : "invalid digit found in string";
        return false;
    }
    uint64_t x = 0;
    for (; p < s.e; ++p) {
        if (*p < '0' || *p > '9') {
            *why = "invalid digit found in string";
            return false;
        }
        const uint64_t d = (uint64_t)(*p - '0');
        if (x > (UINT64_MAX - d) / 10) {
            *why = "number too large to fit in target type";
            return false;
        }
        x = x * 10 + d;
    }
    v = x;
    return true;
}

// Line cursor over the whole buffer.
struct Lines {
    const char *p, *end;
    uint64_t number = 0;  // number of the line last returned
    bool next(Span &out) {
        if (p >= end) return false;
        const char *q = (const char *)memchr(p, '\n', (size_t)(end - p));
        if (!q) q = end;
        out = {p, q};
        p = q < end ? q + 1 : end;
        ++number;
        return true;
    }
};

struct Header {
    bool symmetric = false;
};

int parse_header(Span line, uint64_t ln, Header &h) {
    const std::vector<Span> t = tokens(line);
    auto bad = [&](const std::string &reason) {
        return mtx_fail(MSBFS_MTX_INVALID_HEADER, ln, 0, 0, 0, 0,
                        "Invalid Matrix Market header at line " + std::to_string(ln) + ": " +
                            reason);
    };
    if (t.size() != 5)
        return bad("Expected 5 tokens in header line, found " + std::to_string(t.size()) + ": '" +
                   line.str() + "'");
    if (lower(t[0]) != "%%matrixmarket")
        return bad("Missing '%%MatrixMarket' banner: '" + t[0].str() + "'");
    if (lower(t[1]) != "matrix")
        return bad("Unsupported object '" + t[1].str() + "', only 'matrix' is supported");
    const std::string fmt = lower(t[2]);
    if (fmt == "array")
        return mtx_fail(MSBFS_MTX_UNSUPPORTED_FORMAT, ln, 0, 0, 0, 0,
                        "Unsupported Matrix Market format 'array': only sparse 'coordinate' "
                        "format is supported");
    if (fmt != "coordinate") return bad("Unknown format '" + fmt + "'");
    const std::string field = lower(t[3]);
    if (field != "pattern" && field != "real" && field != "integer" && field != "complex")
        return bad("Unknown field '" + field + "'");
    const std::string sym = lower(t[4]);
    if (sym == "general")
        h.symmetric = false;
    else if (sym == "symmetric" || sym == "skew-symmetric" || sym == "hermitian")
        h.symmetric = true;
    else
        return bad("Unknown symmetry '" + sym + "'");
    return MSBFS_OK;
}

// First error of a chunk (file order), reported after all chunks are parsed.
struct ChunkError {
    int kind = 0;
    uint64_t line = 0, row = 0, col = 0;
    std::string text;
};

void parse_chunk(const char *b, const char *e, uint64_t first_line, uint64_t rows, uint64_t cols,
                 bool symmetric, std::vector<uint64_t> &pairs, ChunkError &err) {
    Lines L{b, e, first_line - 1};
    Span line;
    while (L.next(line)) {
        const Span s = trim(line);
        if (s.b == s.e || *s.b == '%') continue;
        // first two tokens
        const char *p = s.b;
        const char *q = p;
        while (q < s.e && !is_space(*q)) ++q;
        const Span t0{p, q};
        p = q;
        while (p < s.e && is_space(*p)) ++p;
        q = p;
        while (q < s.e && !is_space(*q)) ++q;
        const Span t1{p, q};
        uint64_t row = 0, col = 0;
        const char *why = nullptr;
        if (t1.b == t1.e || !parse_u64(t0, row, &why) || !parse_u64(t1, col, &why)) {
            err = {MSBFS_MTX_INVALID_COORDINATE_LINE, L.number, 0, 0, s.str()};
            return;
        }
        if (row == 0 || col == 0) {
            err = {MSBFS_MTX_ZERO_INDEX, L.number, row, col, {}};
            return;
        }
        if (row > rows || col > cols) {
            err = {MSBFS_MTX_INDEX_OUT_OF_BOUNDS, L.number, row, col, {}};
            return;
        }
        const uint64_t u = row - 1, v = col - 1;
        if (u == v) continue;  // self loop
        pairs.push_back((u << 32) | v);
        if (symmetric) pairs.push_back((v << 32) | u);
    }
}

int parse_buffer(const char *text, uint64_t len, msbfs_host_graph **out) {
    g_mtx_err = msbfs_mtx_error{};
    Lines L{text, text + len};
    Span line;
    // 1. banner: the first non-empty line
    Header h;
    for (;;) {
        if (!L.next(line))
            return mtx_fail(MSBFS_MTX_INVALID_HEADER, L.number + 1, 0, 0, 0, 0,
                            "Invalid Matrix Market header at line " +
                                std::to_string(L.number + 1) +
                                ": Empty input: no MatrixMarket header found");
        const Span s = trim(line);
        if (s.b == s.e) continue;
        const int rc = parse_header(s, L.number, h);
        if (rc) return rc;
        break;
    }
    // 2. size line
    uint64_t rows = 0, cols = 0, nnz = 0;
    for (;;) {
        if (!L.next(line))
            return mtx_fail(MSBFS_MTX_INVALID_DIMENSIONS, L.number + 1, 0, 0, 0, 0,
                            "Invalid dimensions at line " + std::to_string(L.number + 1) +
                                ": Unexpected end of file while searching for dimension line");
        const Span s = trim(line);
        if (s.b == s.e || *s.b == '%') continue;
        const std::vector<Span> t = tokens(s);
        const std::string at = "Invalid dimensions at line " + std::to_string(L.number) + ": ";
        if (t.size() != 3)
            return mtx_fail(MSBFS_MTX_INVALID_DIMENSIONS, L.number, 0, 0, 0, 0,
                            at + "Expected 3 integer tokens (rows cols nnz), got " +
                                std::to_string(t.size()) + ": '" + s.str() + "'");
        const char *why = nullptr;
        const char *names[3] = {"rows", "cols", "nnz"};
        uint64_t *dst[3] = {&rows, &cols, &nnz};
        for (int k = 0; k < 3; ++k)
            if (!parse_u64(t[k], *dst[k], &why))
                return mtx_fail(MSBFS_MTX_INVALID_DIMENSIONS, L.number, 0, 0, 0, 0,
                                at + "Failed to parse " + names[k] + " count '" + t[k].str() +
                                    "': " + why);
        break;
    }
    if (h.symmetric && rows != cols)
        return mtx_fail(MSBFS_MTX_INVALID_DIMENSIONS, L.number, 0, 0, 0, 0,
                        "Invalid dimensions at line " + std::to_string(L.number) +
                            ": Symmetric matrix must be square, but got " +
                            std::to_string(rows) + " rows and " + std::to_string(cols) +
                            " cols");
    const uint64_t V = std::max(rows, cols);
    if (V > 0xFFFFFFFFull)
        return mtx_fail(MSBFS_MTX_INVALID_DIMENSIONS, L.number, 0, 0, 0, 0,
                        "Invalid dimensions at line " + std::to_string(L.number) +
                            ": more than 2^32 - 1 vertices are not supported");

    // 3. entries, in parallel chunks cut at line boundaries
    const char *body = L.p, *end = text + len;
    const uint64_t body_line = L.number + 1;
    const uint64_t blen = (uint64_t)(end - body);
    int T = std::max(1, std::min<int>(msbfs::host_threads(), (int)(blen >> 20) + 1));
    std::vector<const char *> cut(T + 1);
    cut[0] = body;
    cut[T] = end;
    for (int i = 1; i < T; ++i) {
        const char *c = body + blen * i / T;
        if (c < cut[i - 1]) c = cut[i - 1];
        const char *nl = c < end ? (const char *)memchr(c, '\n', (size_t)(end - c)) : nullptr;
        cut[i] = nl ? nl + 1 : end;
    }
    std::vector<uint64_t> nlines(T, 0);
    std::vector<std::vector<uint64_t>> part(T);
    std::vector<ChunkError> errs(T);
    {
        std::vector<std::thread> th;
        for (int i = 0; i < T; ++i)
            th.emplace_back([&, i] {
                uint64_t n = 0;
                for (const char *p = cut[i]; p < cut[i + 1]; ++p) n += *p == '\n';
                nlines[i] = n;
            });
        for (auto &x : th) x.join();
    }
    {
        std::vector<uint64_t> first(T);
        uint64_t ln = body_line;
        for (int i = 0; i < T; ++i) {
            first[i] = ln;
            ln += nlines[i];
        }
        std::vector<std::thread> th;
        for (int i = 0; i < T; ++i)
            th.emplace_back([&, i] {
                part[i].reserve((size_t)((cut[i + 1] - cut[i]) / 8 * (h.symmetric ? 2 : 1)));
                parse_chunk(cut[i], cut[i + 1], first[i], rows, cols, h.symmetric, part[i], errs[i]);
            });
        for (auto &x : th) x.join();
    }
    for (int i = 0; i < T; ++i) {
        const ChunkError &e = errs[i];
        if (!e.kind) continue;
        const std::string at = std::to_string(e.line);
        switch (e.kind) {
        case MSBFS_MTX_INVALID_COORDINATE_LINE:
            return mtx_fail(e.kind, e.line, 0, 0, 0, 0,
                            "Invalid coordinate entry at line " + at + ": '" +
                                e.text.substr(0, 800) + "'");
        case MSBFS_MTX_ZERO_INDEX:
            return mtx_fail(e.kind, e.line, e.row, e.col, 0, 0,
                            "Matrix Market coordinates are 1-based, but found 0 at line " + at +
                                ": row=" + std::to_string(e.row) +
                                ", col=" + std::to_string(e.col));
        default:
            return mtx_fail(e.kind, e.line, e.row, e.col, rows, cols,
                            "Coordinate out of bounds at line " + at + ": entry (" +
                                std::to_string(e.row) + ", " + std::to_string(e.col) +
                                ") exceeds matrix dimensions (" + std::to_string(rows) + ", " +
                                std::to_string(cols) + ")");
        }
    }
    size_t total = 0;
    for (auto &v : part) total += v.size();
    std::vector<uint64_t> pairs;
    pairs.reserve(total);
    for (auto &v : part) {
        pairs.insert(pairs.end(), v.begin(), v.end());
        std::vector<uint64_t>().swap(v);
    }
    // 4. canonical CSR
    msbfs::HostCsr c;
    msbfs::csr_from_pairs(V, pairs, c, msbfs::host_threads());
    *out = msbfs::wrap_host_graph(std::move(c));
    return MSBFS_OK;
}

}  // namespace

extern "C" int msbfs_mtx_parse(const char *text, uint64_t len, msbfs_host_graph **out) {
    msbfs::clear_error();
    if (!out || (len && !text)) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    return parse_buffer(text, len, out);
}

extern "C" int msbfs_mtx_read(const char *path, msbfs_host_graph **out) {
    msbfs::clear_error();
    if (!out || !path) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    g_mtx_err = msbfs_mtx_error{};
    FILE *f = fopen(path, "rb");
    if (!f)
        return mtx_fail(MSBFS_MTX_IO, 0, 0, 0, 0, 0,
                        std::string("I/O error during Matrix Market parsing: ") + path + ": " +
                            strerror(errno));
    std::string buf;
    if (fseek(f, 0, SEEK_END) == 0) {
        const long n = ftell(f);
        if (n > 0) buf.resize((size_t)n);
        fseek(f, 0, SEEK_SET);
    }
    const size_t got = buf.empty() ? 0 : fread(&buf[0], 1, buf.size(), f);
    const bool bad = ferror(f) != 0;
    fclose(f);
    if (bad || got != buf.size())
        return mtx_fail(MSBFS_MTX_IO, 0, 0, 0, 0, 0,
                        std::string("I/O error during Matrix Market parsing: short read of ") + path);
    return parse_buffer(buf.data(), buf.size(), out);
}

extern "C" int msbfs_mtx_last_error(msbfs_mtx_error *out) {
    if (!out) return fail(MSBFS_ERR_INVALID_ARG, "null argument");
    *out = g_mtx_err;
    return MSBFS_OK;
}
