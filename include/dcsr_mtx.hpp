// dcsr_mtx.hpp -- MatrixMarket reader for the Dynamic-CSR engine (host C++,
// header only, no dependencies beyond the standard library).
//
// Replaces the two readers of the reference:
//   * cusp::io::read_matrix_market_file (cusp/io/detail/matrix_market.inl:
//     banner :69-93, coordinate body :117-221, symmetric expansion :227-266,
//     final sort :269), which the reference links but its drivers do not use;
//   * ReadMatrixFile (include/load_matrix.h:54-116), which its drivers use:
//     it skips the banner, forces every value that is missing or 0 to 1.0
//     (:92-93) and does not expand symmetric storage.
// This reader keeps the file's values (pattern entries are 1, as the
// MatrixMarket format defines them), expands symmetric storage to both
// triangles (off-diagonal entries mirrored, diagonal once) and
// skew-symmetric storage with the mirrored value negated, and returns
// 0-based COO triples sorted by (row, column), stable among equal
// coordinates.  Array (dense) storage, complex and hermitian files are
// rejected, as are indices outside the stated shape and files whose entry
// count does not match the size line.
//
//   dcsr::MtxCoo m;
//   std::string err;
//   if (!dcsr::read_mtx("cnr-2000.mtx", m, &err)) { ... }
//   // m.rows, m.cols, m.row[k], m.col[k], m.val[k]  (k < m.row.size())
#pragma once
#include <algorithm>
#include <cctype>
#include <cstdint>
#include <fstream>
#include <istream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

namespace dcsr {

struct MtxCoo {
    int64_t rows = 0, cols = 0;
    std::vector<int32_t> row, col;
    std::vector<float> val;
    std::string type, symmetry;  // banner fields as read (lower case)
};

namespace mtx_detail {

inline std::string lower(std::string s)
{
    for (char &ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

inline bool fail(std::string *err, const std::string &msg)
{
    if (err) *err = msg;
    return false;
}

}  // namespace mtx_detail

// Reads a coordinate MatrixMarket stream.  Returns false (and a message in
// *err when err is not null) on any format error; `out` is then unspecified.
inline bool read_mtx(std::istream &in, MtxCoo &out, std::string *err = nullptr)
{
    using namespace mtx_detail;
    std::string line;
    if (!std::getline(in, line)) return fail(err, "empty MatrixMarket stream");
    std::istringstream banner(line);
    std::string tag, object, storage, type, symmetry;
    banner >> tag >> object >> storage >> type >> symmetry;
    if (tag != "%%MatrixMarket" || lower(object) != "matrix") return fail(err, "invalid MatrixMarket banner");
    storage = lower(storage);
    type = lower(type);
    symmetry = lower(symmetry);
    if (storage != "coordinate") return fail(err, "unsupported MatrixMarket storage [" + storage + "]");
    if (type != "real" && type != "integer" && type != "pattern")
        return fail(err, "unsupported MatrixMarket data type [" + type + "]");
    if (symmetry != "general" && symmetry != "symmetric" && symmetry != "skew-symmetric")
        return fail(err, "unsupported MatrixMarket symmetry [" + symmetry + "]");
    // comments and blank lines up to the size line
    do {
        if (!std::getline(in, line)) return fail(err, "missing MatrixMarket size line");
    } while (line.empty() || line[0] == '%' || line.find_first_not_of(" \t\r") == std::string::npos);
    int64_t R = -1, C = -1, N = -1;
    {
        std::istringstream ss(line);
        ss >> R >> C >> N;
        if (ss.fail() || R < 0 || C < 0 || N < 0) return fail(err, "invalid MatrixMarket size line");
        if (R >= (int64_t(1) << 31) || C >= (int64_t(1) << 31)) return fail(err, "matrix shape exceeds int32 indices");
    }
    const bool pattern = type == "pattern";
    std::vector<int32_t> r, c;
    std::vector<float> v;
    r.reserve((size_t)N);
    c.reserve((size_t)N);
    v.reserve((size_t)N);
    int64_t got = 0;
    while (got < N && std::getline(in, line)) {
        if (line.empty() || line[0] == '%' || line.find_first_not_of(" \t\r") == std::string::npos) continue;
        std::istringstream ss(line);
        int64_t i = 0, j = 0;
        double x = 1.0;
        ss >> i >> j;
        if (!pattern) ss >> x;
        if (ss.fail()) return fail(err, "malformed MatrixMarket entry line " + std::to_string(got + 1));
        if (i < 1 || i > R) return fail(err, "row index out of range at entry " + std::to_string(got + 1));
        if (j < 1 || j > C) return fail(err, "column index out of range at entry " + std::to_string(got + 1));
        r.push_back((int32_t)(i - 1));
        c.push_back((int32_t)(j - 1));
        v.push_back((float)x);
        ++got;
    }
    if (got != N) return fail(err, "unexpected end of MatrixMarket entries");
    if (symmetry != "general") {
        const float sign = symmetry == "skew-symmetric" ? -1.0f : 1.0f;
        const size_t n0 = r.size();
        for (size_t k = 0; k < n0; ++k) {
            if (r[k] == c[k]) continue;
            r.push_back(c[k]);
            c.push_back(r[k]);
            v.push_back(sign * v[k]);
        }
    }
    // stable sort by (row, column): order of the expanded list among equal
    // coordinates is kept
    std::vector<size_t> ord(r.size());
    std::iota(ord.begin(), ord.end(), size_t(0));
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
        return r[a] != r[b] ? r[a] < r[b] : c[a] < c[b];
    });
    out.rows = R;
    out.cols = C;
    out.type = type;
    out.symmetry = symmetry;
    out.row.resize(ord.size());
    out.col.resize(ord.size());
    out.val.resize(ord.size());
    for (size_t k = 0; k < ord.size(); ++k) {
        out.row[k] = r[ord[k]];
        out.col[k] = c[ord[k]];
        out.val[k] = v[ord[k]];
    }
    return true;
}

inline bool read_mtx(const std::string &path, MtxCoo &out, std::string *err = nullptr)
{
    std::ifstream f(path);
    if (!f) return mtx_detail::fail(err, "cannot open " + path);
    return read_mtx(f, out, err);
}

}  // namespace dcsr
