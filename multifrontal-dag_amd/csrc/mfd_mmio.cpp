// MatrixMarket input of the driver program (example.f90:63-145): a symmetric
// matrix stored as its lower triangle in coordinate form becomes the 1-based
// lower CCS, diagonal first in each column, that mfd_analyze takes.
//
// The reader follows the reference driver's reading rules: lines that start
// with '%' before the size line are skipped (example.f90:72-85), the size
// line is "n n nnz" (example.f90:87) and nnz "row col value" entries follow
// (example.f90:89-91).  The column order of the reference's two counting
// passes (example.f90:104-143) is kept: rows ascending within a column and,
// for an entry given twice, the later one first.  Unlike the reference, a
// malformed file is an error rather than undefined behaviour.
#include <cctype>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mfdag.h"

namespace {

struct Reader {
    std::string text;
    size_t pos = 0;

    bool load(const char* path)
    {
        FILE* f = std::fopen(path, "rb");
        if (!f) return false;
        char buf[1 << 16];
        size_t got;
        while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
        std::fclose(f);
        return true;
    }

    // The banner, if any, names a real (or integer) symmetric coordinate
    // matrix; then every '%' line before the size line is skipped.
    int header()
    {
        if (text.compare(0, 14, "%%MatrixMarket") == 0) {
            const size_t end = text.find('\n');
            std::string banner = text.substr(0, end);
            for (auto& ch : banner) ch = (char)std::tolower((unsigned char)ch);
            char obj[64] = {}, fmt[64] = {}, field[64] = {}, sym[64] = {};
            if (std::sscanf(banner.c_str(), "%%%%matrixmarket %63s %63s %63s %63s", obj, fmt, field, sym) != 4)
                return MFD_ERR_ARG;
            if (std::strcmp(obj, "matrix") != 0 || std::strcmp(fmt, "coordinate") != 0) return MFD_ERR_UNSUPPORTED;
            if (std::strcmp(field, "real") != 0 && std::strcmp(field, "integer") != 0 &&
                std::strcmp(field, "double") != 0)
                return MFD_ERR_UNSUPPORTED;
            if (std::strcmp(sym, "symmetric") != 0) return MFD_ERR_UNSUPPORTED;
        }
        while (pos < text.size() && text[pos] == '%') {
            const size_t end = text.find('\n', pos);
            pos = end == std::string::npos ? text.size() : end + 1;
        }
        return MFD_OK;
    }

    // Next whitespace-separated token; empty at the end of the text.
    std::string token()
    {
        while (pos < text.size() && std::isspace((unsigned char)text[pos])) ++pos;
        const size_t start = pos;
        while (pos < text.size() && !std::isspace((unsigned char)text[pos])) ++pos;
        return text.substr(start, pos - start);
    }

    bool integer(int64_t* out)
    {
        const std::string t = token();
        if (t.empty()) return false;
        char* end = nullptr;
        const long long v = std::strtoll(t.c_str(), &end, 10);
        if (*end != '\0') return false;
        *out = v;
        return true;
    }

    // A Fortran-written value may carry a D exponent (1.5D+00).
    bool real(double* out)
    {
        std::string t = token();
        if (t.empty()) return false;
        for (auto& ch : t)
            if (ch == 'd' || ch == 'D') ch = 'e';
        char* end = nullptr;
        *out = std::strtod(t.c_str(), &end);
        return *end == '\0';
    }

    int size(int64_t* n, int64_t* nnz)
    {
        int rc = header();
        if (rc != MFD_OK) return rc;
        int64_t m = 0;
        if (!integer(n) || !integer(&m) || !integer(nnz)) return MFD_ERR_ARG;
        if (*n < 1 || *n != m || *n >= INT_MAX || *nnz < 1 || *nnz >= INT_MAX) return MFD_ERR_ARG;
        return MFD_OK;
    }
};

}  // namespace

extern "C" int mfd_mm_size(const char* path, int* n, int64_t* nnz)
{
    if (!path || !n || !nnz) return MFD_ERR_ARG;
    Reader r;
    if (!r.load(path)) return MFD_ERR_ARG;
    int64_t nn = 0, nz = 0;
    const int rc = r.size(&nn, &nz);
    if (rc != MFD_OK) return rc;
    *n = (int)nn;
    *nnz = nz;
    return MFD_OK;
}

extern "C" int mfd_mm_read(const char* path, int n, int64_t nnz, int* ccs_col, int* ccs_row, double* ccs_val)
{
    if (!path || !ccs_col || !ccs_row || !ccs_val) return MFD_ERR_ARG;
    Reader r;
    if (!r.load(path)) return MFD_ERR_ARG;
    int64_t fn = 0, fnz = 0;
    int rc = r.size(&fn, &fnz);
    if (rc != MFD_OK) return rc;
    if (fn != n || fnz != nnz) return MFD_ERR_ARG;
    std::vector<int> row(nnz), col(nnz);
    std::vector<double> val(nnz);
    for (int64_t e = 0; e < nnz; ++e) {
        int64_t i = 0, j = 0;
        if (!r.integer(&i) || !r.integer(&j) || !r.real(&val[e])) return MFD_ERR_ARG;
        if (j < 1 || i < j || i > n) return MFD_ERR_ARG;  // lower triangle, 1-based
        row[e] = (int)i;
        col[e] = (int)j;
    }
    // Pass 1: entries grouped by row, within a row the later entry first.
    std::vector<int64_t> start(n + 2, 0);
    for (int64_t e = 0; e < nnz; ++e) ++start[row[e] + 1];
    for (int i = 1; i <= n + 1; ++i) start[i] += start[i - 1];
    std::vector<int64_t> by_row(nnz);
    for (int64_t e = nnz - 1; e >= 0; --e) by_row[start[row[e]]++] = e;
    // Pass 2: stable by column, so rows ascend within each column.
    std::vector<int64_t> next(n + 1, 0);
    for (int64_t e = 0; e < nnz; ++e) ++next[col[e]];
    ccs_col[0] = 1;
    for (int j = 1; j <= n; ++j) ccs_col[j] = ccs_col[j - 1] + (int)next[j];
    for (int j = 1; j <= n; ++j) next[j] = ccs_col[j - 1] - 1;
    for (int64_t k = 0; k < nnz; ++k) {
        const int64_t e = by_row[k];
        const int64_t p = next[col[e]]++;
        ccs_row[p] = row[e];
        ccs_val[p] = val[e];
    }
    return MFD_OK;
}
