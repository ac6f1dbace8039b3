// Command-line driver with the reference driver's arguments and output
// (example.f90): reads a MatrixMarket file, analyses, factorizes on the GPU,
// solves A x = b for b = A(1,1) in every entry, and writes the one-line
// record the reference's experiment scripts read (example.f90:47-60).
//
// usage: mfdag_example matrix.mtx max_zero nb record_file
//                      [--natural | --perm FILE] [--fp32] [--analyze-only]
//   The matrix is ordered by the library's graph nested dissection
//   (mfd_order_nd) where the reference's driver has METIS order it
//   (options(4) = use_metis = 1, example.f90:33), unless
//   --natural       the matrix is analysed in its own order (use_metis = 0)
//   --perm FILE     the ordering is given (new -> old, 1-based, n integers)
// and
//   --fp32          FP32 fronts + FP64 iterative refinement
//   --analyze-only  analysis and its getters only (no GPU needed); the
//                   record's factorize / solve times are 0 and its error NaN
//
// The record: max_zero (i8), nb (i8), host threads (i8; the library runs on
// one host thread), nonzero (i12), cost (i20), array_size(1..3) (3 x i20),
// time_info(1..7) (7 x f10.3, seconds) and the error (e10.3)
// ||b - A x||_2 / ||x||_2 (example.f90:147-180), in Fortran's edit forms.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mfdag.h"

namespace {

// Fortran Iw: right-justified, w asterisks when it does not fit.
std::string fmt_i(long long v, int w)
{
    std::string s = std::to_string(v);
    if ((int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}

std::string fmt_special(double v, int w)
{
    std::string s = std::isnan(v) ? "NaN" : (v < 0 ? "-Infinity" : "Infinity");
    if ((int)s.size() > w) s = std::isnan(v) ? "NaN" : (v < 0 ? "-Inf" : "Inf");
    if ((int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}

// Fortran Fw.d.
std::string fmt_f(double v, int w, int d)
{
    if (!std::isfinite(v)) return fmt_special(v, w);
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*f", d, v);
    std::string s = buf;
    if ((int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}

// Fortran Ew.d: [-]0.<d digits>E<sign><2 digits>, or [-]0.<d digits><sign><3
// digits> for exponents beyond 99.
std::string fmt_e(double v, int w, int d)
{
    if (!std::isfinite(v)) return fmt_special(v, w);
    std::string digits(d, '0');
    int exp10 = 0;
    if (v != 0.0) {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.*e", d - 1, std::fabs(v));  // d.ddde+XX
        digits = std::string(1, buf[0]) + std::string(buf + 2, buf + 1 + d);
        exp10 = std::atoi(std::strchr(buf, 'e') + 1) + 1;
    }
    std::string s = (v < 0 ? "-0." : "0.") + digits;
    const char sign = exp10 < 0 ? '-' : '+';
    char e[16];
    if (std::abs(exp10) <= 99)
        std::snprintf(e, sizeof e, "E%c%02d", sign, std::abs(exp10));
    else
        std::snprintf(e, sizeof e, "%c%03d", sign, std::abs(exp10));
    s += e;
    if ((int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}

int fail(const char* what, int rc)
{
    std::fprintf(stderr, "%s: %s (code %d)\n", what, mfd_error_string(rc), rc);
    return 1;
}

// ||rh - A x||_2 / ||x||_2 with A given by its lower CCS (example.f90:147-180).
double calc_error(int n, const std::vector<int>& col, const std::vector<int>& row, const std::vector<double>& val,
                  double rh, const std::vector<double>& x)
{
    std::vector<double> t(n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int p = col[j] - 1; p < col[j + 1] - 1; ++p) {
            const int i = row[p] - 1;
            t[i] += val[p] * x[j];
            if (i != j) t[j] += val[p] * x[i];
        }
    double r2 = 0, x2 = 0;
    for (int i = 0; i < n; ++i) {
        r2 += (rh - t[i]) * (rh - t[i]);
        x2 += x[i] * x[i];
    }
    return std::sqrt(r2) / std::sqrt(x2);
}

bool read_perm(const char* path, int n, std::vector<int>& perm)
{
    FILE* f = std::fopen(path, "r");
    if (!f) return false;
    perm.resize(n);
    int got = 0;
    while (got < n && std::fscanf(f, "%d", &perm[got]) == 1) ++got;
    std::fclose(f);
    return got == n;
}

}  // namespace

int main(int argc, char** argv)
{
    std::vector<const char*> pos;
    const char* perm_file = nullptr;
    bool fp32 = false, analyze_only = false, natural = false;
    for (int a = 1; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--perm") && a + 1 < argc)
            perm_file = argv[++a];
        else if (!std::strcmp(argv[a], "--natural"))
            natural = true;
        else if (!std::strcmp(argv[a], "--fp32"))
            fp32 = true;
        else if (!std::strcmp(argv[a], "--analyze-only"))
            analyze_only = true;
        else
            pos.push_back(argv[a]);
    }
    if (pos.size() != 4 || (natural && perm_file)) {
        std::printf(" command line argument error\n");
        return 1;
    }
    const char* file_name = pos[0];
    const int max_zero = std::atoi(pos[1]);
    const int nb = std::atoi(pos[2]);
    const char* data_name = pos[3];

    int n = 0;
    int64_t nnz = 0;
    int rc = mfd_mm_size(file_name, &n, &nnz);
    if (rc != MFD_OK) return fail(file_name, rc);
    std::vector<int> ccs_col(n + 1), ccs_row(nnz);
    std::vector<double> ccs_val(nnz);
    rc = mfd_mm_read(file_name, n, nnz, ccs_col.data(), ccs_row.data(), ccs_val.data());
    if (rc != MFD_OK) return fail(file_name, rc);
    std::vector<int> perm;
    if (perm_file && !read_perm(perm_file, n, perm)) {
        std::fprintf(stderr, "%s: expected %d integers\n", perm_file, n);
        return 1;
    }

    const double rh = ccs_val[0];
    std::vector<double> b(n, rh);
    std::printf(" I\n");
    mfd_handle* h = nullptr;
    rc = mfd_create(&h, nb, MFD_FLAG_GRAPH | (natural ? 0 : MFD_FLAG_ORDER_ND));
    if (rc != MFD_OK && !(analyze_only && rc == MFD_ERR_NODEVICE)) {
        mfd_destroy(h);
        return fail("mfd_create", rc);
    }
    std::printf(" A\n");
    rc = mfd_analyze(h, n, ccs_col.data(), ccs_row.data(), max_zero, perm.empty() ? nullptr : perm.data());
    if (rc != MFD_OK) {
        mfd_destroy(h);
        return fail("mfd_analyze", rc);
    }
    double error = NAN;
    if (!analyze_only) {
        std::printf(" F\n");
        rc = mfd_factorize(h, ccs_val.data(), fp32 ? MFD_FP32_IR : MFD_FP64);
        if (rc != MFD_OK) {
            mfd_destroy(h);
            return fail("mfd_factorize", rc);
        }
        std::printf(" S\n");
        rc = mfd_solve(h, b.data(), 1, n);
        if (rc != MFD_OK) {
            mfd_destroy(h);
            return fail("mfd_solve", rc);
        }
        error = calc_error(n, ccs_col, ccs_row, ccs_val, rh, b);
    }

    std::string rec = fmt_i(max_zero, 8) + fmt_i(nb, 8) + fmt_i(1, 8) + fmt_i(mfd_get_nonzero(h), 12) +
                      fmt_i(mfd_get_cost(h), 20);
    for (int i = 1; i <= 3; ++i) rec += fmt_i(mfd_get_array_size(h, i), 20);
    for (int i = 1; i <= 7; ++i) rec += fmt_f(mfd_get_time_info(h, i), 10, 3);
    rec += fmt_e(error, 10, 3);
    mfd_destroy(h);

    FILE* f = std::fopen(data_name, "w");
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", data_name);
        return 1;
    }
    std::fprintf(f, "%s\n", rec.c_str());
    std::fclose(f);
    return 0;
}
