// Reference tile storage of the factor.  The offsets below are the
// reference's storage format (they must agree entry for entry with it):
//   block sizes            factors/block_size_calculator_m.f90:12-21
//   work start / border    node_data_m.f90:46-50, 102-115, 156-184
//   supernode array        factors/supernode_matrix_extractor_m.f90:14-54
//   work array             factors/work_matrix_extractor_m.f90:14-74
//   border array           factors/border_matrix_extractor_m.f90:13-54
// Tile (i, j) is row-major with row stride = its column count
// (coefficient_setter_m.f90:45-50).  Indices i, j are 1-based block numbers.
#include "mfd_layout.h"

#include <algorithm>

namespace mfd {

namespace {
int64_t tri(int64_t e) { return e < 1 ? 0 : e * (e + 1) / 2; }
int64_t tri_range(int64_t s, int64_t e) { return e < s ? 0 : tri(e) - tri(s - 1); }
int ceil_div(int a, int p) { return (a - 1) / p + 1; }  // integer_function_m.f90:45-48 (truncating)
int block_size(int idx, int nb, int n) { return (int64_t)idx * nb > n ? n % nb : nb; }
}  // namespace

RefLayout::RefLayout(const std::vector<int>& c, const std::vector<int>& w, int nb) : nb_(nb), c_(c), w_(w)
{
    int m = (int)c.size();
    ss_.resize(m);
    wsz_.resize(m);
    sn_off_.assign(m + 1, 0);
    wk_off_.assign(m + 1, 0);
    bd_off_.assign(m + 1, 0);
    for (int s = 0; s < m; ++s) {
        ss_[s] = c_[s] % nb_;
        wsz_[s] = std::min(nb_ - ss_[s], w_[s]);
    }
    for (int s = 0; s < m; ++s) {
        sn_off_[s + 1] = sn_off_[s] + sn_node(s);
        wk_off_[s + 1] = wk_off_[s] + wk_node(s);
        bd_off_[s + 1] = bd_off_[s] + bd_node(s);
    }
}

int RefLayout::blocks(int s) const { return ceil_div(n_of(s), nb_); }
int RefLayout::work_start(int s) const { return w_[s] == 0 ? ceil_div(c_[s], nb_) + 1 : c_[s] / nb_ + 1; }
int RefLayout::matrix_bs(int idx, int s) const { return block_size(idx, nb_, n_of(s)); }
int RefLayout::supernode_bs(int idx, int s) const { return block_size(idx, nb_, c_[s]); }
int RefLayout::work_bs(int idx, int s) const { return idx == work_start(s) ? wsz_[s] : matrix_bs(idx, s); }

int64_t RefLayout::sn_node(int s) const
{
    return tri_range(w_[s] + 1, n_of(s)) + tri(nb_ - 1) * (c_[s] / nb_) + tri(ss_[s] - 1);
}
int64_t RefLayout::wk_node(int s) const
{
    int B = blocks(s), ws = work_start(s);
    int64_t wn = (B - ws + 1) - 2;
    return (int64_t)w_[s] * wsz_[s] + tri(wn) * nb_ * nb_ + (int64_t)(w_[s] - wsz_[s]) * matrix_bs(B, s);
}
int64_t RefLayout::bd_node(int s) const
{
    if (!has_border(s)) return 0;
    return (int64_t)(n_of(s) - (c_[s] / nb_) * nb_) * matrix_bs(work_start(s), s);
}
int64_t RefLayout::sn_tile(int s, int i, int j) const
{
    int n = n_of(s);
    int64_t left = tri_range((int64_t)n - (int64_t)(j - 1) * nb_ + 1, n) + tri(nb_ - 1) * (j - 1);
    return left + (int64_t)(i - j) * nb_ * supernode_bs(j, s);
}
int64_t RefLayout::wk_tile(int s, int i, int j) const
{
    int ws = work_start(s);
    int64_t up, left;
    if (j == ws) {
        up = std::max<int64_t>((int64_t)nb_ * (i - 1) - c_[s], 0);
        left = 0;
    } else {
        up = (int64_t)(i - j) * nb_;
        left = tri_range((int64_t)n_of(s) - (int64_t)(j - 1) * nb_ + 1, w_[s]) + tri(nb_ - 1) * (j - ws - 1) +
               tri(wsz_[s] - 1);
    }
    return left + up * work_bs(j, s);
}

int64_t RefLayout::size(int which) const
{
    switch (which) {
    case 1: return sn_off_.back();
    case 2: return wk_off_.back();
    case 3: return bd_off_.back();
    default: return -1;
    }
}

void RefLayout::put_front(int s, const double* F, double* sn, double* wk) const
{
    const int n = n_of(s), B = blocks(s), ws = work_start(s), c = c_[s];
    const bool divisible = ss_[s] == 0;
    for (int j = 1; j <= B; ++j) {
        // which arrays own block column j (node_data_m.f90:156-178)
        bool in_sn = divisible ? j <= ws - 1 : j <= ws;
        bool in_wk = divisible ? j > ws - 1 : j >= ws;
        for (int i = j; i <= B; ++i) {
            int rows = matrix_bs(i, s);
            if (in_sn && sn) {
                int cols = supernode_bs(j, s);
                double* t = sn + sn_off_[s] + sn_tile(s, i, j);
                for (int ri = 0; ri < rows; ++ri) {
                    int r = (i - 1) * nb_ + ri;
                    for (int qj = 0; qj < cols; ++qj) {
                        int q = (j - 1) * nb_ + qj;
                        if (q <= r) t[(int64_t)ri * cols + qj] = F[(int64_t)q * n + r];
                    }
                }
            }
            if (in_wk && wk) {
                int wrows = work_bs(i, s), wcols = work_bs(j, s);
                int r0 = i == ws ? c : (i - 1) * nb_;
                int q0 = j == ws ? c : (j - 1) * nb_;
                double* t = wk + wk_off_[s] + wk_tile(s, i, j);
                for (int ri = 0; ri < wrows; ++ri)
                    for (int qj = 0; qj < wcols; ++qj) {
                        int r = r0 + ri, q = q0 + qj;
                        if (q <= r) t[(int64_t)ri * wcols + qj] = F[(int64_t)q * n + r];
                    }
            }
        }
    }
}

}  // namespace mfd
