// The reference's tile storage of the factor (supernode / work / border
// arrays), used to export factors in the reference layout and to report its
// array sizes.  The GPU never computes in this layout.
#pragma once
#include <cstdint>
#include <vector>

namespace mfd {

class RefLayout {
public:
    RefLayout(const std::vector<int>& c, const std::vector<int>& w, int nb);
    int64_t size(int which) const;  // 1 supernode, 2 work, 3 border (hppllt.f90:201-211)
    // Copy node s's dense front (column-major n x n, 0-based s) into the
    // supernode and work arrays (either may be null).
    void put_front(int s, const double* front, double* supernode, double* work) const;

private:
    int nb_;
    std::vector<int> c_, w_, ss_, wsz_;
    std::vector<int64_t> sn_off_, wk_off_, bd_off_;
    int n_of(int s) const { return c_[s] + w_[s]; }
    int blocks(int s) const;
    int work_start(int s) const;
    int matrix_bs(int idx, int s) const;
    int supernode_bs(int idx, int s) const;
    int work_bs(int idx, int s) const;
    bool has_border(int s) const { return ss_[s] != 0 && w_[s] != 0; }
    int64_t sn_node(int s) const;
    int64_t wk_node(int s) const;
    int64_t bd_node(int s) const;
    int64_t sn_tile(int s, int i, int j) const;
    int64_t wk_tile(int s, int i, int j) const;
};

}  // namespace mfd
