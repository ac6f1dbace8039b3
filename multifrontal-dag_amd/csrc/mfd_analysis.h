// Host symbolic analysis of the supernodal multifrontal Cholesky.
//
// Produces the reference's integer structure bit for bit (supernode
// partition, row lists, parent tree, relative-index maps, permutations);
// see mfd_analysis.cpp for the reference routine each step follows.
#pragma once
#include <cstdint>
#include <vector>

namespace mfd {

struct Structure {
    int n = 0;                       // matrix order
    int num_node = 0;                // number of (relaxed) supernodes
    std::vector<int> perm;           // new -> old (1-based), length n
    std::vector<int64_t> ccs_perm;   // a_structure entry -> caller entry (1-based), length nnz
    std::vector<int64_t> a_colptr;   // postordered lower CCS column pointers (0-based), n + 1
    std::vector<int> snidx;          // supernodal_index: front-relative row (1-based) per entry
    std::vector<int> c, w, parent;   // per node: pivot columns, update rows, parent (1-based, 0 root)
    std::vector<int> first;          // first global column (1-based) per node, num_node + 1
    std::vector<int64_t> l_ptr;      // row list (l_structure) pointers, num_node + 1
    std::vector<int> l_rows;         // row lists (global postordered rows, ascending)
    std::vector<int64_t> li_ptr;     // local_index pointers (= l_ptr), num_node + 1
    std::vector<int> lidx;           // local_index: parent-front position (1-based) of each update row
    std::vector<std::vector<int>> children;  // tree_child, ascending
};

// hppllt.f90:24-106 with use_metis = 0 when perm == nullptr, otherwise the
// given ordering (new -> old, 1-based) replaces METIS's.  Input: 1-based
// lower CCS with the diagonal first in each column.  Returns 0 or a negative
// error code.
int analyze(int n, const int* ccs_col, const int* ccs_row, int max_zero, const int* perm, Structure& out);

// Several independent systems analysed one by one (each exactly as
// analyze would) and joined into one elimination forest: system k's
// columns, nodes and entries follow those of systems 0..k-1, so the
// caller's values and right-hand sides are the systems' concatenated.
// perm may be null, and so may each perm[k].
int analyze_batch(int nsys, const int* n, const int* const* ccs_col, const int* const* ccs_row, int max_zero,
                  const int* const* perm, Structure& out);

// Structure given directly (node sizes, parents, local index and the
// supernodal CCS of A), as the reference's driver tests build it.
int structure_from_parts(int num_node, const int* c, const int* w, const int* parent, const int* lidx,
                         const int64_t* a_colptr, const int* snidx, Structure& out);

int64_t cost(const Structure& s);      // cost_calculator_m.f90:9-20
int64_t nonzero(const Structure& s);   // cost_calculator_m.f90:22-33

}  // namespace mfd
