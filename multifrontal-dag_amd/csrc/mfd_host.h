// Host (CPU) numeric factorization of whole subtrees of the assembly tree:
// the host side of the CPU+GPU hybrid.  The subtrees come from a subtree
// partition such as the reference's tree_pruning (tree_pruning_m.f90:13-73,
// mfd_tree_pruning here); host threads take fronts as soon as their
// children are done, each front as seq_factorize treats it
// (seq_factorize_m.f90:13-98: assembly, extend-add of the children's update
// blocks, partial Cholesky of the pivot columns with the Schur update of the
// trailing block), into a staging array laid out like the device front
// pool, so that each run of consecutive host fronts reaches the device in
// one copy.
#pragma once
#include <cstdint>
#include <vector>

#include "mfd_analysis.h"
#include "mfd_plan.h"

namespace mfd {

struct HostFront {
    int node;         // 0-based node id
    int n, c, w;      // front order, pivot columns, update rows
    int first;        // 0-based first global column
    int parent_pos;   // position of the parent in HostPlan::fronts; -1 for a subtree root
    int64_t front;    // offset of the n x n column-major front in the staging array
    int64_t inv;      // offset of its 64 x 64 diagonal-block inverses (L^{-T}, 4096 per block)
    int64_t lidx;     // offset of its local_index entries (w of them, 1-based parent positions)
    int64_t e0, e1;   // its assembly entries [e0, e1) of dst / src
};

// stage[stage .. stage + count) holds pool[pool .. pool + count)
struct HostRun {
    int64_t stage, pool, count;
};

struct HostPlan {
    std::vector<HostFront> fronts;  // subtree by subtree, each in postorder
    std::vector<int> sub_ptr;       // subtree k: fronts [sub_ptr[k], sub_ptr[k + 1])
    std::vector<int> sub_order;     // subtrees by decreasing flop count
    std::vector<int> child_ptr, child_pos;  // front f's children: child_pos[child_ptr[f] .. child_ptr[f + 1])
    std::vector<int64_t> dst, src;  // assembly: stage[dst[e]] = a[src[e]] (0-based caller entries)
    std::vector<HostRun> runs;
    std::vector<char> is_host;      // per node
    std::vector<int> roots;         // the subtree roots (0-based)
    int64_t stage_elems = 0;
    int64_t entries = 0;            // assembly entries of the host fronts
};

// The fronts of the subtrees rooted at roots (0-based, no root inside
// another's subtree), with pool positions from nodes (front and inv
// offsets as the device schedule lays them out).  Returns 0 or a negative
// error code (MFD_ERR_ARG for a bad or nested root).
int build_host_plan(const Structure& S, const std::vector<NodeDev>& nodes, const std::vector<int>& roots,
                    HostPlan& p);

// Assembles and factors every host subtree into stage (stage_elems
// doubles) from the caller's values a (CCS order), with up to `threads`
// threads.  Returns 0, or the 1-based global column of the first pivot that
// is not positive (the smallest over the subtrees).
int host_factor(const HostPlan& p, const Structure& S, const double* a, double* stage, int threads);

}  // namespace mfd
