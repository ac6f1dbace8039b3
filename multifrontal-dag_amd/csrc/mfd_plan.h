// Task descriptors shared by the host scheduler (mfd_runtime.hip) and the
// device kernels (mfd_kernels.hip).  Every launch of the numeric schedule
// processes a contiguous range of one descriptor array that is built once per
// analysis and kept in device memory.
#pragma once
#include <cstdint>

namespace mfd {

constexpr int kTile = 64;        // edge of every front tile a workgroup owns
constexpr int kPanel = 64;       // pivot columns per panel step of a large front
constexpr int kSmallFront = 64;  // fronts up to this order are factored whole in LDS

// One front (supernode).  The front is an n x n column-major block (ld = n)
// in the front pool; its lower part holds, after factorization, L in the
// first c columns and the update matrix (Schur complement) in the trailing
// w x w block.
struct NodeDev {
    int64_t front;   // element offset of the front in the pool
    int64_t lidx;    // offset of this node's local_index (w entries) in the lidx array
    int64_t z;       // per-rhs offset of the solve slice ((n + c) entries per rhs) in the z pool
    int64_t inv;     // element offset (in the pool) of the inverses of the 64 x 64 diagonal blocks
    int n, c, w;
    int parent;      // 0-based parent node, -1 for a root
    int first;       // 0-based first global (postordered) column
    int pad[3];
};

// Extend-add of one 64 x 64 tile (ti >= tj) of child's update matrix.
struct ExtTask {
    int child, ti, tj, pad;
};

// Panel step of a large front: diagonal block [k0, k0+kb) and, for the TRSM
// kernel, the 64-row chunk starting at row0 (row0 >= k0 + kb).
struct PanelTask {
    int node, k0, kb, row0;
};

// One block step of the blocked supernodal solves: the pivot block
// [k0, k0+kb) of a front and, for the row-chunk kernels, the 256 rows from
// row0.  slot/nchunk address the backward sweep's partial-sum scratch
// (slot s holds 64 x nrhs doubles).
struct SolveTask {
    int node, k0, kb, row0;
    int slot, nchunk, pad0, pad1;
};

// C (m x n tile, ldc) -= A (m x k, lda) * B (n x k, ldb)^T, offsets into the
// front pool; tri != 0 keeps only the lower triangle (row >= col) of C.
// k == 0 marks a copy task: the lower m x n block at a (lda) is copied to c.
struct GemmTask {
    int64_t c, a, b;
    int ldc, lda, ldb;
    int m, n, k;
    int tri, pad;
};

}  // namespace mfd
