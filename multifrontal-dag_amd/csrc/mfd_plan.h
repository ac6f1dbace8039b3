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
    int xz;          // 1: the backward solve keeps the pivot part of x in Z
                     // (fronts solved over superblocks), 0: in Y
    int pad[2];
};

// Extend-add of one 64 x 64 tile (ti >= tj) of child's update matrix.
struct ExtTask {
    int child, ti, tj, pad;
};

// Diagonal block [k0, k0+kb) of a large front, to be factored and
// inverted.  kprev = 0: all its updates are done.  kprev > 0: the block
// still needs the previous panel (the kprev columns to its left, block s)
// and, with kpp > 0, the kpp columns before those (panels no in-panel
// update has applied yet); the factorizing workgroup also forms the
// panel-s entries of its own rows (their TRSM with block s's inverse) and
// stores them.
// applied = 1: the kpp pending columns were already applied to the block
// itself (by the previous step's TRSM tile of its rows, GemmTask::syrk); the
// task still applies them to its rows of block s.
struct DiagTask {
    int node, k0, kb, kprev;
    int kpp, applied, pad1, pad2;
};

// One block step of the blocked supernodal solves: the pivot block
// [k0, k0+kb) of a front and, for the row-chunk tasks, the rows from row0
// (forward: 64 rows; backward: 256; row0 < 0: the block's own task).
// Forward: a block task's pad is the order of the previous pivot block;
// pad0 = 1 marks a front of one pivot block, whose task also applies it to
// the front's update rows.
// Backward: a chunk task writes columns 16 pad0 .. 16 pad0 + 15 of
// partial-sum slot `slot` (64 x nrhs doubles); a block task reads the nchunk
// slots from `slot` and pad is the order of the next pivot block (0 for the
// front's last block).
struct SolveTask {
    int node, k0, kb, row0;
    int slot, nchunk, pad0, pad;
};

// Superblocks: the solves of a front with many pivot columns step over
// 256-column superblocks u (columns B_u) with three 256 x 256 matrices per
// superblock (column-major, in a separate array at xoff): X, the inverse of
// the superblock's diagonal block (zero past its width); at xoff + 256^2
// W = X L[B_u, B_{u-1}] (u >= 1, forward); at xoff + 2 256^2
// V = L[B_{u+1}, B_u] X (B_{u+1} not empty, backward).
constexpr int kSuper = 256;
constexpr int64_t kSuperMats = 3 * (int64_t)kSuper * kSuper;  // entries per superblock
// Forward task of superblock step u, 16 rows: kind 1, rows row0 .. row0 +
// 15 of superblock u (width sb; pb the width of superblock u-1, 0 for
// u = 0); kind 0, front rows row0 .. row0 + 15 after superblock u, to which
// superblock u-1 (width pb) is applied.
// Backward task (kind 2), rows row0 .. row0 + 15 of superblock u (width
// sb; pb the width of superblock u+1, 0 if none): x_u = X^T (y_u - q) -
// V^T x_{u+1}, with q summed from the partial-sum slots slot + nchunk j +
// ch (64 columns of block j of the superblock, chunk ch < nchunk).
struct SuperTask {
    int node, u, row0, sb;
    int pb, kind;
    int slot, nchunk;
    int64_t xoff;
};
// One 64 x 64 block (i, j) of superblock u's matrices: j = i copies the
// stored 64 x 64 inverse (and clears the row's blocks above the diagonal,
// or the whole row past the superblock's width); j < i forms
// X_ij = -X_ii sum_{k=j}^{i-1} L_ik X_kj; 4 <= j < 8 forms block (i, j - 4)
// of W; j >= 8 block (i, j - 8) of V.
struct SuperInvTask {
    int node, u, i, j;
    int64_t xoff;
};

// mode 0: C (m x n tile, ldc) -= A (m x k, lda) * B (n x k, ldb)^T;
// mode 1: C = A B^T with B stored transposed (B[j][k] at b + j * ldb + k);
// A may be C itself: the panel TRSM X = P L^{-T}, with b the stored L^{-T}
// (column-major), whose transpose is L^{-1};
// mode 2: C = (C - A B^T) D^T with D = L^{-1} stored transposed at d
// (ld 64) and K = kd: the panel TRSM preceded by the previous panel's
// update of the same rows;
// mode 3: mode 0, and each updated entry also added into the parent front
// at p (order np) at the rows and columns local_index gives (lidx offsets
// li for the tile's rows, lj for its columns): the Schur update with the
// extend-add of its result.  Offsets are into the front pool; tri != 0
// keeps only the lower triangle (row >= col) of C.
// syrk > 0 (a TRSM tile, modes 1 and 2, whose rows are the next pivot block
// but one): afterwards the lower syrk x syrk part of the diagonal block at
// cd (ld ldc) loses R R^T, R the tile's rows over every column from A to the
// end of C (the pending panels and the panel just solved), so that the
// diagonal task of the next step finds them applied.
struct GemmTask {
    int64_t c, a, b, d, p;
    int ldc, lda, ldb;
    int m, n, k;
    int tri, mode;
    int kd, np, li, lj;
    int syrk, pad;
    int64_t cd;
};

// A run of K of one 128 x 128 tile: GEMM task `task`, K columns [k0, k1);
// slot < 0: the whole K, applied at once; otherwise its partial product goes
// to partial-product slot `slot`.
struct SegTask {
    int task, k0, k1, slot;
};
// A tile whose K was split: its partial products in slots slot0 .. slot0 +
// nslot - 1, in K order.
struct FixTask {
    int task, slot0, nslot, pad;
};

}  // namespace mfd
