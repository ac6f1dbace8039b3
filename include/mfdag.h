/*
 * mfdag.h -- C-ABI of the MI355X numeric phase of a supernodal multifrontal
 * Cholesky solver for sparse SPD systems.
 *
 * This library replaces the numeric phase behind the reference's Fortran
 * module `hppllt` (to-mo-ki/multifrontal-DAG, src/main/hppllt.f90:8-9).  The
 * host keeps the ordering and symbolic analysis (bit-identical to the
 * reference's); front assembly, extend-add, the partial front factorizations
 * and the supernodal triangular solves run as HIP kernels on one gfx950 GPU.
 *
 * Conventions
 *   - Index arrays are 1-based int32 as produced by Fortran; sizes are int64.
 *   - Return 0 on success; > 0 is the 1-based global (postordered) column at
 *     which a pivot was not positive (the matrix is not SPD: the reference
 *     ignores this, DPOTRF info at factorize_subroutines_m.f90:18); < 0 is an
 *     argument, state or HIP error (MFD_ERR_*).  No call aborts or throws.
 *   - Host pointers are borrowed for the duration of the call only.  Device
 *     memory and the HIP stream are owned by the handle.  Calls are
 *     synchronous on return.  One host thread per handle.
 */
#ifndef MFDAG_H
#define MFDAG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mfd_handle mfd_handle;

/* return codes */
#define MFD_OK 0
#define MFD_ERR_ARG (-1)
#define MFD_ERR_STATE (-2)
#define MFD_ERR_HIP (-3)
#define MFD_ERR_NOMEM (-4)
#define MFD_ERR_REDUCIBLE (-5)
#define MFD_ERR_INTERNAL (-6)
#define MFD_ERR_NODEVICE (-7)
#define MFD_ERR_UNSUPPORTED (-8)
#define MFD_ERR_NUMERIC (-9)
#define MFD_ERR_SCRATCH (-10) /* mfd_solve_device_noalloc: the solve needs more than the fixed scratch, or would allocate */

/* precision forms for mfd_factorize */
#define MFD_FP64 0     /* plain form: FP64 fronts, FP64 solves */
#define MFD_FP32_IR 1  /* FP32 fronts on MFMA + FP64 iterative refinement */
/* FP32 fronts of the diagonally scaled matrix S A S (S: powers of two, so
   that every scaled diagonal lies in [0.5, 2)) whose in-panel and Schur
   updates multiply FP16- (resp. BF16-) rounded operands on the 16-bit MFMA
   with FP32 accumulation; diagonal blocks and panel TRSM in FP32; FP64
   iterative refinement against the unscaled FP64 matrix.  For
   well-conditioned matrices: a stalled refinement refactors with FP32 fronts,
   and then in FP64. */
#define MFD_FP16_IR 2
#define MFD_BF16_IR 3

/* mfd_create flags */
#define MFD_FLAG_GRAPH 1    /* replay each schedule from a captured hipGraph */
#define MFD_FLAG_PROFILE 2  /* time every kernel class with HIP events */
#define MFD_FLAG_ORDER_ND 4 /* use_metis = 1: mfd_analyze(_batch) orders with mfd_order_nd when no perm is given */

/* timer ids: time_list_m.f90:3-10 */
#define MFD_TIME_ORDERING 1
#define MFD_TIME_ANALYZE 2
#define MFD_TIME_FACTORIZE 3
#define MFD_TIME_SOLVE 4
#define MFD_TIME_FORWARD 5
#define MFD_TIME_BACKWARD 6
#define MFD_TIME_REAL_FACTORIZE 7

/* kernel classes reported by mfd_get_kernel_time */
#define MFD_K_ASSEMBLE 0
#define MFD_K_EXTEND_ADD 1  /* extend-add of small children (n <= 64) */
#define MFD_K_SMALL_FRONT 2
#define MFD_K_POTRF 3       /* per panel step: next diagonal block (its panel updates, Cholesky, inverse), panel TRSM tiles, small updates */
#define MFD_K_TRSM 4        /* a step's GEMM tiles (panel TRSM, small updates) launched apart from its diagonal tasks (levels with many fronts) */
#define MFD_K_UPDATE 5      /* in-panel trailing update (GEMM/SYRK, K = panel width) */
#define MFD_K_SCHUR 6       /* Schur-complement update (SYRK, K = front's pivot count) and its extend-add into the parent */
#define MFD_K_SOLVE_FWD 7
#define MFD_K_SOLVE_BWD 8
#define MFD_K_REFINE 9
#define MFD_K_FRONT 10      /* medium fronts: the whole partial factorization of one front per workgroup */
#define MFD_K_NUM 11

/* Handle; nb is the reference's tile size (options(1), options_m.f90:3),
   used for the reference-layout views (mfd_export_tiles, array sizes). */
int mfd_create(mfd_handle** h, int nb, int flags);
void mfd_destroy(mfd_handle* h);

/* hppllt_analyze (hppllt.f90:24-81, create_data_structure :83-106).
   ccs_col (n+1) / ccs_row: 1-based CCS of the lower triangle with the
   diagonal first in each column.  perm (nullable, new -> old, 1-based) takes
   the place of the METIS ordering (use_metis = 1); NULL analyses the matrix
   as given (use_metis = 0), or, on a handle created with MFD_FLAG_ORDER_ND,
   in the order mfd_order_nd gives (use_metis = 1 with the library's own
   ordering; timed as MFD_TIME_ORDERING).  Builds the GPU schedule and
   allocates all device storage.  Unlike the reference, ccs_row is copied,
   not retained. */
int mfd_analyze(mfd_handle* h, int n, const int* ccs_col, const int* ccs_row, int max_zero, const int* perm);

/* The batching form: nsys independent systems, system k given like
   mfd_analyze's arguments (n[k], ccs_col[k], ccs_row[k], perm[k]; perm and
   each perm[k] nullable), each analysed exactly as mfd_analyze would and
   joined into one elimination forest that the level schedules factor and
   solve together.  Afterwards the handle holds one block-diagonal system:
   mfd_factorize takes the systems' values concatenated (system 0's CCS
   entries first) and mfd_solve the systems' right-hand sides stacked (rows
   of system 0 first).  No counterpart in the reference, whose analysis takes
   one matrix with one elimination tree (a reducible matrix returns
   MFD_ERR_REDUCIBLE from mfd_analyze). */
int mfd_analyze_batch(mfd_handle* h, int nsys, const int* n, const int* const* ccs_col, const int* const* ccs_row,
                      int max_zero, const int* const* perm);

/* The structure given directly, as the reference's driver tests build it
   (seq_factorize_test.f90:23-35): per node supernode_size, work_size and
   parent (1-based, 0 for the root, postordered), local_index (sum of
   work_size entries, create_local_index_m.f90:10-45), and the supernodal CCS
   of A (a_colptr: n+1 0-based offsets over the global postordered columns;
   supernodal_index: front-relative 1-based rows, create_supernodal_index_m.f90:10-40). */
int mfd_set_structure(mfd_handle* h, int num_node, const int* supernode_size, const int* work_size,
                      const int* parent, const int* local_index, const int64_t* a_colptr,
                      const int* supernodal_index);

/* hppllt_factorize (hppllt.f90:123-150): a holds the values in the caller's
   CCS order (after mfd_set_structure: in supernodal CCS order).  A
   lower-precision factorization that meets a non-positive pivot is redone
   in the next form (FP16/BF16 -> FP32 -> FP64; mfd_stats.precision reads the
   form that holds the factor); a positive return means the FP64
   factorization failed too. */
int mfd_factorize(mfd_handle* h, const double* a, int precision);
/* Same, with the values already in device memory (HBM-resident inputs). */
int mfd_factorize_device(mfd_handle* h, const double* d_a, int precision);

/* hppllt_solve (hppllt.f90:152-173) for nrhs right-hand sides: b (ldb x nrhs,
   column-major, caller numbering) is overwritten with x.  For MFD_FP32_IR
   factors the solution is refined in FP64 to a normwise backward error of
   2^-52-ish (see mfd_get_stats). */
int mfd_solve(mfd_handle* h, double* b, int nrhs, int ldb);
/* b in the postordered supernodal numbering (between the two permutations
   of hppllt.f90:157,170). */
int mfd_solve_permuted(mfd_handle* h, double* b, int nrhs, int ldb);
/* Device-resident b (caller numbering), ldb >= n. */
int mfd_solve_device(mfd_handle* h, double* d_b, int nrhs, int ldb);

/* No counterpart in the reference.  For callers that cannot allocate device
   memory (serving loops that solve with one factor many times, harnesses
   that forbid allocation after set-up): mfd_solve_device on a factored
   handle, the same kernels and results, with every solve buffer (Z/Y slices,
   partial sums, permuted b and x, and for the refined forms residual,
   correction and norm slots) taken from one fixed 128 MiB device array of
   the library instead of the handle's allocations.  It calls no hipMalloc
   or hipFree and captures no graph (launches are eager).  Each GPU holds
   its own copy of the array from the library's load until the process
   ends.  Returns MFD_ERR_SCRATCH when a solve of this size (n, fronts, nrhs)
   needs more than 128 MiB, or when a refinement stall would need a
   factorization buffer not yet allocated.  Every call on one GPU shares the
   array and leaves its contents behind: calls that overlap on one GPU (from
   several threads or handles) race.  mfd_get_solve_errors does not see this
   call's solution (MFD_ERR_STATE until the next mfd_solve*). */
int mfd_solve_device_noalloc(mfd_handle* h, double* d_b, int nrhs, int ldb);

/* Errors of the last solve, per right-hand side (nrhs must be that solve's;
   any output may be NULL), evaluated at this call in FP64 on the GPU from
   the solve's b and x and the FP64 values of A (the solve itself is not
   slowed):
     eta[k]  = ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf), the
               normwise backward error;
     res2[k] = ||b - A x||_2 / ||x||_2, the reference driver's measure
               (calc_error, example.f90:147-180);
     ferr[k] = ||d||_inf / ||x||_inf with A d = b - A x solved by the
               current factor: an estimate of the forward error
               ||x - A^{-1} b||_inf / ||x||_inf (one more solve sweep).
   The reference computes only res2, in its driver, for one right-hand side.
   MFD_ERR_STATE before a solve or after a new factorization. */
int mfd_get_solve_errors(mfd_handle* h, int nrhs, double* eta, double* res2, double* ferr);

/* Factor in the reference's tile storage (App. layout of factors_m.f90):
   supernode array (hppllt_get_array_size(1) entries) and work array
   (hppllt_get_array_size(2)); either may be NULL.  The border array is
   scratch in the reference (assembled copies of the border column before its
   split, matrix_splitter_m.f90:5-46) and is not exported. */
int mfd_export_tiles(mfd_handle* h, double* supernode, double* work);
/* Dense front of node (1-based): column-major n x n, n = supernode_size +
   work_size; lower part holds L (first supernode_size columns) and the
   update matrix (trailing block).  After an MFD_FP16_IR / MFD_BF16_IR
   factorization these are the fronts of the scaled matrix S A S. */
int mfd_export_front(mfd_handle* h, int node, double* out);

/* Getters: hppllt_get_cost / get_nonzero / get_array_size / get_time_info
   (hppllt.f90:186-211).  Sizes are int64 (the reference's nonzero is int32). */
int64_t mfd_get_cost(mfd_handle* h);
int64_t mfd_get_nonzero(mfd_handle* h);
int64_t mfd_get_array_size(mfd_handle* h, int which);
double mfd_get_time_info(mfd_handle* h, int id);

/* Structure accessors (for parity checks and the Fortran module):
   sizes: n, num_node, nnz (entries of the lower CCS), sum of work sizes. */
int mfd_get_dims(mfd_handle* h, int* n, int* num_node, int64_t* nnz, int64_t* nlidx);
/* Any pointer may be NULL.  supernode_size/work_size/parent: num_node;
   perm: n (new -> old); local_index: nlidx; supernodal_index, ccs_perm: nnz;
   a_colptr: n + 1. */
int mfd_get_structure(mfd_handle* h, int* supernode_size, int* work_size, int* parent, int* perm,
                      int* local_index, int* supernodal_index, int64_t* ccs_perm, int64_t* a_colptr);

typedef struct mfd_stats {
    double analyze_ms, assemble_ms, factor_ms, solve_ms;  /* last calls, wall (HIP events) */
    double factor_gflops;        /* cost / factor time (cost_calculator_m.f90:9-20) */
    int precision;               /* form of the current factor (after any fallback) */
    int refine_iters;            /* last solve (< 0: the solve changed form; -1 - iterations before that) */
    /* last refined solve, for its final x (-1 after an FP64 solve, which
       computes no residual; mfd_get_solve_errors gives them for any form): */
    double backward_error;       /* max over rhs of ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) */
    double refine_residual;      /* max over rhs of ||b - A x||_2 / ||x||_2 (example.f90:147-180) */
    int64_t front_bytes;         /* device bytes of the front pool */
    int64_t device_bytes;        /* all device bytes held by the handle */
    int num_levels, num_launches_factor, num_launches_solve;
} mfd_stats;
int mfd_get_stats(mfd_handle* h, mfd_stats* st);

/* Per-kernel-class timing of the last factorize + solve (needs
   MFD_FLAG_PROFILE): total ms and number of launches of class k. */
int mfd_get_kernel_time(mfd_handle* h, int k, double* total_ms, int* launches);
/* Algorithmic work of kernel class k over one factorization (k < MFD_K_SOLVE_FWD)
   or one single-RHS solve (solve classes) in the current schedule: flops
   (2 per multiply-add of the entries the kernel must produce) and the HBM
   bytes the class must move at minimum (within each launch every operand
   read once, however many tiles read it, and each result written once) for
   an element size of elem_bytes. */
int mfd_get_kernel_work(mfd_handle* h, int k, int elem_bytes, double* flops, double* bytes);

/* Fill-reducing ordering in place of METIS_NodeND (hppllt.f90:47-51,
   ordering_m.f90:10-38): graph nested dissection of the pattern of the
   1-based lower CCS (as mfd_analyze takes it).  perm (n) receives new -> old,
   1-based, the orientation mfd_analyze's perm argument takes.  Host only. */
int mfd_order_nd(int n, const int* ccs_col, const int* ccs_row, int* perm);

/* Subtree partition of an assembly tree (tree_pruning, tree_pruning_m.f90:
   13-74; host only): subtree roots whose subtrees, dealt out largest first
   to the least loaded of `workers` >= p bins, leave the lightest bin more
   than tol (0 <= tol < 1) of the heaviest's work, work being the
   reference's cost formula (cost_calculator_m.f90:9-20) summed over a
   subtree.  Input: per node (postordered, parents after children) c, w and
   the 1-based parent (0 for a root), as mfd_get_structure returns them.
   Output: subtrees (num_node entries of room) receives the 1-based roots in
   the reference's list order, count their number, and bins (nullable) the
   bin of each (0-based; the number of bins is the largest + 1).  A forest
   starts from all of its roots. */
int mfd_tree_pruning(int num_node, const int* c, const int* w, const int* parent, double tol, int p, int* subtrees,
                     int* bins, int* count);

/* MatrixMarket input, as the reference's driver reads its matrices
   (example.f90:63-145; no counterpart in the hppllt module): a symmetric
   matrix stored as its lower triangle, "row col value" lines in coordinate
   form.  An optional banner must name a real (or integer) symmetric
   coordinate matrix; '%' lines before the size line "n n nnz" are skipped.
   mfd_mm_size returns n and nnz.  mfd_mm_read fills the 1-based lower CCS
   that mfd_analyze / mfd_factorize take: ccs_col (n + 1), ccs_row and
   ccs_val (nnz), rows ascending within each column (so the diagonal comes
   first), an entry given twice kept twice with the later one first -- the
   column order of example.f90:104-143.  Entries above the diagonal, indices
   out of range, n != m, a short file or unreadable numbers are MFD_ERR_ARG;
   a banner naming another kind of matrix is MFD_ERR_UNSUPPORTED. */
int mfd_mm_size(const char* path, int* n, int64_t* nnz);
int mfd_mm_read(const char* path, int n, int64_t nnz, int* ccs_col, int* ccs_row, double* ccs_val);

/* Schedule tuning (no counterpart in the reference, whose schedule is
   StarPU's): the in-panel update of the large fronts runs every `period`
   64-column steps with K = 64 period (default 2, 1..8; 0 = per level: 1
   where the level's longest front has fewer than 48 steps, else 2), the
   choice the tuner (mfdag/tune.py) makes.  On an analysed handle the schedule is
   rebuilt at once and the factorization must be redone (mfd_solve returns
   MFD_ERR_STATE until then); otherwise it applies to the next analysis. */
int mfd_set_update_period(mfd_handle* h, int period);
/* Level rule of the schedule (no counterpart in the reference, which walks
   the fronts in postorder, seq_factorize_m.f90:13-29, or lets StarPU order
   them, starpu_factorize_m.f90:13-31): the fronts of one level share every
   launch.  1 (default): each front one level below its parent, roots on the
   top level, so that siblings' diagonal chains share launches; 0: each front
   one level above its highest child, leaves on level 0.  Rebuilds the
   schedule like mfd_set_update_period (the factorization is lost). */
int mfd_set_level_rule(mfd_handle* h, int rule);
int mfd_get_level_rule(mfd_handle* h);
int mfd_get_update_period(mfd_handle* h);

/* Medium fronts (no counterpart in the reference, whose StarPU tasks are
   per tile): a front of order > 64 is factored either in 64-column steps
   whose tiles spread over the whole GPU (the large-front schedule) or, as a
   medium front, whole by one workgroup (panel, TRSM and update tiles in
   turn; many such fronts of one level run side by side).
   mfd_set_front_split: fronts with at most cmax pivot columns are medium
   (0, the default: none).  mfd_set_front_kinds: one choice per node
   (num_node entries, 1-based node order: 1 medium, 0 large, -1 by cmax;
   NULL clears); fronts of order <= 64 are always factored whole in LDS.
   Both rebuild the schedule of the current analysis like
   mfd_set_update_period.  mfd_get_front_times: per node, the duration in
   microseconds of its workgroup in the last factorization (in-kernel
   clocks), -1 for fronts that are not medium. */
int mfd_set_front_split(mfd_handle* h, int cmax);
int mfd_set_front_kinds(mfd_handle* h, const int* kinds);
int mfd_get_front_times(mfd_handle* h, double* us);

/* Tile shape of the Schur updates (no counterpart in the reference, whose
   tiles are nb x nb): fronts with at least min_rows update rows update
   their Schur complement in 128 x 128 tiles, the others in 64 x 64 ones;
   0 (the default): 64 x 64 only.  Rebuilds the schedule like
   mfd_set_update_period. */
int mfd_set_tile_bound(mfd_handle* h, int min_rows);
int mfd_get_tile_bound(mfd_handle* h);

/* Block width of the solves (no counterpart in the reference, whose sweeps
   are one DTRSV and one DGEMV per front, seq_solve/seq_forward_m.f90:14-81,
   seq_solve/seq_backward_m.f90:14-131): fronts with at least min_c pivot
   columns step forward and backward over 256-column superblocks through
   explicit inverses of their 256 x 256 diagonal blocks and the products
   coupling neighbouring superblocks (formed at the end of each
   factorization), the others over 64-column blocks; 0: 64 columns only.
   Default 256.  Rebuilds the schedule like mfd_set_update_period (the
   factorization is lost). */
int mfd_set_solve_block(mfd_handle* h, int min_c);
int mfd_get_solve_block(mfd_handle* h);

/* CPU+GPU hybrid (the reference's subtree partition, tree_pruning_m.f90:
   13-73, which its main path does not call, put to work): the fronts of the
   subtrees rooted at roots[0 .. count) (0-based node ids, e.g. the subtrees
   mfd_tree_pruning deals to one bin) are assembled and factored by
   `threads` host threads while the GPU factors the rest; their factors and
   diagonal-block inverses are copied into device memory before the first
   GPU launch of the lowest level holding a parent of one of the roots, so
   the solves run on the GPU unchanged.  count = 0 clears the set.  A root
   out of range, given twice or inside another root's subtree is
   MFD_ERR_ARG.  While host subtrees are set, only MFD_FP64 factorizations
   are accepted (others: MFD_ERR_UNSUPPORTED) and the factorization is
   never replayed from a graph.  Rebuilds the schedule like
   mfd_set_update_period (the factorization is lost). */
int mfd_set_host_subtrees(mfd_handle* h, const int* roots, int count, int threads);
/* The host part alone, without a GPU: assembles and factors the host
   subtrees from the caller-order values a; mfd_export_front then returns
   their fronts.  Returns 0 or the 1-based global column of the first pivot
   that is not positive. */
int mfd_factorize_host_subtrees(mfd_handle* h, const double* a);
/* The host part of the last factorization: wall time in ms, number of
   fronts and their flops by the reference's cost formula
   (cost_calculator_m.f90:9-20); each pointer nullable. */
int mfd_get_host_part(mfd_handle* h, double* ms, int* fronts, double* flops);

/* Static check of the factorization schedule (no counterpart in the
   reference, whose StarPU runtime orders tasks by their data accesses):
   the launches of the current analysis on their streams (main, zeroing and
   assembly, large in-panel updates) with their event edges as a DAG, and
   every task's accesses to front entries, stored diagonal-block inverses
   and parent entries.  *conflicts = the number of task pairs that touch a
   common entry, one of them writing, without an order: in one launch, or in
   two launches no path of the DAG orders.  Variant 0 checks the schedule as
   built; variant 1 a model of a schedule the library does not use (each
   front's Schur tiles on a fourth stream forked after the launch with the
   front's last diagonal task).  report (nullable) receives the first few
   pairs.  Runs on the host; works without a GPU. */
int mfd_check_schedule(mfd_handle* h, int variant, int64_t* conflicts, char* report, int report_len);

/* The tuner's cost model (no counterpart in the reference): a perceptron
   with MFD_MLP_INPUTS inputs, MFD_MLP_HIDDEN tanh units and one linear
   output, trained and evaluated by HIP kernels in FP64.  params
   (MFD_MLP_PARAMS): W1[j][i] at j*MFD_MLP_INPUTS + i, then b1, w2 (one per
   hidden unit), b2.  x: n x MFD_MLP_INPUTS row-major; t: n targets.  Loss =
   mean over samples of (y - t)^2.  mfd_mlp_gradient returns its gradient
   (MFD_MLP_PARAMS) and the loss; mfd_mlp_train runs `epochs` full-batch
   Adam steps (learning rate lr; moments start at zero) on params in place
   and returns the loss before the last step.  Host buffers; synchronous;
   1 <= n <= 2^20. */
#define MFD_MLP_INPUTS 4
#define MFD_MLP_HIDDEN 32
#define MFD_MLP_PARAMS (MFD_MLP_HIDDEN * (MFD_MLP_INPUTS + 2) + 1)
int mfd_mlp_predict(int n, const double* x, const double* params, double* y);
int mfd_mlp_gradient(int n, const double* x, const double* t, const double* params, double* grad, double* loss);
int mfd_mlp_train(int n, const double* x, const double* t, double* params, int epochs, double lr, double* loss);

const char* mfd_error_string(int code);

#ifdef __cplusplus
}
#endif
#endif
