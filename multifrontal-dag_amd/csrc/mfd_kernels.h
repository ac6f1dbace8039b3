// Host-side launchers of the numeric kernels (mfd_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "mfd_plan.h"

namespace mfd {

template <typename T> void launch_fill(T* p, int64_t n, hipStream_t st);
template <typename T> void launch_fill_lower(T* pool, const NodeDev* nodes, const int2* tasks, int count, hipStream_t st);
template <typename T>
void launch_scatter_a(T* pool, const double* a, const int64_t* dst, const int64_t* src, int64_t nnz, hipStream_t st);
// Diagonal scaling of the 16-bit forms (powers of two: S A S is exact):
// s[j] from A(j,j) = a[diag_src[j]] so that s_j^2 A(j,j) lies in [0.5, 2);
// the scaled assembly pool[dst[k]] = s[r] s[c] a[src[k]], (r, c) = rc[k];
// out = diag(s) in per right-hand side.
void launch_diag_scale(double* s, const double* a, const int64_t* diag_src, int n, hipStream_t st);
template <typename T>
void launch_scatter_a_scaled(T* pool, const double* a, const int64_t* dst, const int64_t* src, const int2* rc,
                             const double* s, int64_t nnz, hipStream_t st);
void launch_scale_rows(double* out, const double* in, const double* s, int n, int nrhs, hipStream_t st);
template <typename T>
void launch_extend_add(T* pool, const int* lidx, const NodeDev* nodes, const ExtTask* tasks, int count, hipStream_t st);
template <typename T>
void launch_small_front(T* pool, const NodeDev* nodes, const int* tasks, int count, int* info, hipStream_t st);
// Medium fronts: the whole partial factorization of one front per
// workgroup (tasks: node ids); lowp as for launch_gemm's update tiles;
// clocks (nullable): start and end wall_clock64 per task.
template <typename T>
void launch_front_wg(T* pool, const NodeDev* nodes, const int* tasks, int count, int* info, long long* clocks,
                     hipStream_t st, int lowp = 0);
template <typename T>
void launch_diag(T* pool, const NodeDev* nodes, const DiagTask* tasks, int count, int* info, hipStream_t st);
template <typename T>
void launch_upd_diag(T* pool, const NodeDev* nodes, const DiagTask* diags, int ndiag, const GemmTask* tiles, int ntiles,
                     int* info, hipStream_t st);
// kind 0: panel TRSM tiles; 1: in-panel update; 2: Schur update (mode 3
// tiles read local_index from lidx).  lowp (FP32 fronts, kinds 1 and 2): 1
// multiplies FP16-rounded operands, 2 BF16-rounded ones, on the 16-bit MFMA
// with FP32 accumulation; 0 the front's own precision.
template <typename T>
void launch_gemm(T* pool, const GemmTask* tasks, int count, int kind, const int* lidx, hipStream_t st, int lowp = 0);
// 128 x 128 tiles (m, n <= 128) of modes 0 and 3 with their K chunks (16
// columns) divided evenly over `groups`
// workgroups (segments [seg_ptr[w], seg_ptr[w + 1]) each), then the nfix
// tiles whose K was split summed from their partial products in part
// (16384 entries per slot).
template <typename T>
void launch_gemm_seg(T* pool, const GemmTask* tasks, const SegTask* segs, const int* seg_ptr, int groups,
                     const FixTask* fixes, int nfix, T* part, const int* lidx, hipStream_t st, int lowp = 0);
void launch_fwd_init(double* zpool, const double* b, int ldb, int nrhs, const NodeDev* nodes, const int2* tasks,
                     int count, hipStream_t st);
void launch_fwd_add(double* zpool, int nrhs, const NodeDev* nodes, const int* lidx, const int2* tasks, int count,
                    hipStream_t st);
template <typename T>
void launch_fwd_step(const T* pool, double* zpool, int nrhs, const NodeDev* nodes, const SolveTask* tasks, int count,
                     hipStream_t st);
// The chunk tasks of a forward step alone (row0 >= 0), several workgroups
// per CU.
template <typename T>
void launch_fwd_chunk(const T* pool, double* zpool, int nrhs, const NodeDev* nodes, const SolveTask* tasks, int count,
                      hipStream_t st);
// The chunk tasks of a backward step alone (row0 >= 0).
template <typename T>
void launch_bwd_chunk(const T* pool, const double* zpool, double* part, int nrhs, const NodeDev* nodes,
                      const SolveTask* tasks, int count, hipStream_t st);
// 64 x 64 blocks of the superblock matrices: X by block distance (wv =
// false), then the W and V blocks (wv = true, MFMA tile products)
template <typename T>
void launch_superinv(const T* pool, T* sinv, const NodeDev* nodes, const SuperInvTask* tasks, int count, bool wv,
                     hipStream_t st);
// forward sweep step over 256-column superblocks
template <typename T>
void launch_fwd_super(const T* pool, const T* sinv, double* zpool, int nrhs, const NodeDev* nodes,
                      const SuperTask* tasks, int count, hipStream_t st);
// backward top tasks over 256-column superblocks (the partial sums of the
// rows after superblock u+1 come from k_bwd_step chunk tasks)
template <typename T>
void launch_bwd_super(const T* pool, const T* sinv, double* zpool, const double* part, double* x, int ldx, int nrhs,
                      const NodeDev* nodes, const SuperTask* tasks, int count, hipStream_t st);
// One launch of a forward step's superblock tasks (xtasks, workgroups [0,
// nx)) and its 64-column step tasks (the rest).
template <typename T>
void launch_fwd_pair(const T* pool, const T* sinv, double* zpool, int nrhs, const NodeDev* nodes,
                     const SuperTask* xtasks, int nx, const SolveTask* tasks, int count, hipStream_t st);
// One launch of a backward step's superblock top tasks (xtasks, workgroups
// [0, nx)) and its step tasks (the rest).
template <typename T>
void launch_bwd_pair(const T* pool, const T* sinv, double* zpool, double* part, double* x, int ldx, int nrhs,
                     const NodeDev* nodes, const SuperTask* xtasks, int nx, const SolveTask* tasks, int count,
                     hipStream_t st);
void launch_bwd_gather(double* zpool, int nrhs, const NodeDev* nodes, const int* lidx, const int2* tasks, int count,
                       hipStream_t st);
template <typename T>
void launch_bwd_step(const T* pool, double* zpool, double* part, double* x, int ldx, int nrhs, const NodeDev* nodes,
                     const SolveTask* tasks, int count, hipStream_t st);
void launch_set_int(int* p, int v, hipStream_t st);
void launch_residual(double* r, const double* b, const double* x, int n, int nrhs, const int* ptr, const int* col,
                     const double* val, hipStream_t st);
// val[e] = a[src[e]] for e < nnz (the matrix values in CSR order).
void launch_gather_vals(double* val, const double* a, const int64_t* src, int64_t nnz, hipStream_t st);
// Partials per column of the sums of squares in launch_err_norms:
// sumsq_parts(n) <= kSumParts.
constexpr int kSumParts = 256;
int sumsq_parts(int64_t n);
// The maxima of |r|, |x|, |b| and (d != nullptr) |d| per column into out
// (at 0, nrhs, 2 nrhs, 3 nrhs; zeroed by the caller) and the sums of squares
// of r and x as sumsq_parts(n) partials per column (part, then part +
// sumsq_parts(n) nrhs), in one pass.
void launch_err_norms(unsigned long long* out, double* part, const double* r, const double* x, const double* b,
                      const double* d, int64_t n, int nrhs, hipStream_t st);
void launch_norm_inf(unsigned long long* out, int n, const int* ptr, const double* val, hipStream_t st);
void launch_axpy(double* x, const double* d, int64_t n, hipStream_t st);
void launch_zero_u64(unsigned long long* p, int n, hipStream_t st);
void launch_gather_rows(double* out, int ld_out, const double* in, int ld_in, const int* perm, int n, int nrhs,
                        hipStream_t st);
void launch_scatter_rows(double* out, int ld_out, const double* in, int ld_in, const int* perm, int n, int nrhs,
                         hipStream_t st);

}  // namespace mfd
