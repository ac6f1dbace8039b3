// HIP kernels of the numeric phase, written for gfx950 (CDNA4, wave64).
//
// Fronts live in one pool, each an n x n column-major block (see NodeDev in
// mfd_plan.h).  Every kernel processes one range of a task array; a
// workgroup owns one task, so no two workgroups of one launch write the same
// entry and the results do not depend on scheduling.
//
// Reference operations these kernels replace (paths relative to the
// reference root, src/main/):
//   k_scatter_a     set_zero + set_coefficient        zero_setter_m.f90:12-31, coefficient_setter_m.f90:15-50
//   k_extend_add    extend_add_tri / extend_add_rect  seq_factorize/kernel/extend_add_kernel_m.f90:6-34
//   k_small_front   the per-front tile loop           seq_factorize_m.f90:31-79
//   k_panel         DPOTRF("U") + DTRSM("L","U","T")  seq_factorize/kernel/factorize_kernel_m.f90:4-16,
//                                                     border_kernel_m.f90:9-30
//   k_gemm          DSYRK / DGEMM updates             factorize_kernel_m.f90:18-33, border_subroutines_m.f90:41-73
//   k_fwd_*         seq_forward (scatter_add, TRSV, GEMV)   seq_solve/seq_forward_m.f90:14-81
//   k_bwd_*         seq_backward (gather, GEMV, TRSV)       seq_solve/seq_backward_m.f90:14-131
#include <hip/hip_runtime.h>

#include "mfd_kernels.h"

namespace mfd {

namespace {

constexpr int kThreads = 256;
constexpr int kDLd = 2 * kPanel + 1;  // row length of the LDS copy of a diagonal block

// ------------------------------------------------------------ MFMA traits
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// D(16x16) += A(16x4) B(4x16); lane l supplies A[l%16][l/16] and
// B[l/16][l%16].  The accumulator layouts differ (measured on gfx950):
//   f64: element v of lane l is D[l/16 + 4v][l%16]
//   f32: element v of lane l is D[4(l/16) + v][l%16]
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef d4 acc_t;
    static __device__ acc_t mma(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct Mfma<float> {
    typedef f4 acc_t;
    static __device__ acc_t mma(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

__device__ inline double dsqrt(double x) { return sqrt(x); }
__device__ inline float dsqrt(float x) { return sqrtf(x); }

// Orders one wave's LDS accesses (they complete in issue order within a
// wave) against compiler reordering; no workgroup barrier is involved.
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Row kernels keep one 64-entry row per lane in registers as a shift
// register: at step k, x[j] holds the row's entry of column k + j.  The
// step loop is rolled and the inner loop fully unrolled, so every register
// index is static (no scratch memory) and each step shifts the row by one.
// The 64 steps run in four segments of 16 whose inner width shrinks with
// the segment (64 - K0), so the redundant work past the live columns is at
// most 15 columns per step.

// One wave (threads 0..63) factors the first c columns of an n x n lower
// block (n <= 64) held one row per lane (x[j] = A[lane][j], zero above the
// diagonal).  At step k the pivot comes from lane k by a shuffle and the
// finished column is published through LDS (colbuf: 128 entries, the upper
// 64 zero) and read back as broadcasts.  Column k of L is stored to g
// (leading dimension ld).  A pivot that is not positive records the global
// column col0 + k + 1 in *info (smallest wins).  Afterwards x[j] holds the
// Schur complement entry of column c + j.
template <int K0, typename T>
__device__ __forceinline__ void chol_seg(T (&x)[64], T* colbuf, int n, int c, int* info, int col0, T* g, int64_t ld)
{
    constexpr int W = 64 - K0;
    const int lane = threadIdx.x & 63;
    const int kend = min(K0 + 16, c);
#pragma unroll 1
    for (int k = K0; k < kend; ++k) {
        const T d = __shfl(x[0], k);
        if (!(d > T(0)) && lane == 0) atomicMin(info, col0 + k + 1);
        const T s = dsqrt(d);
        const T l = lane == k ? s : x[0] / s;
        colbuf[lane] = l;
        wave_sync();
        if (lane >= k && lane < n) g[lane + (int64_t)k * ld] = l;
        const T* ck = colbuf + k;
#pragma unroll
        for (int j = 1; j < W; ++j) x[j - 1] = x[j] - l * ck[j];
        wave_sync();
    }
}
template <typename T>
__device__ __forceinline__ void chol_rows(T (&x)[64], T* colbuf, int n, int c, int* info, int col0, T* g, int64_t ld)
{
    chol_seg<0>(x, colbuf, n, c, info, col0, g, ld);
    chol_seg<16>(x, colbuf, n, c, info, col0, g, ld);
    chol_seg<32>(x, colbuf, n, c, info, col0, g, ld);
    chol_seg<48>(x, colbuf, n, c, info, col0, g, ld);
}

// x (one row, kb columns) <- x L^{-T} by forward substitution, L given as
// D[k][m] = L[m][k] (D rows are kDLd >= 2 * 64 long and zero past column
// 63, which the shifted reads D[k][k + j] reach).  Column k of the result
// goes to out[k * stride].
template <int K0, typename T>
__device__ __forceinline__ void row_solve_seg(T (&x)[64], const T (*D)[kDLd], int kb, T* out, int64_t stride)
{
    constexpr int W = 64 - K0;
    const int kend = min(K0 + 16, kb);
#pragma unroll 1
    for (int k = K0; k < kend; ++k) {
        const T* dk = D[k] + k;
        const T x0 = x[0] / dk[0];
        out[(int64_t)k * stride] = x0;
#pragma unroll
        for (int j = 1; j < W; ++j) x[j - 1] = x[j] - x0 * dk[j];
    }
}
template <typename T>
__device__ __forceinline__ void row_solve(T (&x)[64], const T (*D)[kDLd], int kb, T* out, int64_t stride)
{
    row_solve_seg<0>(x, D, kb, out, stride);
    row_solve_seg<16>(x, D, kb, out, stride);
    row_solve_seg<32>(x, D, kb, out, stride);
    row_solve_seg<48>(x, D, kb, out, stride);
}

// The inverse of the kb x kb lower block held in D, stored column-major
// (ld 64, zero-padded to 64 x 64) at inv: row i of the identity solved
// against L^T gives row i of L^{-T}, i.e. column i of L^{-1}.  Threads
// 0..63 take one identity row each.
template <typename T>
__device__ __forceinline__ void store_inverse(const T (*D)[kDLd], int kb, T* inv)
{
    const int i = threadIdx.x;
    if (i >= 64) return;
    T x[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) x[j] = j == i ? T(1) : T(0);
    row_solve<T>(x, D, kb, inv + 64 * i, 1);
    for (int k = kb; k < 64; ++k) inv[64 * i + k] = T(0);
}

}  // namespace

// ---------------------------------------------------------------- assembly
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fill(T* p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * kThreads;
    for (; i < n; i += stride) p[i] = T(0);
}

// pool[dst[k]] = a[src[k]]: each stored A entry lands on exactly one front
// position (coefficient_setter_m.f90:45-50 in the front layout).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_scatter_a(T* pool, const double* a, const int64_t* dst, const int64_t* src,
                                                        int64_t nnz)
{
    int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * kThreads;
    for (; k < nnz; k += stride) pool[dst[k]] = T(a[src[k]]);
}

// ------------------------------------------------------------- extend-add
// One 64 x 64 tile (ti >= tj) of the child's update matrix, lower part only,
// added into the parent front at the rows/columns local_index gives.  The
// map is injective within a child and one launch carries at most one child
// per parent, so the read-modify-write is race-free and the parent receives
// its children in ascending order, as in seq_factorize_m.f90:21-27.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_extend_add(T* pool, const int* lidx, const NodeDev* nodes,
                                                         const ExtTask* tasks)
{
    const ExtTask t = tasks[blockIdx.x];
    const NodeDev ch = nodes[t.child];
    const NodeDev pa = nodes[ch.parent];
    __shared__ int64_t colmap[kTile];
    const int tid = threadIdx.x;
    const int i0 = t.ti * kTile, j0 = t.tj * kTile;
    if (tid < kTile) {
        int j = j0 + tid;
        colmap[tid] = j < ch.w ? (int64_t)(lidx[ch.lidx + j] - 1) * pa.n : 0;
    }
    __syncthreads();
    const int i = i0 + (tid & 63);
    if (i >= ch.w) return;
    const int64_t prow = lidx[ch.lidx + i] - 1;
    const T* src = pool + ch.front + (int64_t)(ch.c + i) + (int64_t)ch.c * ch.n;
    T* dst = pool + pa.front + prow;
    for (int jj = tid >> 6; jj < kTile; jj += 4) {
        int j = j0 + jj;
        if (j >= ch.w || j > i) break;
        dst[colmap[jj]] += src[(int64_t)j * ch.n];
    }
}

// -------------------------------------------------------- small fronts
// Whole partial factorization of a front with n <= 64 by one wave: L in
// the first c columns, the Schur complement in the trailing block, and the
// inverse of the c x c diagonal block for the solves.
template <typename T>
__global__ __launch_bounds__(64) void k_small_front(T* pool, const NodeDev* nodes, const int* tasks, int* info)
{
    const NodeDev nd = nodes[tasks[blockIdx.x]];
    __shared__ T D[kPanel][kDLd];
    __shared__ T colbuf[128];
    const int n = nd.n, c = nd.c, lane = threadIdx.x;
    T* g = pool + nd.front;
    for (int e = lane; e < kPanel * kDLd; e += 64) (&D[0][0])[e] = T(0);
    colbuf[lane] = T(0);
    colbuf[lane + 64] = T(0);
    T x[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) x[j] = (lane < n && j <= lane) ? g[lane + (int64_t)j * n] : T(0);
    __syncthreads();
    chol_rows<T>(x, colbuf, n, c, info, nd.first, &D[0][0], kDLd);
    for (int k = 0; k < c; ++k)
        if (lane >= k && lane < n) g[lane + (int64_t)k * n] = D[k][lane];
#pragma unroll
    for (int j = 0; j < 64; ++j)
        if (lane < n && lane >= c + j) g[lane + (int64_t)(c + j) * n] = x[j];
    store_inverse<T>(D, c, pool + nd.inv);
}

// ------------------------------------------------------ large fronts
// One panel step [k0, k0+kb) of a large front.  Wave 0 of every workgroup
// factors the kb x kb diagonal block itself (it is small, and this removes a
// launch and a dependency from the step).  A chunk task (row0 >= 0) then solves its 256
// rows X L^T = P, one thread per row.  The writer task (row0 < 0) stores the
// factored block into scratch slot `slot` -- the front's own copy is
// rewritten by the update launch that follows, once no workgroup of this
// launch can still be reading the unfactored block -- and the inverse of
// the block for the solves.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_panel(T* pool, const NodeDev* nodes, const SolveTask* tasks,
                                                    int64_t scratch, int* info)
{
    const SolveTask t = tasks[blockIdx.x];
    const NodeDev nd = nodes[t.node];
    __shared__ T D[kPanel][kDLd];
    __shared__ T colbuf[128];
    const int n = nd.n, kb = t.kb, tid = threadIdx.x, lane = tid & 63;
    for (int e = tid; e < kPanel * kDLd; e += kThreads) (&D[0][0])[e] = T(0);
    if (tid < 128) colbuf[tid] = T(0);
    __syncthreads();
    if (tid < 64) {
        const T* gd = pool + nd.front + t.k0 + (int64_t)t.k0 * n;
        T xd[64];
#pragma unroll
        for (int j = 0; j < 64; ++j) xd[j] = (lane < kb && j <= lane) ? gd[lane + (int64_t)j * n] : T(0);
        chol_rows<T>(xd, colbuf, kb, kb, info, nd.first + t.k0, &D[0][0], kDLd);
    }
    __syncthreads();
    if (t.row0 >= 0) {
        const int row = t.row0 + tid;
        if (row >= n) return;
        T* gp = pool + nd.front + row + (int64_t)t.k0 * n;
        T x[64];
#pragma unroll
        for (int j = 0; j < 64; ++j) x[j] = j < kb ? gp[(int64_t)j * n] : T(0);
        row_solve<T>(x, D, kb, gp, n);
    } else {
        T* sc = pool + scratch + (int64_t)t.slot * kPanel * kPanel;
        for (int e = tid; e < kPanel * kPanel; e += kThreads) {
            int i = e & 63, k = e >> 6;
            sc[e] = (k < kb && i < kb && i >= k) ? D[k][i] : T(0);
        }
        store_inverse<T>(D, kb, pool + nd.inv + (int64_t)(t.k0 / kPanel) * kPanel * kPanel);
    }
}

// C -= A B^T on one tile (see GemmTask), FP64/FP32 MFMA 16x16x4.  Four waves
// each own a 32 x 32 quarter of the 64 x 64 tile (2 x 2 MFMA blocks).  K is
// staged through LDS in chunks of 16 with the next chunk's global loads
// issued before the current chunk's MFMAs.  The result goes through LDS so
// the read-modify-write of C is coalesced along columns.  k == 0 copies the
// lower part of A into C instead (the factored diagonal block of a panel).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_gemm(T* pool, const GemmTask* tasks)
{
    typedef typename Mfma<T>::acc_t acc_t;
    const GemmTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (t.k == 0) {
        const T* A = pool + t.a;
        T* C = pool + t.c;
        for (int e = tid; e < t.m * t.n; e += kThreads) {
            int i = e % t.m, j = e / t.m;
            if (i >= j) C[i + (int64_t)j * t.ldc] = A[i + (int64_t)j * t.lda];
        }
        return;
    }
    __shared__ T As[16][kTile + 4];
    __shared__ T Bs[16][kTile + 4];
    __shared__ T Cs[kTile][kTile + 1];
    const int wr = (wv & 1) * 32, wc = (wv >> 1) * 32;
    const int li = tid & 63, lp = tid >> 6;
    const T* A = pool + t.a;
    const T* B = pool + t.b;
    acc_t c00 = {0, 0, 0, 0}, c01 = {0, 0, 0, 0}, c10 = {0, 0, 0, 0}, c11 = {0, 0, 0, 0};
    T ra[4], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            int kk = k0 + lp + 4 * p;
            ra[p] = (li < t.m && kk < t.k) ? A[li + (int64_t)kk * t.lda] : T(0);
            rb[p] = (li < t.n && kk < t.k) ? B[li + (int64_t)kk * t.ldb] : T(0);
        }
    };
    load(0);
    for (int k0 = 0; k0 < t.k; k0 += 16) {
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            As[lp + 4 * p][li] = ra[p];
            Bs[lp + 4 * p][li] = rb[p];
        }
        __syncthreads();
        if (k0 + 16 < t.k) load(k0 + 16);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int kr = q * 4 + (lane >> 4);
            const T a0 = As[kr][wr + (lane & 15)], a1 = As[kr][wr + 16 + (lane & 15)];
            const T b0 = Bs[kr][wc + (lane & 15)], b1 = Bs[kr][wc + 16 + (lane & 15)];
            c00 = Mfma<T>::mma(a0, b0, c00);
            c01 = Mfma<T>::mma(a0, b1, c01);
            c10 = Mfma<T>::mma(a1, b0, c10);
            c11 = Mfma<T>::mma(a1, b1, c11);
        }
        __syncthreads();
    }
    const int q0 = wc + (lane & 15);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = wr + Mfma<T>::row(lane, v);
        Cs[q0][r] = c00[v];
        Cs[q0 + 16][r] = c01[v];
        Cs[q0][r + 16] = c10[v];
        Cs[q0 + 16][r + 16] = c11[v];
    }
    __syncthreads();
    T* C = pool + t.c;
    const int row = tid & 63;
    if (row < t.m)
        for (int col = tid >> 6; col < t.n; col += 4)
            if (!t.tri || row >= col) C[row + (int64_t)col * t.ldc] -= Cs[col][row];
}

// ------------------------------------------------------------------ solve
// Blocked supernodal sweeps, level by level (the level's fronts in one
// launch).  Front s's solve slice (per right-hand side, ld n and ld c):
// Z[0:n] is the right-hand side being reduced (pivot part, then the update
// part passed to the parent), Y[0:c] the solved pivot values (forward: y,
// backward: x).  Diagonal blocks are applied through their stored inverses,
// so no kernel runs a sequential loop: forward one launch per 64-column
// block, backward two (seq_forward_m.f90:14-81, seq_backward_m.f90:14-131).
constexpr int kRhsGroup = 8;  // right-hand sides processed together

// Forward, first launch of a level: pivot part of Z from b, update part
// zero, then the children's update parts added in child order (scatter_add,
// sparse_add_subroutines_m.f90:13-30).
__global__ __launch_bounds__(kThreads) void k_fwd_gather(double* zpool, const double* b, int ldb, int nrhs,
                                                         const NodeDev* nodes, const int* lidx, const int* child_ptr,
                                                         const int* child_list, const int* tasks)
{
    const int s = tasks[blockIdx.x];
    const NodeDev nd = nodes[s];
    const int n = nd.n, c = nd.c, tid = threadIdx.x;
    double* Z = zpool + nd.z * nrhs;
    for (int e = tid; e < n * nrhs; e += kThreads) {
        int i = e % n, r = e / n;
        Z[i + (int64_t)r * n] = i < c ? b[(int64_t)(nd.first + i) + (int64_t)r * ldb] : 0.0;
    }
    __syncthreads();
    for (int cp = child_ptr[s]; cp < child_ptr[s + 1]; ++cp) {
        const NodeDev ch = nodes[child_list[cp]];
        const double* Zc = zpool + ch.z * nrhs;
        for (int e = tid; e < ch.w * nrhs; e += kThreads) {
            int i = e % ch.w, r = e / ch.w;
            Z[(lidx[ch.lidx + i] - 1) + (int64_t)r * n] += Zc[(ch.c + i) + (int64_t)r * ch.n];
        }
        __syncthreads();
    }
}

// Stage inv(L_tt) (64 x 64, column-major in the pool) in LDS as LI[col][row].
template <typename T>
__device__ void load_inverse(const T* inv, double (*LI)[65])
{
#pragma unroll
    for (int e = threadIdx.x; e < 64 * 64; e += kThreads) LI[e >> 6][e & 63] = (double)inv[e];
}

// Forward block step: y_t = inv(L_tt) Z[k0:k0+kb] (every workgroup computes
// it from the unmodified Z block); the writer task (row0 < 0) stores y_t in
// Y, each chunk task subtracts L[rows, t] y_t from its 256 rows of Z
// (DTRSV "U","T" and DGEMV "T", solve_kernel_m.f90:4-23).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fwd_step(const T* pool, double* zpool, int nrhs, const NodeDev* nodes,
                                                       const SolveTask* tasks)
{
    const SolveTask t = tasks[blockIdx.x];
    const NodeDev nd = nodes[t.node];
    __shared__ double LI[64][65];
    __shared__ double zt[64][kRhsGroup], yt[64][kRhsGroup];
    const int n = nd.n, c = nd.c, kb = t.kb, tid = threadIdx.x, lane = tid & 63, ty = tid >> 6;
    double* Z = zpool + nd.z * nrhs;
    double* Y = Z + (int64_t)n * nrhs;
    load_inverse<T>(pool + nd.inv + (int64_t)(t.k0 / 64) * 4096, LI);
    const int row = t.row0 + tid;
    const T* L = pool + nd.front + row + (int64_t)t.k0 * n;
    for (int r0 = 0; r0 < nrhs; r0 += kRhsGroup) {
        const int nr = min(kRhsGroup, nrhs - r0);
        for (int e = tid; e < 64 * kRhsGroup; e += kThreads) {
            int k = e & 63, r = e >> 6;
            zt[k][r] = (k < kb && r < nr) ? Z[t.k0 + k + (int64_t)(r0 + r) * n] : 0.0;
        }
        __syncthreads();
        for (int r = ty; r < kRhsGroup; r += 4) {
            double acc = 0.0;
            for (int k = 0; k < kb; ++k) acc += LI[k][lane] * zt[k][r];
            yt[lane][r] = acc;
        }
        __syncthreads();
        if (t.row0 < 0) {
            for (int e = tid; e < kb * nr; e += kThreads) {
                int i = e % kb, r = e / kb;
                Y[t.k0 + i + (int64_t)(r0 + r) * c] = yt[i][r];
            }
        } else if (row < n) {
            double acc[kRhsGroup];
#pragma unroll
            for (int r = 0; r < kRhsGroup; ++r) acc[r] = 0.0;
            // 16 column loads in flight per group (zero past kb; the clamped
            // yt index stays in the array and meets a zero factor)
#pragma unroll 1
            for (int k0 = 0; k0 < kb; k0 += 16) {
                double l[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) l[k] = k0 + k < kb ? (double)L[(int64_t)(k0 + k) * n] : 0.0;
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const int kk = min(k0 + k, 63);
#pragma unroll
                    for (int r = 0; r < kRhsGroup; ++r) acc[r] += l[k] * yt[kk][r];
                }
            }
#pragma unroll
            for (int r = 0; r < kRhsGroup; ++r)
                if (r < nr) Z[row + (int64_t)(r0 + r) * n] -= acc[r];
        }
        __syncthreads();
    }
}

// Backward, first launch of a level: the update part of Z takes the
// parent's solution at the parent-front positions -- pivot positions from
// the parent's Y, update positions from the parent's Z (gather,
// sparse_add_subroutines_m.f90:32-49).
__global__ __launch_bounds__(kThreads) void k_bwd_gather(double* zpool, int nrhs, const NodeDev* nodes, const int* lidx,
                                                         const int* tasks)
{
    const NodeDev nd = nodes[tasks[blockIdx.x]];
    if (nd.parent < 0) return;
    const NodeDev pa = nodes[nd.parent];
    double* Z = zpool + nd.z * nrhs;
    const double* Zp = zpool + pa.z * nrhs;
    const double* Yp = Zp + (int64_t)pa.n * nrhs;
    for (int e = threadIdx.x; e < nd.w * nrhs; e += kThreads) {
        int i = e % nd.w, r = e / nd.w;
        int p = lidx[nd.lidx + i] - 1;
        Z[(nd.c + i) + (int64_t)r * nd.n] = p < pa.c ? Yp[p + (int64_t)r * pa.c] : Zp[p + (int64_t)r * pa.n];
    }
}

// Backward: partial dot products P[slot][k][r] = sum over this chunk's rows
// of L[row][k0+k] x[row][r], x from Y (pivot rows) or Z (update rows)
// (DGEMV "N", solve_kernel_m.f90:25-30).  64-row sub-blocks of L are staged
// transposed through LDS so the global reads and the per-column dot
// products are both conflict-free.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_bwd_partial(const T* pool, double* zpool, double* part, int nrhs,
                                                          const NodeDev* nodes, const SolveTask* tasks)
{
    const SolveTask t = tasks[blockIdx.x];
    const NodeDev nd = nodes[t.node];
    __shared__ double Ls[64][65];  // Ls[k][i] = L[row_i][k0+k]
    __shared__ double Xs[64][kRhsGroup];
    const int n = nd.n, c = nd.c, kb = t.kb, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const double* Z = zpool + nd.z * nrhs;
    const double* Y = Z + (int64_t)n * nrhs;
    const T* L = pool + nd.front + (int64_t)t.k0 * n;
    const int rend = min(n, t.row0 + 256);
    double* P = part + (int64_t)t.slot * 64 * nrhs;
    for (int r0 = 0; r0 < nrhs; r0 += kRhsGroup) {
        const int nr = min(kRhsGroup, nrhs - r0);
        double acc0 = 0.0, acc1 = 0.0;  // (k = tx, r = ty) and (k = tx, r = ty + 4)
        for (int rb = t.row0; rb < rend; rb += 64) {
            const int rows = min(64, rend - rb);
            __syncthreads();
#pragma unroll
            for (int q = ty; q < 64; q += 4)
                if (q < kb) Ls[q][tx] = tx < rows ? (double)L[rb + tx + (int64_t)q * n] : 0.0;
            for (int e = threadIdx.x; e < 64 * nr; e += kThreads) {
                int i = e & 63, r = e >> 6, g = rb + i;
                Xs[i][r] = i >= rows ? 0.0 : g < c ? Y[g + (int64_t)(r0 + r) * c] : Z[g + (int64_t)(r0 + r) * n];
            }
            __syncthreads();
            if (tx < kb) {
                if (ty < nr)
                    for (int i = 0; i < rows; ++i) acc0 += Ls[tx][i] * Xs[i][ty];
                if (ty + 4 < nr)
                    for (int i = 0; i < rows; ++i) acc1 += Ls[tx][i] * Xs[i][ty + 4];
            }
        }
        if (tx < kb) {
            if (ty < nr) P[tx + (int64_t)(r0 + ty) * 64] = acc0;
            if (ty + 4 < nr) P[tx + (int64_t)(r0 + ty + 4) * 64] = acc1;
        }
    }
}

// Backward block step: z = Y[k0:k0+kb] - sum of the chunk partials (in
// chunk order), x_t = inv(L_tt)^T z into Y (DTRSV "U","N",
// solve_kernel_m.f90:11-16).  The front's first block writes its pivot part
// of x out.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_bwd_diag(const T* pool, double* zpool, const double* part, double* x,
                                                       int ldx, int nrhs, const NodeDev* nodes, const SolveTask* tasks)
{
    const SolveTask t = tasks[blockIdx.x];
    const NodeDev nd = nodes[t.node];
    __shared__ double LI[64][65];
    __shared__ double zt[64][kRhsGroup];
    const int n = nd.n, c = nd.c, kb = t.kb, tid = threadIdx.x, lane = tid & 63, ty = tid >> 6;
    double* Y = zpool + nd.z * nrhs + (int64_t)n * nrhs;
    load_inverse<T>(pool + nd.inv + (int64_t)(t.k0 / 64) * 4096, LI);
    for (int r0 = 0; r0 < nrhs; r0 += kRhsGroup) {
        const int nr = min(kRhsGroup, nrhs - r0);
        for (int e = tid; e < 64 * kRhsGroup; e += kThreads) {
            int k = e & 63, r = e >> 6;
            double z = 0.0;
            if (k < kb && r < nr) {
                z = Y[t.k0 + k + (int64_t)(r0 + r) * c];
                for (int ch = 0; ch < t.nchunk; ++ch) z -= part[(int64_t)(t.slot + ch) * 64 * nrhs + k + (int64_t)(r0 + r) * 64];
            }
            zt[k][r] = z;
        }
        __syncthreads();
        for (int r = ty; r < nr; r += 4) {
            double acc = 0.0;
            for (int m = 0; m < kb; ++m) acc += LI[lane][m] * zt[m][r];
            if (lane < kb) {
                Y[t.k0 + lane + (int64_t)(r0 + r) * c] = acc;
                if (t.k0 == 0) x[(int64_t)(nd.first + lane) + (int64_t)(r0 + r) * ldx] = acc;
            }
        }
        __syncthreads();
    }
    // The front's first block is its last one solved: the other blocks'
    // values were written to Y by earlier launches.
    if (t.k0 == 0 && c > kb) {
        const int rest = c - kb;
        for (int e = tid; e < rest * nrhs; e += kThreads) {
            int i = kb + e % rest, r = e / rest;
            x[(int64_t)(nd.first + i) + (int64_t)r * ldx] = Y[i + (int64_t)r * c];
        }
    }
}

// ----------------------------------------------- permutations and residual
// out[i + r*ld_out] = in[perm[i] + r*ld_in] (perm 0-based)
__global__ __launch_bounds__(kThreads) void k_gather_rows(double* out, int ld_out, const double* in, int ld_in,
                                                          const int* perm, int n, int nrhs)
{
    int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (e >= (int64_t)n * nrhs) return;
    int i = (int)(e % n), r = (int)(e / n);
    out[i + (int64_t)r * ld_out] = in[perm[i] + (int64_t)r * ld_in];
}
// out[perm[i] + r*ld_out] = in[i + r*ld_in]
__global__ __launch_bounds__(kThreads) void k_scatter_rows(double* out, int ld_out, const double* in, int ld_in,
                                                           const int* perm, int n, int nrhs)
{
    int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (e >= (int64_t)n * nrhs) return;
    int i = (int)(e % n), r = (int)(e / n);
    out[perm[i] + (int64_t)r * ld_out] = in[i + (int64_t)r * ld_in];
}

// ------------------------------------------------------ FP64 refinement
// r = b - A x for the full symmetric matrix in CSR (postordered numbering);
// entry e's value is a[src[e]] (values stay in the caller's order).  One
// wave per row, so the row sum is reduced in a fixed order.
__global__ __launch_bounds__(kThreads) void k_residual(double* r, const double* b, const double* x, int n, int nrhs,
                                                       const int* ptr, const int* col, const int64_t* src,
                                                       const double* a)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= n) return;
    for (int k = 0; k < nrhs; ++k) {
        double acc = 0.0;
        for (int e = ptr[row] + lane; e < ptr[row + 1]; e += 64) acc += a[src[e]] * x[col[e] + (int64_t)k * n];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if (lane == 0) r[row + (int64_t)k * n] = b[row + (int64_t)k * n] - acc;
    }
}

// out[k] = max(out[k], max_i |x[i + k*ld]|) as IEEE bits (order-preserving
// for non-negative doubles, so the atomic max is exact and deterministic).
__global__ __launch_bounds__(kThreads) void k_absmax(unsigned long long* out, const double* x, int64_t n, int ld,
                                                     int nrhs)
{
    const int k = blockIdx.y;
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        m = fmax(m, fabs(x[i + (int64_t)k * ld]));
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out + k, (unsigned long long)__double_as_longlong(m));
}

// out[0] = max over rows of sum_e |a[src[e]]|  (the infinity norm of A)
__global__ __launch_bounds__(kThreads) void k_norm_inf(unsigned long long* out, int n, const int* ptr,
                                                       const int64_t* src, const double* a)
{
    const int row = blockIdx.x * kThreads + threadIdx.x;
    double s = 0.0;
    if (row < n)
        for (int e = ptr[row]; e < ptr[row + 1]; ++e) s += fabs(a[src[e]]);
    for (int off = 32; off > 0; off >>= 1) s = fmax(s, __shfl_xor(s, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(s));
}

__global__ __launch_bounds__(kThreads) void k_axpy(double* x, const double* d, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) x[i] += d[i];
}

__global__ void k_zero_u64(unsigned long long* p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

void launch_residual(double* r, const double* b, const double* x, int n, int nrhs, const int* ptr, const int* col,
                     const int64_t* src, const double* a, hipStream_t st)
{
    if (n > 0)
        hipLaunchKernelGGL(k_residual, dim3((n + 3) / 4), dim3(kThreads), 0, st, r, b, x, n, nrhs, ptr, col, src, a);
}
void launch_absmax(unsigned long long* out, const double* x, int64_t n, int ld, int nrhs, hipStream_t st)
{
    int g = (int)std::min<int64_t>((n + kThreads - 1) / kThreads, 1024);
    if (n > 0) hipLaunchKernelGGL(k_absmax, dim3(std::max(g, 1), nrhs), dim3(kThreads), 0, st, out, x, n, ld, nrhs);
}
void launch_norm_inf(unsigned long long* out, int n, const int* ptr, const int64_t* src, const double* a, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(k_norm_inf, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, st, out, n, ptr, src, a);
}
void launch_axpy(double* x, const double* d, int64_t n, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, x, d, n);
}
void launch_zero_u64(unsigned long long* p, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, p, n);
}

__global__ void k_set_int(int* p, int v) { *p = v; }

// ------------------------------------------------------------- launchers
static inline int grid_for(int64_t n) { return (int)std::min<int64_t>((n + kThreads - 1) / kThreads, 65536); }

void launch_set_int(int* p, int v, hipStream_t st) { hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, p, v); }

template <typename T> void launch_fill(T* p, int64_t n, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n)), dim3(kThreads), 0, st, p, n);
}
template <typename T>
void launch_scatter_a(T* pool, const double* a, const int64_t* dst, const int64_t* src, int64_t nnz, hipStream_t st)
{
    if (nnz > 0) hipLaunchKernelGGL(k_scatter_a<T>, dim3(grid_for(nnz)), dim3(kThreads), 0, st, pool, a, dst, src, nnz);
}
template <typename T>
void launch_extend_add(T* pool, const int* lidx, const NodeDev* nodes, const ExtTask* tasks, int count, hipStream_t st)
{
    if (count > 0) hipLaunchKernelGGL(k_extend_add<T>, dim3(count), dim3(kThreads), 0, st, pool, lidx, nodes, tasks);
}
template <typename T>
void launch_small_front(T* pool, const NodeDev* nodes, const int* tasks, int count, int* info, hipStream_t st)
{
    if (count > 0) hipLaunchKernelGGL(k_small_front<T>, dim3(count), dim3(64), 0, st, pool, nodes, tasks, info);
}
template <typename T>
void launch_panel(T* pool, const NodeDev* nodes, const SolveTask* tasks, int count, int64_t scratch, int* info,
                  hipStream_t st)
{
    if (count > 0) hipLaunchKernelGGL(k_panel<T>, dim3(count), dim3(kThreads), 0, st, pool, nodes, tasks, scratch, info);
}
template <typename T> void launch_gemm(T* pool, const GemmTask* tasks, int count, hipStream_t st)
{
    if (count > 0) hipLaunchKernelGGL(k_gemm<T>, dim3(count), dim3(kThreads), 0, st, pool, tasks);
}
void launch_fwd_gather(double* zpool, const double* b, int ldb, int nrhs, const NodeDev* nodes, const int* lidx,
                       const int* child_ptr, const int* child_list, const int* tasks, int count, hipStream_t st)
{
    if (count > 0)
        hipLaunchKernelGGL(k_fwd_gather, dim3(count), dim3(kThreads), 0, st, zpool, b, ldb, nrhs, nodes, lidx, child_ptr,
                           child_list, tasks);
}
template <typename T>
void launch_fwd_step(const T* pool, double* zpool, int nrhs, const NodeDev* nodes, const SolveTask* tasks, int count,
                     hipStream_t st)
{
    if (count > 0) hipLaunchKernelGGL(k_fwd_step<T>, dim3(count), dim3(kThreads), 0, st, pool, zpool, nrhs, nodes, tasks);
}
void launch_bwd_gather(double* zpool, int nrhs, const NodeDev* nodes, const int* lidx, const int* tasks, int count,
                       hipStream_t st)
{
    if (count > 0) hipLaunchKernelGGL(k_bwd_gather, dim3(count), dim3(kThreads), 0, st, zpool, nrhs, nodes, lidx, tasks);
}
template <typename T>
void launch_bwd_partial(const T* pool, double* zpool, double* part, int nrhs, const NodeDev* nodes,
                        const SolveTask* tasks, int count, hipStream_t st)
{
    if (count > 0)
        hipLaunchKernelGGL(k_bwd_partial<T>, dim3(count), dim3(kThreads), 0, st, pool, zpool, part, nrhs, nodes, tasks);
}
template <typename T>
void launch_bwd_diag(const T* pool, double* zpool, const double* part, double* x, int ldx, int nrhs,
                     const NodeDev* nodes, const SolveTask* tasks, int count, hipStream_t st)
{
    if (count > 0)
        hipLaunchKernelGGL(k_bwd_diag<T>, dim3(count), dim3(kThreads), 0, st, pool, zpool, part, x, ldx, nrhs, nodes,
                           tasks);
}
void launch_gather_rows(double* out, int ld_out, const double* in, int ld_in, const int* perm, int n, int nrhs,
                        hipStream_t st)
{
    int64_t tot = (int64_t)n * nrhs;
    if (tot > 0)
        hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((tot + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, out,
                           ld_out, in, ld_in, perm, n, nrhs);
}
void launch_scatter_rows(double* out, int ld_out, const double* in, int ld_in, const int* perm, int n, int nrhs,
                         hipStream_t st)
{
    int64_t tot = (int64_t)n * nrhs;
    if (tot > 0)
        hipLaunchKernelGGL(k_scatter_rows, dim3((unsigned)((tot + kThreads - 1) / kThreads)), dim3(kThreads), 0, st,
                           out, ld_out, in, ld_in, perm, n, nrhs);
}

#define MFD_INSTANTIATE(T)                                                                                             \
    template void launch_fill<T>(T*, int64_t, hipStream_t);                                                            \
    template void launch_scatter_a<T>(T*, const double*, const int64_t*, const int64_t*, int64_t, hipStream_t);        \
    template void launch_extend_add<T>(T*, const int*, const NodeDev*, const ExtTask*, int, hipStream_t);              \
    template void launch_small_front<T>(T*, const NodeDev*, const int*, int, int*, hipStream_t);                       \
    template void launch_panel<T>(T*, const NodeDev*, const SolveTask*, int, int64_t, int*, hipStream_t);              \
    template void launch_gemm<T>(T*, const GemmTask*, int, hipStream_t);                                               \
    template void launch_fwd_step<T>(const T*, double*, int, const NodeDev*, const SolveTask*, int, hipStream_t);       \
    template void launch_bwd_partial<T>(const T*, double*, double*, int, const NodeDev*, const SolveTask*, int,        \
                                        hipStream_t);                                                                  \
    template void launch_bwd_diag<T>(const T*, double*, const double*, double*, int, int, const NodeDev*,              \
                                     const SolveTask*, int, hipStream_t);
MFD_INSTANTIATE(double)
MFD_INSTANTIATE(float)

}  // namespace mfd
