// HIP kernels of the numeric phase, written for gfx950 (CDNA4, wave64).
//
// Fronts live in one pool, each an n x n column-major block (see NodeDev in
// mfd_plan.h).  Every kernel processes one range of a task array; a
// workgroup owns one task, so no two workgroups of one launch write the same
// front entry and the results do not depend on scheduling.
//
// Reference operations these kernels replace (paths relative to the
// reference root, src/main/):
//   k_scatter_a     set_zero + set_coefficient      zero_setter_m.f90:12-31, coefficient_setter_m.f90:15-50
//   k_extend_add    extend_add_tri / extend_add_rect seq_factorize/kernel/extend_add_kernel_m.f90:6-34
//   k_small_front   the whole per-front tile loop    seq_factorize_m.f90:31-79
//   k_potrf         DPOTRF("U") on a diagonal tile   seq_factorize/kernel/factorize_kernel_m.f90:4-9
//   k_trsm          DTRSM("L","U","T","N")           factorize_kernel_m.f90:11-16, border_kernel_m.f90:21-30
//   k_gemm          DSYRK / DGEMM updates            factorize_kernel_m.f90:18-33, border_subroutines_m.f90:41-73
//   k_solve_fwd     seq_forward (TRSV + GEMV + scatter_add)  seq_solve/seq_forward_m.f90:14-81
//   k_solve_bwd     seq_backward (gather + GEMV + TRSV)      seq_solve/seq_backward_m.f90:14-131
#include <hip/hip_runtime.h>

#include "mfd_kernels.h"

namespace mfd {

namespace {

constexpr int kThreads = 256;

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

// Right-looking Cholesky of the first c columns of an n x n lower matrix held
// in LDS as F[col][row] (n <= 64); the trailing (n-c) x (n-c) block receives
// the Schur complement.  Division (not a reciprocal) keeps integer-valued
// factors exact.  A non-positive pivot reports its global column.
template <typename T>
__device__ void chol_lds(T (*F)[65], int n, int c, int* info, int col0)
{
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int k = 0; k < c; ++k) {
        T d = F[k][k];
        T s = dsqrt(d);
        if (!(d > T(0)) && threadIdx.x == 0) atomicMin(info, col0 + k + 1);
        if (ty == 0 && tx > k && tx < n) F[k][tx] = F[k][tx] / s;
        __syncthreads();
        if (threadIdx.x == 0) F[k][k] = s;
        for (int jj = k + 1 + ty; jj < n; jj += 4)
            if (tx >= jj && tx < n) F[jj][tx] -= F[k][tx] * F[k][jj];
        __syncthreads();
    }
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
// Whole partial factorization of a front with n <= 64 inside LDS.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_small_front(T* pool, const NodeDev* nodes, const int* tasks, int* info)
{
    const NodeDev nd = nodes[tasks[blockIdx.x]];
    __shared__ T F[kSmallFront][65];
    const int n = nd.n, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    T* g = pool + nd.front;
    for (int q = ty; q < n; q += 4)
        if (tx >= q && tx < n) F[q][tx] = g[tx + (int64_t)q * n];
    __syncthreads();
    chol_lds<T>(F, n, nd.c, info, nd.first);
    for (int q = ty; q < n; q += 4)
        if (tx >= q && tx < n) g[tx + (int64_t)q * n] = F[q][tx];
}

// ------------------------------------------------------ large fronts
// Cholesky of the kb x kb diagonal block [k0, k0+kb) of a large front.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_potrf(T* pool, const NodeDev* nodes, const PanelTask* tasks, int* info)
{
    const PanelTask t = tasks[blockIdx.x];
    const NodeDev nd = nodes[t.node];
    __shared__ T D[kPanel][65];
    const int n = nd.n, kb = t.kb, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    T* g = pool + nd.front + t.k0 + (int64_t)t.k0 * n;
    for (int q = ty; q < kb; q += 4)
        if (tx >= q && tx < kb) D[q][tx] = g[tx + (int64_t)q * n];
    __syncthreads();
    chol_lds<T>(D, kb, kb, info, nd.first + t.k0);
    for (int q = ty; q < kb; q += 4)
        if (tx >= q && tx < kb) g[tx + (int64_t)q * n] = D[q][tx];
}

// X L^T = P for the 256 rows [row0, row0+256) of the panel [k0, k0+kb),
// L the factored diagonal block.  One thread per row.  The row lives in
// registers as a shift register: at step k, x[0] is column k; it is
// finished by one division, stored, and the remaining columns are updated
// and shifted down by one.  The inner loop is fully unrolled with static
// register indices (no scratch); L sits in LDS as D[k][m] = L[m][k], zero
// past the block, and is read as a broadcast.  No barrier inside the solve.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_trsm(T* pool, const NodeDev* nodes, const PanelTask* tasks)
{
    const PanelTask t = tasks[blockIdx.x];
    const NodeDev nd = nodes[t.node];
    __shared__ T D[kPanel][2 * kPanel + 1];
    const int n = nd.n, kb = t.kb, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const T* gd = pool + nd.front + t.k0 + (int64_t)t.k0 * n;
    for (int q = ty; q < kPanel; q += 4) {
        D[q][tx] = (q < kb && tx < kb && tx >= q) ? gd[tx + (int64_t)q * n] : T(0);
        D[q][tx + kPanel] = T(0);
    }
    __syncthreads();
    const int row = t.row0 + threadIdx.x;
    if (row >= n) return;
    T* gp = pool + nd.front + row + (int64_t)t.k0 * n;
    T x[kPanel];
#pragma unroll
    for (int j = 0; j < kPanel; ++j) x[j] = j < kb ? gp[(int64_t)j * n] : T(0);
    for (int k = 0; k < kb; ++k) {
        const T x0 = x[0] / D[k][k];
        gp[(int64_t)k * n] = x0;
        const T* dk = &D[k][k];
#pragma unroll
        for (int j = 1; j < kPanel; ++j) x[j - 1] = x[j] - x0 * dk[j];
        x[kPanel - 1] = T(0);
    }
}

// C -= A B^T on one tile (see GemmTask), FP64/FP32 MFMA 16x16x4.  Four waves
// each own a 32 x 32 quarter of the 64 x 64 tile (2 x 2 MFMA blocks).  K is
// staged through LDS in chunks of 16 with the next chunk's global loads
// issued before the current chunk's MFMAs.  The result goes through LDS so
// the read-modify-write of C is coalesced along columns.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_gemm(T* pool, const GemmTask* tasks)
{
    typedef typename Mfma<T>::acc_t acc_t;
    const GemmTask t = tasks[blockIdx.x];
    __shared__ T As[16][kTile + 4];
    __shared__ T Bs[16][kTile + 4];
    __shared__ T Cs[kTile][kTile + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
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
// Forward sweep of one front (the level's fronts run in parallel): gather
// b and the children's update parts (in child order), then L11^{-1} on the
// pivot part and the GEMV update of the rows below, block by block.  The
// diagonal 64 x 64 solves are done by wave 0 with lane shuffles.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_solve_fwd(const T* pool, double* zpool, const double* b, int ldb, int nrhs,
                                                        const NodeDev* nodes, const int* lidx, const int* child_ptr,
                                                        const int* child_list, const int* tasks)
{
    const int s = tasks[blockIdx.x];
    const NodeDev nd = nodes[s];
    const int n = nd.n, c = nd.c, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double* Z = zpool + nd.z * nrhs;
    const T* L = pool + nd.front;
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
    for (int k0 = 0; k0 < c; k0 += 64) {
        const int kb = min(64, c - k0);
        if (wv == 0) {
            for (int r = 0; r < nrhs; ++r) {
                double zi = lane < kb ? Z[k0 + lane + (int64_t)r * n] : 0.0;
                for (int k = 0; k < kb; ++k) {
                    double zk = __shfl(zi, k) / (double)L[(k0 + k) + (int64_t)(k0 + k) * n];
                    if (lane == k) zi = zk;
                    else if (lane > k && lane < kb) zi -= (double)L[(k0 + lane) + (int64_t)(k0 + k) * n] * zk;
                }
                if (lane < kb) Z[k0 + lane + (int64_t)r * n] = zi;
            }
        }
        __syncthreads();
        for (int row = k0 + kb + tid; row < n; row += kThreads)
            for (int r = 0; r < nrhs; ++r) {
                double acc = 0.0;
                for (int k = 0; k < kb; ++k) acc += (double)L[row + (int64_t)(k0 + k) * n] * Z[k0 + k + (int64_t)r * n];
                Z[row + (int64_t)r * n] -= acc;
            }
        __syncthreads();
    }
}

// Backward sweep of one front: gather the parent's solution at the update
// rows, then for pivot blocks in descending order subtract L21^T x and solve
// with the diagonal block transposed; write the pivot part of x.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_solve_bwd(const T* pool, double* zpool, double* x, int ldx, int nrhs,
                                                        const NodeDev* nodes, const int* lidx, const int* tasks)
{
    const int s = tasks[blockIdx.x];
    const NodeDev nd = nodes[s];
    const int n = nd.n, c = nd.c, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double* Z = zpool + nd.z * nrhs;
    const T* L = pool + nd.front;
    if (nd.parent >= 0) {
        const NodeDev pa = nodes[nd.parent];
        const double* Zp = zpool + pa.z * nrhs;
        for (int e = tid; e < nd.w * nrhs; e += kThreads) {
            int i = e % nd.w, r = e / nd.w;
            Z[(c + i) + (int64_t)r * n] = Zp[(lidx[nd.lidx + i] - 1) + (int64_t)r * pa.n];
        }
        __syncthreads();
    }
    const int nblk = (c + 63) / 64;
    for (int bk = nblk - 1; bk >= 0; --bk) {
        const int k0 = bk * 64, kb = min(64, c - k0);
        // z_blk -= L[below, blk]^T z[below]: wave wv takes columns wv, wv+4, ...
        for (int r = 0; r < nrhs; ++r) {
            for (int k = wv; k < kb; k += 4) {
                double acc = 0.0;
                const T* col = L + (int64_t)(k0 + k) * n;
                for (int row = k0 + kb + lane; row < n; row += 64) acc += (double)col[row] * Z[row + (int64_t)r * n];
                for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
                if (lane == 0) Z[k0 + k + (int64_t)r * n] -= acc;
            }
        }
        __syncthreads();
        if (wv == 0) {
            for (int r = 0; r < nrhs; ++r) {
                double zi = lane < kb ? Z[k0 + lane + (int64_t)r * n] : 0.0;
                for (int k = kb - 1; k >= 0; --k) {
                    double xk = __shfl(zi, k) / (double)L[(k0 + k) + (int64_t)(k0 + k) * n];
                    if (lane == k) zi = xk;
                    else if (lane < k) zi -= (double)L[(k0 + k) + (int64_t)(k0 + lane) * n] * xk;
                }
                if (lane < kb) Z[k0 + lane + (int64_t)r * n] = zi;
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < c * nrhs; e += kThreads) {
        int i = e % c, r = e / c;
        x[(int64_t)(nd.first + i) + (int64_t)r * ldx] = Z[i + (int64_t)r * n];
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
    if (count > 0) hipLaunchKernelGGL(k_small_front<T>, dim3(count), dim3(kThreads), 0, st, pool, nodes, tasks, info);
}
template <typename T>
void launch_potrf(T* pool, const NodeDev* nodes, const PanelTask* tasks, int count, int* info, hipStream_t st)
{
    if (count > 0) hipLaunchKernelGGL(k_potrf<T>, dim3(count), dim3(kThreads), 0, st, pool, nodes, tasks, info);
}
template <typename T>
void launch_trsm(T* pool, const NodeDev* nodes, const PanelTask* tasks, int count, hipStream_t st)
{
    if (count > 0) hipLaunchKernelGGL(k_trsm<T>, dim3(count), dim3(kThreads), 0, st, pool, nodes, tasks);
}
template <typename T> void launch_gemm(T* pool, const GemmTask* tasks, int count, hipStream_t st)
{
    if (count > 0) hipLaunchKernelGGL(k_gemm<T>, dim3(count), dim3(kThreads), 0, st, pool, tasks);
}
template <typename T>
void launch_solve_fwd(const T* pool, double* zpool, const double* b, int ldb, int nrhs, const NodeDev* nodes,
                      const int* lidx, const int* child_ptr, const int* child_list, const int* tasks, int count,
                      hipStream_t st)
{
    if (count > 0)
        hipLaunchKernelGGL(k_solve_fwd<T>, dim3(count), dim3(kThreads), 0, st, pool, zpool, b, ldb, nrhs, nodes, lidx,
                           child_ptr, child_list, tasks);
}
template <typename T>
void launch_solve_bwd(const T* pool, double* zpool, double* x, int ldx, int nrhs, const NodeDev* nodes, const int* lidx,
                      const int* tasks, int count, hipStream_t st)
{
    if (count > 0)
        hipLaunchKernelGGL(k_solve_bwd<T>, dim3(count), dim3(kThreads), 0, st, pool, zpool, x, ldx, nrhs, nodes, lidx,
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
    template void launch_potrf<T>(T*, const NodeDev*, const PanelTask*, int, int*, hipStream_t);                       \
    template void launch_trsm<T>(T*, const NodeDev*, const PanelTask*, int, hipStream_t);                              \
    template void launch_gemm<T>(T*, const GemmTask*, int, hipStream_t);                                               \
    template void launch_solve_fwd<T>(const T*, double*, const double*, int, int, const NodeDev*, const int*,          \
                                      const int*, const int*, const int*, int, hipStream_t);                           \
    template void launch_solve_bwd<T>(const T*, double*, double*, int, int, const NodeDev*, const int*, const int*,    \
                                      int, hipStream_t);
MFD_INSTANTIATE(double)
MFD_INSTANTIATE(float)

}  // namespace mfd
