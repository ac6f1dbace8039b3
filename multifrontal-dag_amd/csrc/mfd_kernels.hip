// HIP kernels of the numeric phase, written for gfx950 (CDNA4, wave64).
//
// Fronts live in one pool, each an n x n column-major block (see NodeDev in
// mfd_plan.h).  Every kernel processes one range of a task array; a
// workgroup owns one task, so no two workgroups of one launch write the same
// entry and the results do not depend on scheduling.
//
// Reference operations these kernels replace (paths relative to the
// reference root, src/main/):
//   k_fill_lower,   set_zero + set_coefficient        zero_setter_m.f90:12-31, coefficient_setter_m.f90:15-50
//   k_scatter_a
//   k_extend_add    extend_add_tri / extend_add_rect  seq_factorize/kernel/extend_add_kernel_m.f90:6-34
//                   (small children; k_gemm mode 3 for the others)
//   k_small_front   the per-front tile loop           seq_factorize_m.f90:31-79
//   k_diag,         DPOTRF("U") of a diagonal block,  seq_factorize/kernel/factorize_kernel_m.f90:4-16,
//   k_upd_diag      with its panel's DTRSM tiles      border_kernel_m.f90:9-30
//   k_gemm          DTRSM (modes 1, 2), DSYRK / DGEMM factorize_kernel_m.f90:11-33, border_kernel_m.f90:21-30,
//                   updates (modes 0, 3)              border_subroutines_m.f90:41-73
//   k_fwd_*         seq_forward (scatter_add, TRSV, GEMV)   seq_solve/seq_forward_m.f90:14-81
//   k_bwd_*         seq_backward (gather, GEMV, TRSV)       seq_solve/seq_backward_m.f90:14-131
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

// 16-bit operands with FP32 accumulation (the FP16_IR / BF16_IR forms):
// D(16x16) += A(16x16) B(16x16); lane l supplies A[l%16][4(l/16) + i] and
// B[4(l/16) + i][l%16], i < 4; the accumulator layout is the f32 one above.
// Operands are rounded to nearest even from FP32.
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));
enum LowPrec { kNative = 0, kF16 = 1, kBF16 = 2 };
__device__ __forceinline__ short bf16_bits(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (short)(u >> 16);
}
template <int LP> struct Mfma16;
template <> struct Mfma16<kF16> {
    typedef h4 op_t;
    static __device__ _Float16 cvt(float x) { return (_Float16)x; }
    static __device__ f4 mma(op_t a, op_t b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0); }
};
template <> struct Mfma16<kBF16> {
    typedef s4 op_t;
    static __device__ short cvt(float x) { return bf16_bits(x); }
    static __device__ f4 mma(op_t a, op_t b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
};


// 1 / sqrt(d): the hardware estimate (v_rsq) refined by Newton-Raphson
// steps r <- r + (r / 2) (1 - d r^2), each squaring the relative error (two
// for double, one for float): within a few ulp of the exact value, with a
// dependent chain of seven steps where a correctly rounded sqrt and
// division take about twenty.
__device__ __forceinline__ double rsqrt_nr(double d)
{
    double r = __builtin_amdgcn_rsq(d);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double e = fma(-(d * r), r, 1.0);
        r = fma(0.5 * r, e, r);
    }
    return r;
}
__device__ __forceinline__ float rsqrt_nr(float d)
{
    const float r = __builtin_amdgcn_rsqf(d);
    const float e = fmaf(-(d * r), r, 1.0f);
    return fmaf(0.5f * r, e, r);
}
// a / b given rb ~ 1 / b (within a few ulp): q = a rb, then one correction
// with the exact residual a - q b (Markstein), which returns a / b exactly
// whenever it is representable and is within an ulp of it otherwise.  Its
// chain is a multiply and two fmas after rb and b, against the ten
// dependent steps of a correctly rounded division.
template <typename T> __device__ __forceinline__ T quot(T a, T b, T rb)
{
    const T q = a * rb;
    return fma(fma(-q, b, a), rb, q);
}

// p[i] for a uniform base p and a 32-bit unsigned index: the address is
// base + zero-extended byte offset, which keeps one VGPR per address when
// many loads are in flight.
template <typename T> __device__ __forceinline__ T ld_off(const T* p, unsigned i)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + (size_t)(i * (unsigned)sizeof(T)));
}

// Value of v in lane src (uniform src; only its low 6 bits select the lane).
__device__ __forceinline__ double bcast(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float bcast(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// Orders one wave's LDS accesses (they complete in issue order within a
// wave) against compiler reordering; no workgroup barrier is involved.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------- dense 64 x 64 partial Cholesky
// A front (or diagonal block) of order n <= 64 in LDS, factored in its first
// c columns by a blocked right-looking algorithm with 16-column panels:
//   - each panel is factored by one wave, one row per lane, in a 16-step
//     chain: the pivot and the finished column's entry of the next column's
//     row come from their lanes by readlane, the entries the later columns
//     need are read back from LDS, where each column is stored as soon as it
//     is finished;
//   - the panel's 16 x 16 diagonal block is inverted by another wave during
//     the next panel's chain (lane r solves L u = e_r);
//   - the rank-16 update of the trailing blocks runs on MFMA across the four
//     waves;
//   - the off-diagonal 16 x 16 blocks of L^{-1} follow on MFMA, by distance
//     from the diagonal: X_ij = -inv(L_ii) sum_{k=j}^{i-1} L_ik X_kj.
// Layout: A[col][row] holds the lower triangle (zero past n); afterwards it
// holds L in columns < c and the Schur complement in the trailing block.
// X[col][row] = L^{-1}[row][col] for row, col < c (other entries are
// scratch).  Upper-triangle entries of A are scratch too.  rs[k] = 1 / L_kk.
template <typename T> struct Front64 {
    T A[64][65];
    T X[64][65];
    T Ts[4][16][17];
    T rs[64];
};

// sqrt(d) from r ~ 1/sqrt(d) (within a few ulp): s = d r and one Newton
// step s + (d - s^2) r / 2 with the exact residual (fma); exact whenever
// sqrt(d) is representable, within an ulp otherwise, in three dependent
// operations after r.
template <typename T> __device__ __forceinline__ T sqrt_from_rs(T d, T r)
{
    const T s = d * r;
    return fma(fma(-s, s, d), T(0.5) * r, s);
}

// Panel b16 .. b16 + cb - 1 (cb <= 16; FULL: cb == 16) by one wave, rows
// b16 .. 63 on their lanes.  Column k: d = x_k on lane k, L_kk = sqrt(d),
// l = x_k / L_kk (quot: exact whenever the quotient is representable); the
// column goes to LDS at once, its entry for row k+1 is broadcast by readlane
// and applied to column k+1 (the next pivot needs it), and its entries for
// rows k+2 .. 15 are read back from LDS (the wave's LDS accesses complete in
// issue order, so the reads see the column) and applied to those columns in
// step k+1, beside that step's pivot chain, by when the reads have landed.
// A partial panel also updates the block's columns past cb (trailing
// columns of the front) by all of its columns, and stores them.  Every x_j
// receives its updates in column order, as in the plain loop.  The first
// non-positive pivot is recorded.
template <typename T, bool FULL>
__device__ __forceinline__ void panel16(Front64<T>& S, int b16, int cb, int n, int* info, int col0)
{
    const int r = threadIdx.x & 63, row = b16 + r;
    const bool live = row < n;
    T x[16], v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        x[j] = (live && (r >= 16 || j <= r)) ? S.A[b16 + j][row] : T(0);
        v[j] = T(0);
    }
    int bad = 16;
    T lp = T(0);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (FULL || k < cb) {  // no early exit: the loop stays unrolled, x and v in registers
            const T d = bcast(x[k], k);
            bad = (!(d > T(0)) && bad == 16) ? k : bad;
            const T rs = rsqrt_nr(d), sd = sqrt_from_rs(d, rs);
            // column k-1's updates of columns k+1 .. 15 (its entries read in step k-1)
#pragma unroll
            for (int j = k + 1; j < 16; ++j) x[j] -= lp * v[j];
            const T l = quot(x[k], sd, rs);
            x[k] = r == k ? sd : l;
            if (live) S.A[b16 + k][row] = x[k];
            if (r == 0) S.rs[b16 + k] = rs;
            if (k + 1 < 16 && (FULL || k + 1 < cb)) x[k + 1] -= l * bcast(l, k + 1);
            lp = l;
            __builtin_amdgcn_wave_barrier();  // the reads below stay after the store above
#pragma unroll
            for (int j = k + 2; j < 16; ++j) v[j] = S.A[b16 + k][b16 + j];
            if (!FULL && k + 1 == cb && k + 1 < 16) v[k + 1] = S.A[b16 + k][b16 + k + 1];
        }
    }
    if (!FULL) {
        if (r == 0)
            for (int k = cb; k < 16; ++k) S.rs[b16 + k] = T(1);  // finite stand-ins for the block's inverse
        // the last column's updates of the block's columns past cb
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (j >= cb) x[j] -= lp * v[j];
        if (live) {
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (j >= cb) S.A[b16 + j][row] = x[j];
        }
    }
    if (bad < 16 && r == 0) atomicMin(info, col0 + b16 + bad + 1);
}

// Inverse of panel b16's 16 x 16 diagonal block of L (one wave): lane r < 16
// solves L u = e_r by forward substitution, u = column r of the inverse,
// stored as X[b16 + r][b16 + k] = u_k.  For a partial panel the rows and
// columns past its cb pivots are finite stand-ins (rs = 1 there; the
// entries of L in them are trailing-block values), which no stored
// inverse entry uses: the inverse's row k < cb depends only on the rows
// and columns <= k.
template <typename T>
__device__ __forceinline__ void diaginv16(Front64<T>& S, int b16)
{
    const int r = threadIdx.x & 63;
    T y[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) y[j] = j == r ? T(1) : T(0);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const T y0 = y[k] * S.rs[b16 + k];
        y[k] = y0;
#pragma unroll
        for (int j = k + 1; j < 16; ++j) y[j] -= y0 * S.A[b16 + k][b16 + j];
    }
    if (r < 16) {
#pragma unroll
        for (int k = 0; k < 16; ++k) S.X[b16 + r][b16 + k] = y[k];
    }
}

// A_IJ -= L_Ib L_Jb^T over the cb pivot columns of panel b (one wave).
template <typename T>
__device__ __forceinline__ void trailing16(Front64<T>& S, int b16, int cb, int I, int J)
{
    typedef typename Mfma<T>::acc_t acc_t;
    const int lane = threadIdx.x & 63, li = lane & 15;
    acc_t acc = {0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int kk = 4 * q + (lane >> 4);
        const bool ok = kk < cb;
        const T a = ok ? S.A[b16 + kk][16 * I + li] : T(0);
        const T b = ok ? S.A[b16 + kk][16 * J + li] : T(0);
        acc = Mfma<T>::mma(a, b, acc);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) S.A[16 * J + li][16 * I + Mfma<T>::row(lane, v)] -= acc[v];
}

// X_ij = -inv(L_ii) sum_{k=j}^{i-1} L_ik X_kj (one wave; staging in Ts[w]).
template <typename T>
__device__ __forceinline__ void inverse16(Front64<T>& S, int i, int j, int w)
{
    typedef typename Mfma<T>::acc_t acc_t;
    const int lane = threadIdx.x & 63, li = lane & 15;
    acc_t acc = {0, 0, 0, 0};
    for (int k = j; k < i; ++k) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int kk = 4 * q + (lane >> 4);
            acc = Mfma<T>::mma(S.A[16 * k + kk][16 * i + li], S.X[16 * j + li][16 * k + kk], acc);
        }
    }
    wave_sync();
#pragma unroll
    for (int v = 0; v < 4; ++v) S.Ts[w][Mfma<T>::row(lane, v)][li] = acc[v];
    wave_sync();
    acc_t acc2 = {0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int m = 4 * q + (lane >> 4);
        acc2 = Mfma<T>::mma(S.X[16 * i + m][16 * i + li], S.Ts[w][m][li], acc2);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) S.X[16 * j + li][16 * i + Mfma<T>::row(lane, v)] = -acc2[v];
}

// All 256 threads: loads the lower n x n block at g (leading dimension ld),
// factors its first c columns, writes L and the Schur complement back and
// stores L^{-T} of the c x c factor column-major at inv (64 x 64,
// zero-padded: inv[r + 64 q] = L^{-1}[q][r]).  A pivot that is not positive
// records the global column col0 + k + 1 in *info (smallest wins).
// Same, with S.A already holding the block (lower part, zero elsewhere)
// and S.X zeroed, both visible to the workgroup.
template <typename T>
__device__ __forceinline__ void factor_dense64_loaded(Front64<T>& S, T* g, int64_t ld, int n, int c, int* info,
                                                      int col0, T* inv)
{
    const int tid = threadIdx.x, wv = tid >> 6;
    const int nblk = (c + 15) >> 4, nall = (n + 15) >> 4;
    // Phase b: wave 0 runs panel b's chain; wave 1 inverts panel b-1's
    // diagonal block; waves 2 and 3 form row b-2 of the inverse's
    // off-diagonal blocks (it needs the diagonal inverses up to b-2 and the
    // rows above it).  Then all four waves apply panel b to the trailing
    // blocks.  The last diagonal inverse and the last two rows follow the
    // loop.  Only the last panel can be partial.
    for (int b = 0; b < nblk; ++b) {
        const int b16 = 16 * b, cb = min(16, c - b16);
        if (wv == 0) {
            if (cb == 16)
                panel16<T, true>(S, b16, 16, n, info, col0);
            else
                panel16<T, false>(S, b16, cb, n, info, col0);
        } else if (wv == 1) {
            if (b >= 1) diaginv16<T>(S, b16 - 16);
        } else if (wv - 2 < b - 2) {
            inverse16<T>(S, b - 2, wv - 2, wv);
        }
        __syncthreads();
        int idx = 0;
        for (int J = b + 1; J < nall; ++J)
            for (int I = J; I < nall; ++I, ++idx)
                if ((idx & 3) == wv) trailing16<T>(S, b16, cb, I, J);
        __syncthreads();
    }
    const int last = nblk - 1;
    if (last >= 0) {
        if (wv == 1)
            diaginv16<T>(S, 16 * last);
        else if (wv >= 2 && wv - 2 < last - 1)
            inverse16<T>(S, last - 1, wv - 2, wv);
        __syncthreads();
        if (wv < last) inverse16<T>(S, last, wv, wv);
    }
    __syncthreads();
    for (int e = tid; e < 64 * 64; e += kThreads) {
        const int row = e & 63, col = e >> 6;
        if (row < n && col <= row) g[row + (int64_t)col * ld] = S.A[col][row];
        inv[e] = (row < c && col < c) ? S.X[row][col] : T(0);
    }
}
template <typename T>
__device__ __forceinline__ void factor_dense64(Front64<T>& S, T* g, int64_t ld, int n, int c, int* info, int col0,
                                               T* inv)
{
    // every load from a clamped in-bounds position, all issued before the
    // masked stores (no branch around each load)
    T v[16];
    const int row = threadIdx.x & 63, rc = min(row, n - 1);
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = ld_off(g, (unsigned)(rc + min((int)(threadIdx.x >> 6) + 4 * j, n - 1) * ld));
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int col = (threadIdx.x >> 6) + 4 * j;
        S.A[col][row] = (row < n && col <= row) ? v[j] : T(0);
        S.X[col][row] = T(0);
    }
    __syncthreads();
    factor_dense64_loaded<T>(S, g, ld, n, c, info, col0, inv);
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

// Zeroes the lower part (row >= col) of 64 columns of one front: task
// (node, first column).  Only the lower triangle of a front is ever read.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fill_lower(T* pool, const NodeDev* nodes, const int2* tasks)
{
    const int2 t = tasks[blockIdx.x];
    const NodeDev nd = nodes[t.x];
    const int n = nd.n, jend = min(t.y + 64, n);
    constexpr int V = 16 / sizeof(T);  // elements per 16-byte store
    T* g = pool + nd.front;
    for (int j = t.y; j < jend; ++j) {
        // rows [j, n) of column j: scalar head up to a 16-byte boundary,
        // 16-byte stores, scalar tail
        T* col = g + (int64_t)j * n;
        const int head = min((int)((V - ((reinterpret_cast<uintptr_t>(col + j) / sizeof(T)) % V)) % V), n - j);
        const int i0 = j + head, nv = (n - i0) / V, tail = i0 + nv * V;
        if ((int)threadIdx.x < head) col[j + threadIdx.x] = T(0);
        uint4* v = reinterpret_cast<uint4*>(col + i0);
        for (int p = threadIdx.x; p < nv; p += kThreads) v[p] = make_uint4(0u, 0u, 0u, 0u);
        if ((int)threadIdx.x < n - tail) col[tail + threadIdx.x] = T(0);
    }
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

// Diagonal scaling of the 16-bit forms: s[j] = 2^-e with e = floor(E / 2)
// for A(j,j) = m 2^E (0.5 <= m < 1), so that s_j^2 A(j,j) lies in [0.5, 2);
// a diagonal that is not positive and finite keeps s[j] = 1.  Powers of two
// make the scaled matrix S A S exact.  diag_src[j]: caller index of A(j,j)
// (j postordered).
__global__ __launch_bounds__(kThreads) void k_diag_scale(double* s, const double* a, const int64_t* diag_src, int n)
{
    const int j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= n) return;
    const double d = a[diag_src[j]];
    double v = 1.0;
    if (d > 0.0 && d <= 1.7976931348623157e308) {
        int E;
        (void)frexp(d, &E);
        const int e = E >= 0 ? E / 2 : -((1 - E) / 2);  // floor(E / 2)
        v = ldexp(1.0, -e);
    }
    s[j] = v;
}

// k_scatter_a for the scaled forms: pool[dst[k]] = s[r] s[c] a[src[k]] with
// (r, c) = rc[k], the entry's global (postordered, 0-based) row and column.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_scatter_a_scaled(T* pool, const double* a, const int64_t* dst,
                                                               const int64_t* src, const int2* rc, const double* s,
                                                               int64_t nnz)
{
    const int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (k >= nnz) return;
    const int2 q = rc[k];
    pool[dst[k]] = (T)(a[src[k]] * s[q.x] * s[q.y]);
}

// out[i + k n] = s[i] in[i + k n] (out may be in).
__global__ __launch_bounds__(kThreads) void k_scale_rows(double* out, const double* in, const double* s, int n, int nrhs)
{
    const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (e >= (int64_t)n * nrhs) return;
    out[e] = s[e % n] * in[e];
}

// ------------------------------------------------------------- extend-add
// One 64 x 64 tile (ti >= tj) of the child's update matrix, lower part only,
// added into the parent front at the rows/columns local_index gives (used
// for the children of order <= 64; the Schur tiles of larger children add
// their results themselves).  The map is injective within a child and one
// launch carries at most one child per parent, so the read-modify-write is
// race-free; a launch per child rank keeps the order of the additions fixed.
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
// Whole partial factorization of a front with n <= 64 in one workgroup: L in
// the first c columns, the Schur complement in the trailing block, and the
// inverse of the c x c diagonal block for the solves.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_small_front(T* pool, const NodeDev* nodes, const int* tasks, int* info)
{
    const NodeDev nd = nodes[tasks[blockIdx.x]];
    __shared__ Front64<T> S;
    factor_dense64<T>(S, pool + nd.front, nd.n, nd.n, nd.c, info, nd.first, pool + nd.inv);
}

// ------------------------------------------------------ large fronts
// LDS staging of one 64 x 64 MFMA tile product.
template <typename T> struct TileLds {
    T As[16][kTile + 4];
    T Bs[16][kTile + 4];
};

// sum_kk A[row][kk] B[col][kk] over a 64 x 64 tile (rows past m and columns
// past n give 0), FP64/FP32 MFMA 16x16x4.  Four waves each own a 32 x 32
// quarter (2 x 2 MFMA blocks c[i][j]: rows wr + 16 i, columns wc + 16 j).  K
// is staged through LDS in chunks of 16 with the next chunk's global loads
// issued before the current chunk's MFMAs.  Every read of A and B happens
// before the function returns (it ends with a barrier), so the caller may
// overwrite A afterwards.
//
// tb != 0 takes B transposed in memory (B[col][kk] at B + col * ldb + kk);
// its chunks are then read 16 entries at a time along kk.
//
// LP = kF16 / kBF16 (FP32 operands only) rounds each operand to 16 bits as
// it leaves LDS and multiplies on the 16-bit MFMA with FP32 accumulation.
// TB: 0 = B stored like A (B[col + kk * ldb]), 1 = B transposed in memory,
// -1 = as the runtime flag tb says (the panel TRSM kernel holds both).
template <typename T, int LP = kNative, int TB = -1>
__device__ __forceinline__ void tile_mma(const T* A, int lda, const T* B, int ldb, bool tb, int m, int n, int k,
                                         TileLds<T>& L, typename Mfma<T>::acc_t (&c)[2][2], int tid = threadIdx.x)
{
    static_assert(LP == kNative || sizeof(T) == 4, "16-bit operands come from FP32 fronts");
    typedef typename Mfma<T>::acc_t acc_t;
    const bool trb = TB < 0 ? tb : TB == 1;
    const int lane = tid & 63, wv = tid >> 6;
    const int wr = (wv & 1) * 32, wc = (wv >> 1) * 32;
    const int li = tid & 63, lp = tid >> 6;
    acc_t c00 = {0, 0, 0, 0}, c01 = {0, 0, 0, 0}, c10 = {0, 0, 0, 0}, c11 = {0, 0, 0, 0};
    T ra[4], rb[4];
    // transposed-B chunk: element e = tid + 256 p is (col e / 16, kk e % 16)
    const int tcol = tid >> 4, tkk = tid & 15;
    // Every load from a clamped in-bounds position (m, n, k >= 1), its value
    // kept raw and masked only as it is stored to LDS: no branch around the
    // loads, which all go out together.
    const int lic = min(li, m - 1), lib = min(li, n - 1);
    auto load = [&](int k0) {
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int kc = min(k0 + lp + 4 * p, k - 1);
            ra[p] = A[lic + (int64_t)kc * lda];
            if (trb)
                rb[p] = B[(int64_t)min(tcol + 16 * p, n - 1) * ldb + min(k0 + tkk, k - 1)];
            else
                rb[p] = B[lib + (int64_t)kc * ldb];
        }
    };
    load(0);
    for (int k0 = 0; k0 < k; k0 += 16) {
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int kk = k0 + lp + 4 * p;
            L.As[lp + 4 * p][li] = (li < m && kk < k) ? ra[p] : T(0);
            if (trb) {
                const int col = tcol + 16 * p;
                L.Bs[tkk][col] = (col < n && k0 + tkk < k) ? rb[p] : T(0);
            } else {
                L.Bs[lp + 4 * p][li] = (li < n && kk < k) ? rb[p] : T(0);
            }
        }
        __syncthreads();
        if (k0 + 16 < k) load(k0 + 16);
        if constexpr (LP == kNative) {
            // the operands of the chunk's four k-steps, then its 16 MFMAs (the
            // compiler interleaves the reads with the MFMAs)
            T a0[4], a1[4], b0[4], b1[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int kr = q * 4 + (lane >> 4);
                a0[q] = L.As[kr][wr + (lane & 15)];
                a1[q] = L.As[kr][wr + 16 + (lane & 15)];
                b0[q] = L.Bs[kr][wc + (lane & 15)];
                b1[q] = L.Bs[kr][wc + 16 + (lane & 15)];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                c00 = Mfma<T>::mma(a0[q], b0[q], c00);
                c01 = Mfma<T>::mma(a0[q], b1[q], c01);
                c10 = Mfma<T>::mma(a1[q], b0[q], c10);
                c11 = Mfma<T>::mma(a1[q], b1[q], c11);
            }
        } else {
            typedef Mfma16<LP> M16;
            typename M16::op_t a0, a1, b0, b1;
            const int kr = 4 * (lane >> 4);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a0[i] = M16::cvt(L.As[kr + i][wr + (lane & 15)]);
                a1[i] = M16::cvt(L.As[kr + i][wr + 16 + (lane & 15)]);
                b0[i] = M16::cvt(L.Bs[kr + i][wc + (lane & 15)]);
                b1[i] = M16::cvt(L.Bs[kr + i][wc + 16 + (lane & 15)]);
            }
            c00 = M16::mma(a0, b0, c00);
            c01 = M16::mma(a0, b1, c01);
            c10 = M16::mma(a1, b0, c10);
            c11 = M16::mma(a1, b1, c11);
        }
        __syncthreads();
    }
    c[0][0] = c00;
    c[0][1] = c01;
    c[1][0] = c10;
    c[1][1] = c11;
}
// The same product staged as Cs[col][row]; ends with a barrier.
template <typename T, int LP = kNative, int TB = -1>
__device__ __forceinline__ void tile_product(const T* A, int lda, const T* B, int ldb, bool tb, int m, int n, int k,
                                             TileLds<T>& L, T (*Cs)[kTile + 1])
{
    typename Mfma<T>::acc_t c[2][2];
    tile_mma<T, LP, TB>(A, lda, B, ldb, tb, m, n, k, L, c);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int wr = (wv & 1) * 32, wc = (wv >> 1) * 32, q0 = wc + (lane & 15);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = wr + Mfma<T>::row(lane, v);
        Cs[q0][r] = c[0][0][v];
        Cs[q0 + 16][r] = c[0][1][v];
        Cs[q0][r + 16] = c[1][0][v];
        Cs[q0 + 16][r + 16] = c[1][1][v];
    }
    __syncthreads();
}

// Products with operands already in LDS, same wave layout as tile_mma.
// Am[kk][row] holds A (k columns, zero from column k up to a multiple of
// 16), B comes from global memory like tile_mma's B (staged through L.Bs in
// chunks of 16).  Ends with a barrier.
template <typename T>
__device__ __forceinline__ void tile_mma_ldsA(const T (*Am)[kTile + 1], const T* B, int ldb, bool tb, int n, int k,
                                              TileLds<T>& L, typename Mfma<T>::acc_t (&c)[2][2])
{
    typedef typename Mfma<T>::acc_t acc_t;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = (wv & 1) * 32, wc = (wv >> 1) * 32;
    const int li = tid & 63, lp = tid >> 6, tcol = tid >> 4, tkk = tid & 15;
    acc_t c00 = {0, 0, 0, 0}, c01 = {0, 0, 0, 0}, c10 = {0, 0, 0, 0}, c11 = {0, 0, 0, 0};
    for (int k0 = 0; k0 < k; k0 += 16) {
        // loads from clamped in-bounds positions, all issued before the
        // masked stores (no branch around each load)
        T rb[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            if (tb)
                rb[p] = B[(int64_t)min(tcol + 16 * p, n - 1) * ldb + min(k0 + tkk, k - 1)];
            else
                rb[p] = B[min(li, n - 1) + (int64_t)min(k0 + lp + 4 * p, k - 1) * ldb];
        }
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            if (tb) {
                const int col = tcol + 16 * p, kt = k0 + tkk;
                L.Bs[tkk][col] = (col < n && kt < k) ? rb[p] : T(0);
            } else {
                const int kk = k0 + lp + 4 * p;
                L.Bs[lp + 4 * p][li] = (li < n && kk < k) ? rb[p] : T(0);
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int kr = q * 4 + (lane >> 4);
            const T a0 = Am[k0 + kr][wr + (lane & 15)], a1 = Am[k0 + kr][wr + 16 + (lane & 15)];
            const T b0 = L.Bs[kr][wc + (lane & 15)], b1 = L.Bs[kr][wc + 16 + (lane & 15)];
            c00 = Mfma<T>::mma(a0, b0, c00);
            c01 = Mfma<T>::mma(a0, b1, c01);
            c10 = Mfma<T>::mma(a1, b0, c10);
            c11 = Mfma<T>::mma(a1, b1, c11);
        }
        __syncthreads();
    }
    c[0][0] = c00;
    c[0][1] = c01;
    c[1][0] = c10;
    c[1][1] = c11;
}
// c += sum_kk Am[kk][row] Bm[kk][col] over kk < 64 with both operands in
// LDS (zero past the product's true K).  Same wave layout as tile_mma;
// fully unrolled so the accumulators stay in place; no barrier.
template <typename T>
__device__ __forceinline__ void lds_mma(const T (*Am)[kTile + 1], const T (*Bm)[kTile + 1],
                                        typename Mfma<T>::acc_t (&c)[2][2])
{
    typedef typename Mfma<T>::acc_t acc_t;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int wr = (wv & 1) * 32 + (lane & 15), wc = (wv >> 1) * 32 + (lane & 15), kl = lane >> 4;
    acc_t c00 = c[0][0], c01 = c[0][1], c10 = c[1][0], c11 = c[1][1];
#pragma unroll
    for (int k0 = 0; k0 < kTile; k0 += 4) {
        const T a0 = Am[k0 + kl][wr], a1 = Am[k0 + kl][wr + 16];
        const T b0 = Bm[k0 + kl][wc], b1 = Bm[k0 + kl][wc + 16];
        c00 = Mfma<T>::mma(a0, b0, c00);
        c01 = Mfma<T>::mma(a0, b1, c01);
        c10 = Mfma<T>::mma(a1, b0, c10);
        c11 = Mfma<T>::mma(a1, b1, c11);
    }
    c[0][0] = c00;
    c[0][1] = c01;
    c[1][0] = c10;
    c[1][1] = c11;
}
// Cs[col][row] = c (sub = false) or -= c (sub = true), in the layout of
// tile_mma's result.
template <typename T>
__device__ __forceinline__ void stage_acc(const typename Mfma<T>::acc_t (&c)[2][2], T (*Cs)[kTile + 1], bool sub)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int wr = (wv & 1) * 32, wc = (wv >> 1) * 32, q0 = wc + (lane & 15);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                T& dst = Cs[q0 + 16 * j][wr + 16 * i + Mfma<T>::row(lane, v)];
                dst = sub ? dst - c[i][j][v] : c[i][j][v];
            }
}

// One GemmTask per workgroup (see mfd_plan.h): C -= A B^T (mode 0: the
// DSYRK/DGEMM updates) or C = A B^T with B transposed in memory (mode 1: the
// panel TRSM as a product with the stored L^{-T} of the diagonal block, C
// and A the same rows).
// The result goes through LDS so the write of C is coalesced along columns.
// KIND only names the instance after the kernel class that launches it
// (TRSM, in-panel update, Schur update), so that profiles tell them apart.
// The result goes through LDS so the write of C is coalesced along columns
// (written straight from the accumulators it would be 32-byte pieces).
template <typename T, int LP = kNative, bool PRE = true, int TB = -1>
__device__ __forceinline__ void gemm_task(T* pool, const GemmTask& t, TileLds<T>& L, T (*Cs)[kTile + 1],
                                          const int* lidx = nullptr)
{
    T* C = pool + t.c;
    const int tid = threadIdx.x, row = tid & 63;
    // the entries of C this thread reads afterwards (row, column tid / 64 +
    // 4 j), loaded before the K loop so that their latency hides behind it
    // (from clamped positions, masked where used); PRE = false (the Schur
    // tiles, whose K loop is long) loads them after it
    T cpre[16];
    if (!PRE && t.mode != 1) {
        // loaded below, after the product
    } else if (t.mode != 1) {
        const int rc = min(row, t.m - 1);
#pragma unroll
        for (int j = 0; j < 16; ++j) cpre[j] = C[rc + (int64_t)min((tid >> 6) + 4 * j, t.n - 1) * t.ldc];
    }
    // mode 3 in the instances that load C after the K loop (the Schur
    // tiles): the parent positions of this thread's row and columns
    // (local_index), loaded before it so that the parent entries can be read
    // as soon as it ends (the others read them in the epilogue)
    int lrow = 1, lcol[16];
    if (!PRE && t.mode == 3) {
        lrow = lidx[t.li + min(row, t.m - 1)];
#pragma unroll
        for (int j = 0; j < 16; ++j) lcol[j] = lidx[t.lj + min((tid >> 6) + 4 * j, t.n - 1)];
    }
    tile_product<T, LP, TB>(pool + t.a, t.lda, pool + t.b, t.ldb, t.mode == 1, t.m, t.n, t.k, L, Cs);
    if (!PRE && t.mode != 1) {
        const int rc = min(row, t.m - 1);
#pragma unroll
        for (int j = 0; j < 16; ++j) cpre[j] = C[rc + (int64_t)min((tid >> 6) + 4 * j, t.n - 1) * t.ldc];
    }
    if (t.mode == 2) {
        // Cs = C - A B^T, then C = Cs D^T (D = L^{-1}, stored transposed)
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int col = (tid >> 6) + 4 * j;
            Cs[col][row] = (row < t.m && col < t.n) ? cpre[j] - Cs[col][row] : T(0);
        }
        __syncthreads();
        typename Mfma<T>::acc_t c[2][2];
        tile_mma_ldsA<T>(Cs, pool + t.d, kPanel, true, t.n, t.kd, L, c);
        stage_acc<T>(c, Cs, false);
        __syncthreads();
        if (row < t.m)
            for (int col = tid >> 6; col < t.n; col += 4) C[row + (int64_t)col * t.ldc] = Cs[col][row];
        return;
    }
    if (row >= t.m) return;
    if (t.mode == 3 || t.mode == 0) {
        // C -= the product; mode 3 also adds the new entry into the parent
        // (one launch holds at most one child of each parent, and local_index
        // is injective within a child: no two lanes add into one entry).
        // Four columns at a time, their loads (from clamped columns) all
        // issued before the stores, which may alias them as far as the
        // compiler knows.
        const bool ea = t.mode == 3;
        T* P = ea ? pool + t.p + ((PRE ? lidx[t.li + row] : lrow) - 1) : C;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int c0 = (tid >> 6) + 16 * k;
            if (c0 >= t.n) break;
            T pv[4];
            int64_t po[4];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                po[q] = ea ? (int64_t)((PRE ? lidx[t.lj + min(c0 + 4 * q, t.n - 1)] : lcol[4 * k + q]) - 1) * t.np : 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) pv[q] = ea ? P[po[q]] : T(0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int col = c0 + 4 * q;
                if (col < t.n && (!t.tri || row >= col)) {
                    const T v = cpre[4 * k + q] - Cs[col][row];
                    C[row + (int64_t)col * t.ldc] = v;
                    if (ea) P[po[q]] = pv[q] + v;
                }
            }
        }
    } else {
        for (int col = tid >> 6; col < t.n; col += 4) C[row + (int64_t)col * t.ldc] = Cs[col][row];
    }
}

// A TRSM tile with syrk > 0, after gemm_task: the diagonal block at cd
// loses R R^T over its lower syrk x syrk part, R = the tile's rows over the
// pending panels and the panel just solved (contiguous columns, A then C,
// which this workgroup has just written).
template <typename T>
__device__ __forceinline__ void trsm_syrk(T* pool, const GemmTask& t, TileLds<T>& L, T (*Cs)[kTile + 1])
{
    __syncthreads();
    const T* R = pool + (t.mode == 2 ? t.a : t.c);
    const int k = t.mode == 2 ? t.k + t.kd : t.k;
    tile_product<T, kNative, 0>(R, t.lda, R, t.lda, false, t.syrk, t.syrk, k, L, Cs);
    T* D = pool + t.cd;
    const int row = threadIdx.x & 63;
    if (row < t.syrk)
        for (int col = threadIdx.x >> 6; col <= row; col += 4) D[row + (int64_t)col * t.ldc] -= Cs[col][row];
}

// LP != kNative: the update and Schur tiles (KIND 1, 2) of the FP16_IR /
// BF16_IR forms multiply 16-bit operands.
template <typename T, int KIND, int LP = kNative>
__global__ __launch_bounds__(kThreads, KIND == 0 ? 3 : 4) void k_gemm(T* pool, const GemmTask* tasks, const int* lidx)
{
    static_assert(LP == kNative || KIND != 0, "the panel TRSM stays in the front's precision");
    if constexpr (KIND == 0) {
        // panel TRSM tiles (modes 1 and 2): mode 2 keeps its first product
        // in Cs while staging the inverse through L, so the two are separate
        __shared__ TileLds<T> L;
        __shared__ T Cs[kTile][kTile + 1];
        const GemmTask t = tasks[blockIdx.x];
        gemm_task<T>(pool, t, L, Cs);
        if (t.syrk > 0) trsm_syrk<T>(pool, t, L, Cs);
    } else {
        // modes 0 and 3 (B never transposed): the staged result Cs is written
        // after the last read of the operand staging L, so the two share LDS
        // (four workgroups per CU instead of three)
        __shared__ union {
            TileLds<T> L;
            T Cs[kTile][kTile + 1];
        } u;
        gemm_task<T, LP, KIND != 2, 0>(pool, tasks[blockIdx.x], u.L, u.Cs, lidx);
    }
}

// One launch per panel step s of the large fronts: GemmTasks from workgroup
// ndiag on (the panel TRSM of the rows after block s+1, preceded by the
// update of those rows by the panels still pending for block column s, and,
// where small, the periodic in-panel update of the column blocks after s+1)
// and, below ndiag, the factorization of block s+1.  A diagonal workgroup
// applies the pending panels and panel s to its own block itself, forming
// (and storing) its rows' panel-s entries on the way, so no task of the
// launch reads what another one writes: the chain of diagonal blocks, the
// critical path, advances one block per launch.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_upd_diag(T* pool, const NodeDev* nodes, const DiagTask* diags, int ndiag,
                                                       const GemmTask* tiles, int* info)
{
    __shared__ Front64<T> S;
    __shared__ union {
        TileLds<T> L;         // GEMM tasks
        T B[2][64][kTile + 1];  // diagonal task: staged operands
    } U;
    if ((int)blockIdx.x >= ndiag) {
        const GemmTask g = tiles[blockIdx.x - ndiag];
        gemm_task<T>(pool, g, U.L, S.A);
        if (g.syrk > 0) trsm_syrk<T>(pool, g, U.L, S.A);
        return;
    }
    const DiagTask t = diags[blockIdx.x];
    const NodeDev nd = nodes[t.node];
    const int n = nd.n, kb = t.kb, kp = t.kprev, kpp = t.kpp;
    T* front = pool + nd.front;
    const T* C = front + t.k0 + (int64_t)t.k0 * n;                         // block s+1
    T* Pg = front + t.k0 + (int64_t)(t.k0 - kp) * n;                       // its rows in panel s
    const T* Xd = front + t.k0 + (int64_t)(t.k0 - kp - kpp) * n;          // its rows in the pending panels
    const T* Xs = front + (t.k0 - kp) + (int64_t)(t.k0 - kp - kpp) * n;   // block s's rows in them
    const T* inv = pool + nd.inv + (int64_t)((t.k0 - kp) / kPanel) * kPanel * kPanel;  // L_s^{-T}
    const int tid = threadIdx.x, row = tid & 63, c4 = tid >> 6;
    // Stage every operand with all loads of a batch in flight: entry (row,
    // c4 + 4 j) of each 64 x 64 block; positions past the block read entry 0
    // and are masked after the loads.
    {
        T r0[16], r1[16];
        const unsigned ra = row < kb ? row : 0u, rb = row < kp ? row : 0u;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int col = c4 + 4 * j;
            const unsigned cd = col < kpp ? col : 0u;
            r0[j] = ld_off(Xd, ra + cd * (unsigned)n);
            r1[j] = ld_off(Xs, rb + cd * (unsigned)n);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int col = c4 + 4 * j;
            U.B[0][col][row] = (row < kb && col < kpp) ? r0[j] : T(0);
            U.B[1][col][row] = (row < kp && col < kpp) ? r1[j] : T(0);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int col = c4 + 4 * j;
            const unsigned ca = col <= row && row < kb ? col : 0u, cp = col < kp ? col : 0u;
            r0[j] = ld_off(C, ra + ca * (unsigned)n);
            r1[j] = ld_off((const T*)Pg, ra + cp * (unsigned)n);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int col = c4 + 4 * j;
            S.A[col][row] = (row < kb && col <= row) ? r0[j] : T(0);
            S.X[col][row] = (row < kb && col < kp) ? r1[j] : T(0);
        }
    }
    // inverse of block s in flight during the first products:
    // B[k][j] = L_s^{-1}[j][k] = inv[k + 64 j]
    T rv[16];
#pragma unroll
    for (int p = 0; p < 16; ++p) rv[p] = ld_off(inv, (unsigned)(tid + kThreads * p));
    __syncthreads();
    typedef typename Mfma<T>::acc_t acc_t;
    acc_t c[2][2] = {{{0, 0, 0, 0}, {0, 0, 0, 0}}, {{0, 0, 0, 0}, {0, 0, 0, 0}}};
    if (kpp > 0) {
        // the pending panels (64 columns per round): S.A -= Xd Xd^T (unless
        // the previous step's TRSM tile of these rows has applied it),
        // S.X -= Xd Xs^T
        acc_t c2[2][2] = {{{0, 0, 0, 0}, {0, 0, 0, 0}}, {{0, 0, 0, 0}, {0, 0, 0, 0}}};
        for (int q = 0; q < kpp; q += 64) {
            if (q > 0) {
                T r0[16], r1[16];
                const unsigned ra = row < kb ? row : 0u, rb = row < kp ? row : 0u;
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const int col = q + c4 + 4 * j;
                    const unsigned cd = col < kpp ? col : 0u;
                    r0[j] = ld_off(Xd, ra + cd * (unsigned)n);
                    r1[j] = ld_off(Xs, rb + cd * (unsigned)n);
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const int col = c4 + 4 * j;
                    U.B[0][col][row] = (row < kb && q + col < kpp) ? r0[j] : T(0);
                    U.B[1][col][row] = (row < kp && q + col < kpp) ? r1[j] : T(0);
                }
                __syncthreads();
            }
            if (!t.applied) lds_mma<T>(U.B[0], U.B[0], c);
            lds_mma<T>(U.B[0], U.B[1], c2);
        }
        if (!t.applied) stage_acc<T>(c, S.A, true);
        stage_acc<T>(c2, S.X, true);
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const int e = tid + kThreads * p, k = e & 63, j = e >> 6;
        U.B[0][k][j] = (k < kp && j < kp) ? rv[p] : T(0);
    }
    __syncthreads();
    // X = S.X L_s^{-T}, stored into the front (these rows are this task's)
    c[0][0] = c[0][1] = c[1][0] = c[1][1] = acc_t{0, 0, 0, 0};
    lds_mma<T>(S.X, U.B[0], c);
    __syncthreads();
    stage_acc<T>(c, S.X, false);
    __syncthreads();
    for (int col = c4; col < kp; col += 4)
        if (row < kb) Pg[row + (int64_t)col * n] = S.X[col][row];
    // panel s: S.A -= X X^T
    c[0][0] = c[0][1] = c[1][0] = c[1][1] = acc_t{0, 0, 0, 0};
    lds_mma<T>(S.X, S.X, c);
    stage_acc<T>(c, S.A, true);
    __syncthreads();
    for (int col = c4; col < kTile; col += 4) {
        S.X[col][row] = T(0);
        if (!(row < kb && col <= row)) S.A[col][row] = T(0);
    }
    __syncthreads();
    factor_dense64_loaded<T>(S, front + t.k0 + (int64_t)t.k0 * n, n, kb, kb, info, nd.first + t.k0,
                             pool + nd.inv + (int64_t)(t.k0 / kPanel) * kPanel * kPanel);
}

// C -= A B^T on 128 x 128 tiles (GemmTask with m, n <= 128, mode 0 or 3):
// the Schur update of fronts with many update rows (and their extend-add,
// mode 3).  Each of the four waves owns a 64 x 64 quarter as 4 x 4 MFMA
// 16x16x4 blocks, so every operand read from LDS feeds four MFMAs (half the
// LDS reads and half the L2 traffic per flop of the 64 x 64 tile).  K is
// staged through LDS in chunks of 16, double buffered: chunk j+1's global
// loads are issued before chunk j's MFMAs and stored to the other buffer
// after them, one barrier per chunk.  A diagonal tile (tri) skips the
// quarter above the diagonal.  The result goes back one quarter at a time
// through LDS (coalesced along columns); mode 3 also adds each new entry
// into the parent front at the rows and columns local_index gives (one
// launch holds at most one child of each parent).
template <typename T> struct WideLds {
    union {
        struct {
            T As[2][16][132];
            T Bs[2][16][132];
        } k;
        T Cs[64][65];
    };
};
template <typename T> using WideAcc = typename Mfma<T>::acc_t[4][4];

// acc = sum over K columns [k0, k1) of A B^T (a multiple of 16 apart, or k1
// = K); ends with a barrier.
template <typename T, int LP>
__device__ __forceinline__ void wide_product(const T* pool, const GemmTask& t, int k0, int k1, WideAcc<T>& acc,
                                             WideLds<T>& u)
{
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, li = lane & 15;
    const int wr = (wv & 1) * 64, wc = (wv >> 1) * 64;
    const bool skip = t.tri && wc > wr;  // quarter strictly above the diagonal
    const T* A = pool + t.a;
    const T* B = pool + t.b;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = typename Mfma<T>::acc_t{0, 0, 0, 0};
    const int lr = tid & 127, lk = tid >> 7;
    T ra[8], rb[8];
    auto load = [&](int kc) {
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            const int kk = kc + lk + 2 * p;
            ra[p] = (lr < t.m && kk < k1) ? A[lr + (int64_t)kk * t.lda] : T(0);
            rb[p] = (lr < t.n && kk < k1) ? B[lr + (int64_t)kk * t.ldb] : T(0);
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            u.k.As[buf][lk + 2 * p][lr] = ra[p];
            u.k.Bs[buf][lk + 2 * p][lr] = rb[p];
        }
    };
    load(k0);
    store(0);
    __syncthreads();
    const int nch = (k1 - k0 + 15) / 16;
    for (int j = 0; j < nch; ++j) {
        const int buf = j & 1;
        if (j + 1 < nch) load(k0 + 16 * (j + 1));
        if (!skip) {
            if constexpr (LP == kNative) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int kr = 4 * q + (lane >> 4);
                    T a[4], b[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        a[i] = u.k.As[buf][kr][wr + 16 * i + li];
                        b[i] = u.k.Bs[buf][kr][wc + 16 * i + li];
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int jj = 0; jj < 4; ++jj) acc[i][jj] = Mfma<T>::mma(a[i], b[jj], acc[i][jj]);
                }
            } else {
                typedef Mfma16<LP> M16;
                const int kr = 4 * (lane >> 4);
                typename M16::op_t a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        a[i][e] = M16::cvt(u.k.As[buf][kr + e][wr + 16 * i + li]);
                        b[i][e] = M16::cvt(u.k.Bs[buf][kr + e][wc + 16 * i + li]);
                    }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj) acc[i][jj] = M16::mma(a[i], b[jj], acc[i][jj]);
            }
        }
        if (j + 1 < nch) store(buf ^ 1);
        __syncthreads();
    }
}

// C -= acc (and, mode 3, each new entry added into the parent front), or,
// with part != nullptr, acc stored as partial product `part` (four 64 x 64
// quarters, column-major); both through LDS one quarter at a time; ends with
// a barrier.
template <typename T>
__device__ __forceinline__ void wide_apply(T* pool, const GemmTask& t, const int* lidx, const WideAcc<T>& acc,
                                           WideLds<T>& u, T* part)
{
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, li = lane & 15;
    T* C = pool + t.c;
    const bool ea = t.mode == 3;
    for (int w = 0; w < 4; ++w) {
        const int qr = (w & 1) * 64, qc = (w >> 1) * 64;
        if ((t.tri && qc > qr) || qr >= t.m || qc >= t.n) continue;  // uniform
        if (wv == w) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int jj = 0; jj < 4; ++jj)
#pragma unroll
                    for (int v = 0; v < 4; ++v) u.Cs[16 * jj + li][16 * i + Mfma<T>::row(lane, v)] = acc[i][jj][v];
        }
        __syncthreads();
        const int row = tid & 63, gr = qr + row;
        if (part) {
            T* Q = part + w * 4096 + row;
            for (int col = tid >> 6; col < 64; col += 4) Q[col * 64] = u.Cs[col][row];
        } else if (gr < t.m) {
            T* P = ea ? pool + t.p + (lidx[t.li + gr] - 1) : C;
            for (int c0 = tid >> 6; c0 < 64; c0 += 16) {
                T cv[4], pv[4];
                int64_t po[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int gc = min(qc + c0 + 4 * q, t.n - 1);
                    cv[q] = C[gr + (int64_t)gc * t.ldc];
                    po[q] = ea ? (int64_t)(lidx[t.lj + gc] - 1) * t.np : 0;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) pv[q] = ea ? P[po[q]] : T(0);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int col = c0 + 4 * q, gc = qc + col;
                    if (gc < t.n && (!t.tri || gr >= gc)) {
                        const T v = cv[q] - u.Cs[col][row];
                        C[gr + (int64_t)gc * t.ldc] = v;
                        if (ea) P[po[q]] = pv[q] + v;
                    }
                }
            }
        }
        __syncthreads();
    }
}

// The tiles' K loops divided evenly over the workgroups (every workgroup
// takes a run of 16-column K chunks, which may start and end inside a
// tile), so that a launch has no partly filled last round: workgroup w
// takes segments [seg_ptr[w], seg_ptr[w + 1]).  A segment covering its
// tile's whole K applies its product at once; the others store their
// partial products (slot: four 64 x 64 quarters, column-major, staged
// through LDS) for k_gemm_fix, which adds them in segment order and applies
// the sum.  Nothing waits on another workgroup.
template <typename T, int LP = kNative>
__global__ __launch_bounds__(kThreads, 2) void k_gemm_seg(T* pool, const GemmTask* tasks, const SegTask* segs,
                                                          const int* seg_ptr, T* part, const int* lidx)
{
    __shared__ WideLds<T> u;
    for (int s = seg_ptr[blockIdx.x]; s < seg_ptr[blockIdx.x + 1]; ++s) {
        const SegTask g = segs[s];
        const GemmTask t = tasks[g.task];
        WideAcc<T> acc;
        wide_product<T, LP>(pool, t, g.k0, g.k1, acc, u);
        wide_apply<T>(pool, t, lidx, acc, u, g.slot < 0 ? nullptr : part + (int64_t)g.slot * 16384);
    }
}

// Tiles whose K was split: C -= the partial products of slots slot0 ..
// slot0 + nslot - 1 added in that order (and, mode 3, each new entry added
// into the parent front), quarter by quarter like wide_apply.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_gemm_fix(T* pool, const GemmTask* tasks, const FixTask* fixes,
                                                       const T* part, const int* lidx)
{
    const FixTask f = fixes[blockIdx.x];
    const GemmTask t = tasks[f.task];
    const int tid = threadIdx.x, row = tid & 63;
    T* C = pool + t.c;
    const bool ea = t.mode == 3;
    for (int w = 0; w < 4; ++w) {
        const int qr = (w & 1) * 64, qc = (w >> 1) * 64, gr = qr + row;
        if ((t.tri && qc > qr) || qr >= t.m || qc >= t.n || gr >= t.m) continue;
        const T* Q = part + (int64_t)f.slot0 * 16384 + w * 4096 + row;
        T* P = ea ? pool + t.p + (lidx[t.li + gr] - 1) : C;
        for (int col = tid >> 6; col < 64; col += 4) {
            const int gc = qc + col;
            if (gc >= t.n || (t.tri && gr < gc)) continue;
            T sum = Q[col * 64];
            for (int sl = 1; sl < f.nslot; ++sl) sum += Q[(int64_t)sl * 16384 + col * 64];
            const T v = C[gr + (int64_t)gc * t.ldc] - sum;
            C[gr + (int64_t)gc * t.ldc] = v;
            if (ea) {
                T* pp = P + (int64_t)(lidx[t.lj + gc] - 1) * t.np;
                *pp = *pp + v;
            }
        }
    }
}

// Diagonal block [k0, k0+kb) of a large front, already updated by every
// earlier panel (DPOTRF("U"), factorize_kernel_m.f90:4-9): L written back
// into the front, and the inverse of L stored for the panel TRSM and the
// solves.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_diag(T* pool, const NodeDev* nodes, const DiagTask* tasks, int* info)
{
    const DiagTask t = tasks[blockIdx.x];
    const NodeDev nd = nodes[t.node];
    __shared__ Front64<T> S;
    factor_dense64<T>(S, pool + nd.front + t.k0 + (int64_t)t.k0 * nd.n, nd.n, t.kb, t.kb, info, nd.first + t.k0,
                      pool + nd.inv + (int64_t)(t.k0 / kPanel) * kPanel * kPanel);
}

// ---------------------------------------------------------- medium fronts
// The whole partial factorization of one front per workgroup, for fronts
// larger than one 64 x 64 LDS block whose 64-column steps would leave most
// of the chip idle (many such fronts share a level): panel by panel,
//   1. the kb x kb diagonal block factored and inverted in LDS
//      (factor_dense64: DPOTRF, factorize_kernel_m.f90:4-9), its inverse
//      stored for the solves;
//   2. the rows below it solved against that inverse, in place, one 64-row
//      MFMA tile at a time (the panel TRSM, factorize_kernel_m.f90:11-16);
//   3. the trailing lower triangle -- pivot columns and update matrix --
//      updated by the panel, one 64 x 64 MFMA tile at a time with K = kb
//      (DSYRK / DGEMM, factorize_kernel_m.f90:18-33).
// The workgroup owns its front, so workgroup barriers order the phases; no
// tile of a phase reads what another tile of it writes.  LP rounds the
// update tiles' operands to 16 bits (the FP16_IR / BF16_IR forms).  With
// clocks != nullptr, task i's start and end (wall_clock64, 100 MHz) go to
// clocks[2 i], clocks[2 i + 1]: the per-front timings of the schedule tuner.
// One tile of k_front_wg, kept out of line: the medium-front kernel's
// registers are then those of its diagonal phase (two waves per SIMD, the
// call's saved registers on the stack) instead of the sum of both phases.
template <typename T, int LP>
__device__ __attribute__((noinline)) void front_tile(T* pool, GemmTask g, TileLds<T>& L, T (*Cs)[kTile + 1])
{
    gemm_task<T, LP, false>(pool, g, L, Cs);
}
template <typename T, int LP>
__global__ __launch_bounds__(kThreads) void k_front_wg(T* pool, const NodeDev* nodes, const int* tasks, int* info,
                                                       long long* clocks)
{
    long long t_start = 0;
    if (clocks) t_start = wall_clock64();
    const NodeDev nd = nodes[tasks[blockIdx.x]];
    __shared__ union {
        Front64<T> D;
        struct {
            TileLds<T> L;
            T Cs[kTile][kTile + 1];
        } G;
    } U;
    const int n = nd.n, c = nd.c;
    for (int k0 = 0; k0 < c; k0 += kPanel) {
        const int kb = min(kPanel, c - k0);
        const int64_t inv = nd.inv + (int64_t)(k0 / kPanel) * kPanel * kPanel;
        factor_dense64<T>(U.D, pool + nd.front + k0 + (int64_t)k0 * n, n, kb, kb, info, nd.first + k0, pool + inv);
        __syncthreads();
        for (int r0 = k0 + kb; r0 < n; r0 += kTile) {
            GemmTask g{};
            g.c = g.a = nd.front + r0 + (int64_t)k0 * n;
            g.ldc = g.lda = n;
            g.b = inv;  // L^{-T}, read transposed
            g.ldb = kPanel;
            g.m = min(kTile, n - r0);
            g.n = g.k = kb;
            g.mode = 1;
            front_tile<T, kNative>(pool, g, U.G.L, U.G.Cs);
            __syncthreads();
        }
        for (int c0 = k0 + kb; c0 < n; c0 += kTile)
            for (int r0 = c0; r0 < n; r0 += kTile) {
                GemmTask g{};
                g.c = nd.front + r0 + (int64_t)c0 * n;
                g.a = nd.front + r0 + (int64_t)k0 * n;
                g.b = nd.front + c0 + (int64_t)k0 * n;
                g.ldc = g.lda = g.ldb = n;
                g.m = min(kTile, n - r0);
                g.n = min(kTile, n - c0);
                g.k = kb;
                g.tri = r0 == c0;
                front_tile<T, LP>(pool, g, U.G.L, U.G.Cs);
                __syncthreads();
            }
    }
    if (clocks && threadIdx.x == 0) {
        clocks[2 * blockIdx.x] = t_start;
        clocks[2 * blockIdx.x + 1] = wall_clock64();
    }
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

// LDS of the solve tasks, named so that a launch holding two kinds of task
// keeps them in one union.
struct FwdStepLds {
    union {
        struct {
            double LI[64][65];
            double Lp[64][65];  // Lp[k][i] = L[k0 + i][k0 - 64 + k], block task
        } b;
        double red[4][64][kRhsGroup + 1];  // chunk task: partial sums per column quarter
    } u;
    double zt[64][kRhsGroup], yt[64][kRhsGroup];
};
struct SuperLds {
    double zs[kSuper][kRhsGroup + 1], ys[kSuper][kRhsGroup + 1];  // the backward task's x_{u+1} in ys
    double red[16][16][kRhsGroup + 1];
};
// block and chunk tasks of the backward step use disjoint layouts (two
// workgroups per CU)
struct BwdChunkLds {
    double Ls[64][65];              // L columns, staged rows
    double Xs[256][kRhsGroup + 1];  // x of the chunk's rows
};
union BwdStepLds {
    struct {
        double LI[64][65];             // the diagonal block's inverse
        double Ln[64][65];             // Ln[k][i] = L[k0 + 64 + i][k0 + k]
        double zt[64][kRhsGroup];
        double Xn[64][kRhsGroup + 1];  // x of block t+1
    } b;
    BwdChunkLds c;
};

// Forward, first launch of a level, task (front, first row): 256 rows of
// the front's Z, pivot part from b, update part zero.
__global__ __launch_bounds__(kThreads) void k_fwd_init(double* zpool, const double* b, int ldb, int nrhs,
                                                       const NodeDev* nodes, const int2* tasks)
{
    const int2 t = tasks[blockIdx.x];
    const NodeDev nd = nodes[t.x];
    const int i = t.y + threadIdx.x;
    if (i >= nd.n) return;
    double* Z = zpool + nd.z * nrhs;
    for (int r = 0; r < nrhs; ++r)
        Z[i + (int64_t)r * nd.n] = i < nd.c ? b[(int64_t)(nd.first + i) + (int64_t)r * ldb] : 0.0;
}

// Forward, one launch per child rank: task (child, first row) adds 256 rows
// of the child's update part into its parent's Z (scatter_add,
// sparse_add_subroutines_m.f90:13-30).  A launch holds at most one child of
// each parent, and the ranks run in child order.
__global__ __launch_bounds__(kThreads) void k_fwd_add(double* zpool, int nrhs, const NodeDev* nodes, const int* lidx,
                                                      const int2* tasks)
{
    const int2 t = tasks[blockIdx.x];
    const NodeDev ch = nodes[t.x];
    const int i = t.y + threadIdx.x;
    if (i >= ch.w) return;
    const NodeDev pa = nodes[ch.parent];
    double* Zp = zpool + pa.z * nrhs;
    const double* Zc = zpool + ch.z * nrhs;
    const int p = lidx[ch.lidx + i] - 1;
    for (int r = 0; r < nrhs; ++r) Zp[p + (int64_t)r * pa.n] += Zc[(ch.c + i) + (int64_t)r * ch.n];
}

// Forward sweep, one launch per block step t (DTRSV "U","T" and DGEMV "T",
// solve_kernel_m.f90:4-23), pipelined one block behind:
//   - the block task (row0 < 0) solves block t: z = Z_t - L[B_t, B_{t-1}]
//     y_{t-1} (pad = order of block t-1; 0 for the first block), y_t =
//     inv(L_tt) z, written to Y_t;
//   - a chunk task (row0 >= 0) applies block k0/64 (= t - 1) to its 64
//     rows, all after block t: Z[rows] -= L[rows, B_{t-1}] y_{t-1} (four
//     threads per row, 16 columns each, summed through LDS).
// No task of a launch depends on another task of the same launch, and every
// global read of a task is issued before its first wait.
// Forward chunk task (row0 >= 0, pad0 != 1): block B = k0/64 applied to the
// 64 rows from row0 (all after the block the step solves), Z[rows] -=
// L[rows, B] y_B (yt: staging of y_B; red: the four column quarters' partial
// sums).
template <typename T>
__device__ __forceinline__ void fwd_chunk_task(const T* pool, double* zpool, int nrhs, const NodeDev* nodes,
                                               const SolveTask t, double (&yt)[64][kRhsGroup],
                                               double (&red)[4][64][kRhsGroup + 1])
{
    const NodeDev nd = nodes[t.node];
    const int n = nd.n, c = nd.c, kb = t.kb, tid = threadIdx.x, lane = tid & 63, ty = tid >> 6;
    double* Z = zpool + nd.z * nrhs;
    const double* Y = Z + (int64_t)n * nrhs;
    const T* Lb = pool + nd.front + (int64_t)t.k0 * n;
    // 64 rows; thread (row lane, quarter ty) takes columns 16 ty .. 16 ty + 15
    const int row = t.row0 + lane;
    const bool live = row < n;
    // Unconditional loads from clamped (in-bounds) positions, kept in the
    // raw element type (a conditional or converted load would wait for
    // the previous one) and not masked: past column kb they multiply the
    // zeros of yt, and rows past n are not stored, so the z and y loads
    // below go out without waiting for these.
    const int rowc = min(row, n - 1);
    T l[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) l[j] = ld_off(Lb, (unsigned)(rowc + min(16 * ty + j, kb - 1) * n));
    for (int r0 = 0; r0 < nrhs; r0 += kRhsGroup) {
        const int nr = min(kRhsGroup, nrhs - r0);
        // this thread stores right-hand sides r0 + ty and r0 + ty + 4; the
        // right-hand side of every load below is uniform over the wave (no
        // load for the group's missing ones), its row clamped in bounds
        const double z0 = ty < nr ? Z[rowc + (int64_t)(r0 + ty) * n] : 0.0;
        const double z1 = ty + 4 < nr ? Z[rowc + (int64_t)(r0 + ty + 4) * n] : 0.0;
        static_assert(64 * kRhsGroup == 2 * kThreads, "two staged entries per thread");
        double yv[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int e = tid + kThreads * h, k = e & 63, r = e >> 6;
            yv[h] = r < nr ? Y[t.k0 + min(k, kb - 1) + (int64_t)(r0 + r) * c] : 0.0;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int e = tid + kThreads * h, k = e & 63, r = e >> 6;
            yt[k][r] = (k < kb && r < nr) ? yv[h] : 0.0;
        }
        __syncthreads();
        double acc[kRhsGroup];
#pragma unroll
        for (int r = 0; r < kRhsGroup; ++r) acc[r] = 0.0;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const double lk = (double)l[j];
#pragma unroll
            for (int r = 0; r < kRhsGroup; ++r) acc[r] += lk * yt[16 * ty + j][r];
        }
#pragma unroll
        for (int r = 0; r < kRhsGroup; ++r) red[ty][lane][r] = acc[r];
        __syncthreads();
        if (live) {
            if (ty < nr)
                Z[row + (int64_t)(r0 + ty) * n] =
                    z0 - ((red[0][lane][ty] + red[1][lane][ty]) + (red[2][lane][ty] + red[3][lane][ty]));
            if (ty + 4 < nr)
                Z[row + (int64_t)(r0 + ty + 4) * n] =
                    z1 - ((red[0][lane][ty + 4] + red[1][lane][ty + 4]) +
                          (red[2][lane][ty + 4] + red[3][lane][ty + 4]));
        }
        __syncthreads();
    }
}

template <typename T>
__device__ __forceinline__ void fwd_step_task(const T* pool, double* zpool, int nrhs, const NodeDev* nodes,
                                              const SolveTask t, FwdStepLds& S)
{
    const NodeDev nd = nodes[t.node];
    auto& u = S.u;
    auto& zt = S.zt;
    auto& yt = S.yt;
    double(*LI)[65] = u.b.LI;
    double(*Lp)[65] = u.b.Lp;
    const int n = nd.n, c = nd.c, kb = t.kb, tid = threadIdx.x, lane = tid & 63, ty = tid >> 6;
    double* Z = zpool + nd.z * nrhs;
    double* Y = Z + (int64_t)n * nrhs;
    const T* Lb = pool + nd.front + (int64_t)t.k0 * n;
    if (t.pad0 == 1) {
        // a front of one pivot block (k0 = 0, kb = c): y = inv(L_00) z, then
        // its update rows 64 at a time, Z[rows] -= L[rows, 0:c] y
        const T* inv = pool + nd.inv;
        T li[16];
#pragma unroll
        for (int p = 0; p < 16; ++p) {
            const int e = tid + kThreads * p, i = e & 63, k = e >> 6;
            li[p] = ld_off(inv, (unsigned)((i < kb && k < kb) ? e : 0));
        }
#pragma unroll
        for (int p = 0; p < 16; ++p) {
            const int e = tid + kThreads * p, i = e & 63, k = e >> 6;
            li[p] = (i < kb && k < kb) ? li[p] : T(0);
        }
        for (int r0 = 0; r0 < nrhs; r0 += kRhsGroup) {
            const int nr = min(kRhsGroup, nrhs - r0);
#pragma unroll
            for (int p = 0; p < 16; ++p) {
                const int e = tid + kThreads * p;
                LI[e & 63][e >> 6] = (double)li[p];
            }
            double zv[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int e = tid + kThreads * h, k = e & 63, r = e >> 6;  // r: uniform over the wave
                zv[h] = r < nr ? Z[min(k, kb - 1) + (int64_t)(r0 + r) * n] : 0.0;
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int e = tid + kThreads * h, k = e & 63, r = e >> 6;
                zt[k][r] = (k < kb && r < nr) ? zv[h] : 0.0;
            }
            __syncthreads();
            for (int r = ty; r < kRhsGroup; r += 4) {
                double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int k = 0; k < 64; ++k) acc[k & 3] += LI[k][lane] * zt[k][r];
                const double y = (acc[0] + acc[1]) + (acc[2] + acc[3]);
                if (lane < kb && r < nr) Y[lane + (int64_t)(r0 + r) * c] = y;
                yt[lane][r] = (lane < kb && r < nr) ? y : 0.0;
            }
            __syncthreads();
            for (int rr = c; rr < n; rr += 64) {
                const int row = rr + lane;
                const bool live = row < n;
                const int rowc = min(row, n - 1);
                T l[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) l[j] = ld_off(Lb, (unsigned)(rowc + min(16 * ty + j, kb - 1) * n));
                // rows clamped in bounds (used only where in range); ty: the wave
                const double z0 = ty < nr ? Z[rowc + (int64_t)(r0 + ty) * n] : 0.0;
                const double z1 = ty + 4 < nr ? Z[rowc + (int64_t)(r0 + ty + 4) * n] : 0.0;
                double acc[kRhsGroup];
#pragma unroll
                for (int r = 0; r < kRhsGroup; ++r) acc[r] = 0.0;
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const double lk = (double)l[j];
#pragma unroll
                    for (int r = 0; r < kRhsGroup; ++r) acc[r] += lk * yt[16 * ty + j][r];
                }
#pragma unroll
                for (int r = 0; r < kRhsGroup; ++r) u.red[ty][lane][r] = acc[r];
                __syncthreads();
                if (live) {
                    if (ty < nr)
                        Z[row + (int64_t)(r0 + ty) * n] =
                            z0 - ((u.red[0][lane][ty] + u.red[1][lane][ty]) + (u.red[2][lane][ty] + u.red[3][lane][ty]));
                    if (ty + 4 < nr)
                        Z[row + (int64_t)(r0 + ty + 4) * n] =
                            z1 - ((u.red[0][lane][ty + 4] + u.red[1][lane][ty + 4]) +
                                  (u.red[2][lane][ty + 4] + u.red[3][lane][ty + 4]));
                }
                __syncthreads();
            }
        }
        return;
    }
    if (t.row0 >= 0) {
        fwd_chunk_task<T>(pool, zpool, nrhs, nodes, t, yt, u.red);
        return;
    }
    // block task
    const int kp = t.pad;
    const T* inv = pool + nd.inv + (int64_t)(t.k0 / 64) * 4096;
    // block (B_t, B_{t-1}): rows k0.., columns k0 - 64.. (only read when kp > 0)
    const T* Lq = kp > 0 ? pool + nd.front + t.k0 + (int64_t)(t.k0 - 64) * n : Lb;
    // entries outside the kb x kb block read entry 0 (one line for all of
    // them instead of the whole padded 64 x 64 block) and are zeroed below
    T li[16], lp[16];
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const int e = tid + kThreads * p, i = e & 63, k = e >> 6;
        li[p] = ld_off(inv, (unsigned)((i < kb && k < kb) ? e : 0));
        lp[p] = ld_off(Lq, (unsigned)(min(i, kb - 1) + min(k, max(kp, 1) - 1) * n));
    }
    // the first right-hand-side group's z and y, in flight with the above
    double zv[2], yv[2];
    // from clamped in-bounds positions (y of block t-1 from entry 0 when
    // there is none), masked as they are staged
    auto load_zy = [&](int r0) {
        const int nr = min(kRhsGroup, nrhs - r0);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int e = tid + kThreads * h, k = e & 63, r = e >> 6;  // r: uniform over the wave
            const int64_t rc = r0 + r;
            zv[h] = r < nr ? Z[t.k0 + min(k, kb - 1) + rc * n] : 0.0;
            yv[h] = (r < nr && kp > 0) ? Y[t.k0 - 64 + min(k, kp - 1) + rc * c] : 0.0;
        }
    };
    load_zy(0);
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const int e = tid + kThreads * p, i = e & 63, k = e >> 6;
        lp[p] = (i < kb && k < kp) ? lp[p] : T(0);
        li[p] = (i < kb && k < kb) ? li[p] : T(0);
    }
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const int e = tid + kThreads * p;
        LI[e & 63][e >> 6] = li[p];
        Lp[e >> 6][e & 63] = lp[p];
    }
    // Lp, LI, zt and yt are zero past kp / kb, so the products run over all
    // 64 entries, unrolled, with four partial sums.
    for (int r0 = 0; r0 < nrhs; r0 += kRhsGroup) {
        const int nr = min(kRhsGroup, nrhs - r0);
        if (r0 > 0) load_zy(r0);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int e = tid + kThreads * h, k = e & 63, r = e >> 6;
            zt[k][r] = (k < kb && r < nr) ? zv[h] : 0.0;
            yt[k][r] = (k < kp && r < nr) ? yv[h] : 0.0;
        }
        __syncthreads();
        for (int r = ty; r < nr; r += 4) {
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 64; ++k) acc[k & 3] += Lp[k][lane] * yt[k][r];
            zt[lane][r] -= (acc[0] + acc[1]) + (acc[2] + acc[3]);
        }
        __syncthreads();
        for (int r = ty; r < nr; r += 4) {
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 64; ++k) acc[k & 3] += LI[k][lane] * zt[k][r];
            if (lane < kb) Y[t.k0 + lane + (int64_t)(r0 + r) * c] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
        }
        __syncthreads();
    }
}
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fwd_step(const T* pool, double* zpool, int nrhs, const NodeDev* nodes,
                                                       const SolveTask* tasks)
{
    __shared__ FwdStepLds S;
    fwd_step_task<T>(pool, zpool, nrhs, nodes, tasks[blockIdx.x], S);
}

// A forward step's chunk tasks alone (steps with many of them): without the
// block task's LDS and registers several workgroups share a CU.
template <typename T>
__global__ __launch_bounds__(kThreads, 6) void k_fwd_chunk(const T* pool, double* zpool, int nrhs, const NodeDev* nodes,
                                                        const SolveTask* tasks)
{
    __shared__ double yt[64][kRhsGroup];
    __shared__ double red[4][64][kRhsGroup + 1];
    fwd_chunk_task<T>(pool, zpool, nrhs, nodes, tasks[blockIdx.x], yt, red);
}

// ------------------------------------------------------------ superblocks
// The three 256 x 256 matrices of superblock u, column-major, X at xoff, W
// at xoff + 256^2 (u >= 1), V at xoff + 2 256^2 (superblock u+1 not empty):
//   X = inverse of the superblock's diagonal block of L (zero past its
//       width sb), by 64 x 64 blocks (i, j), j <= i: X_ii is the stored
//       64 x 64 inverse (j = i; the row's blocks above the diagonal cleared,
//       the whole row past sb) and X_ij = -X_ii sum_{k=j}^{i-1} L_ik X_kj;
//   W = X L[B_u, B_{u-1}] (4 <= j < 8: block (i, j - 4)), the coupling of
//       the superblock to the previous one, so that y_u = X z_u - W y_{u-1};
//   V = L[B_{u+1}, B_u] X (j >= 8: block (i, j - 8), rows of superblock
//       u+1), so that x_u = X^T (y_u - q_u) - V^T x_{u+1}.
// One task per block, launched by block distance i - j; the W and V blocks
// follow in one launch of k_superwv.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_superinv(const T* pool, T* sinv, const NodeDev* nodes,
                                                       const SuperInvTask* tasks)
{
    const SuperInvTask t = tasks[blockIdx.x];
    const NodeDev nd = nodes[t.node];
    T* X = sinv + t.xoff;
    const int n = nd.n, c0 = kSuper * t.u, tid = threadIdx.x;
    const int i = t.i, j = t.j, b0 = c0 + 64 * i;  // first global pivot column of block row i
    const int kbi = max(0, min(64, nd.c - b0));     // pivot rows of block row i
    if (i == j) {
        const bool live = kbi > 0;
        const T* inv = pool + nd.inv + (int64_t)(live ? b0 / 64 : 0) * 4096;  // inv[r + 64 q] = L^{-1}(q, r)
        for (int e = tid; e < 4096; e += kThreads) {
            const int r = e & 63, q = e >> 6;  // block row q, column r
            X[(64 * i + q) + kSuper * (64 * i + r)] = (r < kbi && q < kbi) ? inv[r + 64 * q] : T(0);
            for (int jj = 0; jj < 4; ++jj)
                if (jj > i || (!live && jj < i)) X[(64 * i + q) + kSuper * (64 * jj + r)] = T(0);
        }
        return;
    }
    __shared__ T As[64][65], Bs[64][65];  // As[m][r] = A(r, m), Bs[col][m] = B(m, col)
    const int r = tid & 63, cq = tid >> 6;
    T acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = T(0);
    for (int k = j; k < i; ++k) {
        for (int e = tid; e < 4096; e += kThreads) {
            const int rr = e & 63, m = e >> 6;
            // A = L_ik, B = X_kj
            As[m][rr] = rr < kbi ? pool[nd.front + (b0 + rr) + (int64_t)(c0 + 64 * k + m) * n] : T(0);
            Bs[m][rr] = X[(64 * k + rr) + kSuper * (64 * j + m)];
        }
        __syncthreads();
#pragma unroll 4
        for (int m = 0; m < 64; ++m) {
            const T a = As[m][r];
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[q] += a * Bs[16 * cq + q][m];
        }
        __syncthreads();
    }
    // X_ij = -X_ii T1: As[m][rr] = X_ii(rr, m), Bs[col][m] = T1(m, col)
    for (int e = tid; e < 4096; e += kThreads) {
        const int rr = e & 63, m = e >> 6;
        As[m][rr] = X[(64 * i + rr) + kSuper * (64 * i + m)];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) Bs[16 * cq + q][r] = acc[q];
    __syncthreads();
    T out[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) out[q] = T(0);
#pragma unroll 4
    for (int m = 0; m < 64; ++m) {
        const T a = As[m][r];
#pragma unroll
        for (int q = 0; q < 16; ++q) out[q] += a * Bs[16 * cq + q][m];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) X[(64 * i + r) + kSuper * (64 * j + 16 * cq + q)] = -out[q];
}

// The W and V blocks of the superblock matrices (tasks with j >= 4, see
// k_superinv) as single MFMA tile products with K up to 256:
//   W block (i, j - 4) = X[block row i, :] L[B_u, B_{u-1} block j - 4]
//       (X is zero right of its diagonal block, so K = min(64 (i + 1), sb));
//   V block (i, j - 8) = L[B_{u+1} block i, B_u from block j - 8]
//       X[B_u from block j - 8, block j - 8] (X is zero above it).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_superwv(const T* pool, T* sinv, const NodeDev* nodes,
                                                      const SuperInvTask* tasks)
{
    const SuperInvTask t = tasks[blockIdx.x];
    const NodeDev nd = nodes[t.node];
    T* X = sinv + t.xoff;
    const int n = nd.n, c0 = kSuper * t.u, sb = min(kSuper, nd.c - c0), i = t.i;
    const T* F = pool + nd.front;
    __shared__ union {
        TileLds<T> L;
        T Cs[kTile][kTile + 1];
    } u;
    T* dst;
    if (t.j >= 8) {
        const int jv = t.j - 8, rv = c0 + kSuper + 64 * i;
        const int m = max(0, min(64, nd.c - rv));
        tile_product<T, kNative, 1>(F + rv + (int64_t)(c0 + 64 * jv) * n, n, X + kSuper * 64 * jv + 64 * jv, kSuper, true, m, 64,
                        kSuper - 64 * jv, u.L, u.Cs);
        dst = X + 2 * kSuper * kSuper + 64 * i + kSuper * 64 * jv;
    } else {
        const int jw = t.j - 4, cp = c0 - kSuper;
        tile_product<T, kNative, 1>(X + 64 * i, kSuper, F + c0 + (int64_t)(cp + 64 * jw) * n, n, true, 64, 64, min(64 * (i + 1), sb),
                        u.L, u.Cs);
        dst = X + kSuper * kSuper + 64 * i + kSuper * 64 * jw;
    }
    const int row = threadIdx.x & 63;
    for (int col = threadIdx.x >> 6; col < 64; col += 4) dst[row + kSuper * col] = u.Cs[col][row];
}

// Forward sweep over superblocks, one launch per superblock step u, every
// task 16 rows wide (thread: row tid % 16, columns 16 g .. 16 g + 15 of
// each 256-column operand, g = tid / 16; the 16 partial sums per row summed
// through LDS):
//   - top task (kind 1, rows row0 .. row0 + 15 of superblock u):
//     y_u = X z_u - W y_{u-1}, written to Y (z_u: Z at the superblock's
//     rows, already holding every superblock before u - 1);
//   - chunk task (kind 0, front rows row0 .. row0 + 15, all after
//     superblock u): Z[rows] -= L[rows, B_{u-1}] y_{u-1}.
// Every matrix load of a task is issued before its first wait.
template <typename T>
__device__ __forceinline__ void fwd_super_task(const T* pool, const T* sinv, double* zpool, int nrhs,
                                               const NodeDev* nodes, const SuperTask t, SuperLds& S)
{
    const NodeDev nd = nodes[t.node];
    const int n = nd.n, c = nd.c, c0 = kSuper * t.u, cp = c0 - kSuper, tid = threadIdx.x;
    const int i = tid & 15, g = tid >> 4;
    double* Z = zpool + nd.z * nrhs;
    double* Y = Z + (int64_t)n * nrhs;
    auto& zs = S.zs;
    auto& ys = S.ys;
    auto& red = S.red;
    const int pb = t.pb, sb = t.sb;
    const bool top = t.kind == 1;
    T a[16], w[16];
    if (top) {
        const T* X = sinv + t.xoff;
        const int r = t.row0 + i;
        // all loads inside the superblock's X and W (W of the first
        // superblock is unused storage), masked where used
#pragma unroll
        for (int q = 0; q < 16; ++q) a[q] = X[r + kSuper * (16 * g + q)];
#pragma unroll
        for (int q = 0; q < 16; ++q) w[q] = X[kSuper * kSuper + r + kSuper * (16 * g + q)];
    } else {
        const T* Lr = pool + nd.front + min(t.row0 + i, n - 1) + (int64_t)cp * n;
#pragma unroll
        for (int q = 0; q < 16; ++q) w[q] = Lr[(int64_t)min(16 * g + q, pb - 1) * n];
#pragma unroll
        for (int q = 0; q < 16; ++q) a[q] = T(0);
    }
    const bool lower = 16 * g <= t.row0 + 15;  // X is lower triangular
    for (int r0 = 0; r0 < nrhs; r0 += kRhsGroup) {
        const int nr = min(kRhsGroup, nrhs - r0);
        // y of superblock u-1 and z of superblock u from clamped positions
        // (entry 0 where there is none), all loads before the masked stores
        static_assert(kSuper == kThreads, "one superblock entry per thread");
        double yv[kRhsGroup], zv[kRhsGroup];
#pragma unroll
        for (int rr = 0; rr < kRhsGroup; ++rr) {
            // rr < nr, pb > 0 and top are uniform
            yv[rr] = (rr < nr && pb > 0) ? Y[cp + min(tid, pb - 1) + (int64_t)(r0 + rr) * c] : 0.0;
            zv[rr] = (rr < nr && top) ? Z[c0 + min(tid, sb - 1) + (int64_t)(r0 + rr) * n] : 0.0;
        }
#pragma unroll
        for (int rr = 0; rr < kRhsGroup; ++rr) {
            ys[tid][rr] = (tid < pb && rr < nr) ? yv[rr] : 0.0;
            if (top) zs[tid][rr] = (tid < sb && rr < nr) ? zv[rr] : 0.0;
        }
        __syncthreads();
        double acc[kRhsGroup];
#pragma unroll
        for (int rr = 0; rr < kRhsGroup; ++rr) acc[rr] = 0.0;
        if (top) {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const double av = lower ? (double)a[q] : 0.0;
#pragma unroll
                for (int rr = 0; rr < kRhsGroup; ++rr) acc[rr] += av * zs[16 * g + q][rr];
            }
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const double wv = 16 * g + q < pb ? (double)w[q] : 0.0;  // pb = 0: w unused
#pragma unroll
            for (int rr = 0; rr < kRhsGroup; ++rr) acc[rr] -= wv * ys[16 * g + q][rr];
        }
#pragma unroll
        for (int rr = 0; rr < kRhsGroup; ++rr) red[g][i][rr] = acc[rr];
        __syncthreads();
        if (tid < 16 * kRhsGroup) {
            const int i2 = tid & 15, rr = tid >> 4;
            double v = 0.0;
#pragma unroll
            for (int gg = 0; gg < 16; ++gg) v += red[gg][i2][rr];
            if (rr < nr) {
                if (top) {
                    if (t.row0 + i2 < sb) Y[c0 + t.row0 + i2 + (int64_t)(r0 + rr) * c] = v;
                } else if (t.row0 + i2 < n) {
                    Z[t.row0 + i2 + (int64_t)(r0 + rr) * n] += v;
                }
            }
        }
        __syncthreads();
    }
}
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fwd_super(const T* pool, const T* sinv, double* zpool, int nrhs,
                                                        const NodeDev* nodes, const SuperTask* tasks)
{
    __shared__ SuperLds S;
    fwd_super_task<T>(pool, sinv, zpool, nrhs, nodes, tasks[blockIdx.x], S);
}

// A forward step's superblock tasks (workgroups [0, nx)) and 64-column step
// tasks (the rest) in one launch: they touch different fronts.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fwd_pair(const T* pool, const T* sinv, double* zpool, int nrhs,
                                                       const NodeDev* nodes, const SuperTask* xtasks, int nx,
                                                       const SolveTask* tasks)
{
    __shared__ union {
        SuperLds x;
        FwdStepLds s;
    } u;
    if ((int)blockIdx.x < nx)
        fwd_super_task<T>(pool, sinv, zpool, nrhs, nodes, xtasks[blockIdx.x], u.x);
    else
        fwd_step_task<T>(pool, zpool, nrhs, nodes, tasks[blockIdx.x - nx], u.s);
}

// Backward sweep over superblocks, top tasks (kind 2), 16 rows of
// superblock u each, in the launch of backward step u (thread: row
// tid % 16, entries 16 g .. 16 g + 15 of each 256-long operand, g = tid /
// 16; the 16 partial sums per row summed through LDS):
//   x_u = X^T (y_u - q_u) - V^T x_{u+1},
// y_u from Y, q_u summed in a fixed order from the partial-sum slots that
// the previous launch's chunk tasks wrote (the rows after superblock u+1),
// x_{u+1} from Z (the pivot part of x of these fronts lives in Z, so that
// no task of a launch reads what another writes); x_u goes to Z and to x.
template <typename T>
__device__ __forceinline__ void bwd_super_task(const T* pool, const T* sinv, double* zpool, const double* part,
                                               double* x, int ldx, int nrhs, const NodeDev* nodes, const SuperTask t,
                                               SuperLds& S)
{
    const NodeDev nd = nodes[t.node];
    const int n = nd.n, c = nd.c, c0 = kSuper * t.u, cn = c0 + kSuper, tid = threadIdx.x;
    const int i = tid & 15, g = tid >> 4;
    double* Z = zpool + nd.z * nrhs;
    const double* Y = Z + (int64_t)n * nrhs;
    auto& zs = S.zs;
    auto& xs = S.ys;
    auto& red = S.red;
    const int pb = t.pb, sb = t.sb;
    const T* X = sinv + t.xoff;
    const T* V = X + 2 * kSuper * kSuper;
    const int col = t.row0 + i;  // column of X and V: the row of x_u
    T a[16], v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) a[q] = X[(16 * g + q) + kSuper * col];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = V[(16 * g + q) + kSuper * col];  // unused storage when pb = 0
    static_assert(kSuper == kThreads, "one superblock entry per thread");
    const int m = tid;  // the entry of y_u, q_u and x_{u+1} this thread stages
    // every load below from a clamped in-bounds position (x_{u+1} from entry
    // 0 when there is none), masked as it is staged
    const int mc = min(m, sb - 1);
    const double* pp = part + (int64_t)(t.slot + t.nchunk * (mc >> 6)) * 64 * nrhs + (mc & 63);
    const int64_t cs = (int64_t)64 * nrhs;  // stride between chunks
    for (int r0 = 0; r0 < nrhs; r0 += kRhsGroup) {
        const int nr = min(kRhsGroup, nrhs - r0);
        // the chunks' partial sums two at a time, every load of a pair (all
        // right-hand sides of the group) issued before the subtractions
        double zv[kRhsGroup], xv[kRhsGroup];
#pragma unroll
        for (int rr = 0; rr < kRhsGroup; ++rr) {
            // rr < nr and pb > 0 are uniform
            zv[rr] = rr < nr ? Y[c0 + mc + (int64_t)(r0 + rr) * c] : 0.0;
            xv[rr] = (rr < nr && pb > 0) ? Z[cn + min(m, pb - 1) + (int64_t)(r0 + rr) * n] : 0.0;
        }
        for (int ch = 0; ch < t.nchunk; ch += 2) {
            const int ch2 = min(ch + 1, t.nchunk - 1);  // the last chunk twice when odd
            double p0[kRhsGroup], p1[kRhsGroup];
#pragma unroll
            for (int rr = 0; rr < kRhsGroup; ++rr) {
                const double* q = pp + (int64_t)(r0 + rr) * 64;
                p0[rr] = rr < nr ? q[ch * cs] : 0.0;
                p1[rr] = rr < nr ? q[ch2 * cs] : 0.0;
            }
            const bool two = ch + 1 < t.nchunk;
#pragma unroll
            for (int rr = 0; rr < kRhsGroup; ++rr) zv[rr] = (zv[rr] - p0[rr]) - (two ? p1[rr] : 0.0);
        }
#pragma unroll
        for (int rr = 0; rr < kRhsGroup; ++rr) {
            zs[m][rr] = (rr < nr && m < sb) ? zv[rr] : 0.0;
            xs[m][rr] = (rr < nr && m < pb) ? xv[rr] : 0.0;
        }
        __syncthreads();
        double acc[kRhsGroup];
#pragma unroll
        for (int rr = 0; rr < kRhsGroup; ++rr) acc[rr] = 0.0;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const double av = (double)a[q], vv = pb > 0 ? (double)v[q] : 0.0;
#pragma unroll
            for (int rr = 0; rr < kRhsGroup; ++rr) acc[rr] += av * zs[16 * g + q][rr] - vv * xs[16 * g + q][rr];
        }
#pragma unroll
        for (int rr = 0; rr < kRhsGroup; ++rr) red[g][i][rr] = acc[rr];
        __syncthreads();
        if (tid < 16 * kRhsGroup) {
            const int i2 = tid & 15, rr = tid >> 4, row = t.row0 + i2;
            double s = 0.0;
#pragma unroll
            for (int gg = 0; gg < 16; ++gg) s += red[gg][i2][rr];
            if (rr < nr && row < sb) {
                Z[c0 + row + (int64_t)(r0 + rr) * n] = s;
                x[(int64_t)(nd.first + c0 + row) + (int64_t)(r0 + rr) * ldx] = s;
            }
        }
        __syncthreads();
    }
}
template <typename T>
__global__ __launch_bounds__(kThreads) void k_bwd_super(const T* pool, const T* sinv, double* zpool, const double* part,
                                                        double* x, int ldx, int nrhs, const NodeDev* nodes,
                                                        const SuperTask* tasks)
{
    __shared__ SuperLds S;
    bwd_super_task<T>(pool, sinv, zpool, part, x, ldx, nrhs, nodes, tasks[blockIdx.x], S);
}

// Backward, first launch of a level, task (front, first row): 256 rows of
// the update part of Z take the parent's solution at the parent-front
// positions -- pivot positions from the parent's Y, update positions from
// the parent's Z (gather, sparse_add_subroutines_m.f90:32-49).
__global__ __launch_bounds__(kThreads) void k_bwd_gather(double* zpool, int nrhs, const NodeDev* nodes, const int* lidx,
                                                         const int2* tasks)
{
    const int2 t = tasks[blockIdx.x];
    const NodeDev nd = nodes[t.x];
    const int i = t.y + threadIdx.x;
    if (nd.parent < 0 || i >= nd.w) return;
    const NodeDev pa = nodes[nd.parent];
    double* Z = zpool + nd.z * nrhs;
    const double* Zp = zpool + pa.z * nrhs;
    const double* Yp = Zp + (int64_t)pa.n * nrhs;
    const int p = lidx[nd.lidx + i] - 1;
    for (int r = 0; r < nrhs; ++r)
        Z[(nd.c + i) + (int64_t)r * nd.n] = (p < pa.c && !pa.xz) ? Yp[p + (int64_t)r * pa.c] : Zp[p + (int64_t)r * pa.n];
}

// Backward sweep, one launch per block step t (DGEMV "N" + DTRSV "U","N",
// solve_kernel_m.f90:11-16,25-30), pipelined one block ahead:
//   - the block task (row0 < 0) solves block t:
//       z = Y_t - sum of the nchunk partials at part[slot..] (rows after
//           block t+1, formed by the previous launch) - L[B_{t+1}, B_t]^T x_{t+1}
//       x_t = inv(L_tt)^T z, written to Y_t (and, for t = 0, the front's
//       pivot part of x);  pad = kb of block t+1 (0: no such block);
//   - a chunk task (row0 >= 0) forms, for block k0/64 (= t - 1), the partial
//     products P[slot][k][r] = sum over its 256 rows of L[row][k0+k] x[row][r]
//     for its 16 of the block's columns (pad0 = which 16; four tasks per
//     chunk) or, with pad0 < 0, for all 64 (launches with many chunks) on
//     MFMA (x from Y on pivot rows, from Z on update rows; every row lies
//     after block t, so its x is final).
// No task of a launch depends on another task of the same launch, and the
// partials are summed in a fixed order.
// Backward chunk task (row0 >= 0): the partial sums of block k0/64 (all 64
// columns, pad0 < 0, or columns 16 pad0 .. 16 pad0 + 15) over the 256 rows
// from row0, into slot `slot` of part.
template <typename T>
__device__ __forceinline__ void bwd_chunk_task(const T* pool, const double* zpool, double* part, int nrhs,
                                               const NodeDev* nodes, const SolveTask t, BwdChunkLds& S)
{
    const NodeDev nd = nodes[t.node];
    const int n = nd.n, c = nd.c, kb = t.kb, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const double* Z = zpool + nd.z * nrhs;
    const double* Y = Z + (int64_t)n * nrhs;
    const T* Lb = pool + nd.front + (int64_t)t.k0 * n;
    if (t.row0 >= 0 && t.pad0 < 0) {
        // all 64 columns in one task: wave w takes columns 16 w .. 16 w + 15,
        // the 256 rows staged 64 at a time through LDS
        typedef typename Mfma<double>::acc_t acc_t;
        const int row = t.row0 + tid;
        const bool live = row < n;
        // Unconditional loads from clamped (in-bounds) positions, kept in the
        // raw element type (a conditional or converted load would wait for
        // the previous one) and not masked: rows past n meet the zero x of
        // Xs, and partials past column kb are never read, so the x loads
        // below go out without waiting for these.
        const int rowc = min(row, n - 1);
        T l[64];
#pragma unroll
        for (int k = 0; k < 64; ++k) l[k] = ld_off(Lb, (unsigned)(rowc + min(k, kb - 1) * n));
        for (int r0 = 0; r0 < nrhs; r0 += kRhsGroup) {
            const int nr = min(kRhsGroup, nrhs - r0);
#pragma unroll
            for (int r = 0; r < kRhsGroup; ++r)
                S.Xs[tid][r] = (live && r < nr) ? ((row < c && !nd.xz) ? Y[row + (int64_t)(r0 + r) * c]
                                                                          : Z[row + (int64_t)(r0 + r) * n])
                                              : 0.0;
            acc_t acc = {0, 0, 0, 0};
            const int rc = lane & 15;
            for (int sb = 0; sb < 4; ++sb) {
                __syncthreads();
                if (wv == sb) {
#pragma unroll
                    for (int k = 0; k < 64; ++k) S.Ls[k][lane] = l[k];
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int rr = 4 * q + (lane >> 4);
                    const double a = S.Ls[16 * wv + rc][rr];
                    const double b = rc < nr ? S.Xs[64 * sb + rr][rc] : 0.0;
                    acc = Mfma<double>::mma(a, b, acc);
                }
            }
            if (rc < nr) {
                double* P = part + (int64_t)t.slot * 64 * nrhs + (int64_t)(r0 + rc) * 64;
#pragma unroll
                for (int v = 0; v < 4; ++v) P[16 * wv + Mfma<double>::row(lane, v)] = acc[v];
            }
            __syncthreads();
        }
        return;
    }
    if (t.row0 >= 0) {
        // columns 16 pad0 .. 16 pad0 + 15 of block k0/64 over the chunk's 256
        // rows: one row per thread, a quarter of the rows per wave, the four
        // waves' 16 x 16 products added through LDS
        typedef typename Mfma<double>::acc_t acc_t;
        const int col0 = 16 * t.pad0, row = t.row0 + tid;
        const bool live = row < n;
        // Unconditional loads from clamped (in-bounds) positions, kept in the
        // raw element type (a conditional or converted load would wait for
        // the previous one) and not masked: rows past n meet the zero x of
        // Xs, and partials past column kb are never read, so the x loads
        // below go out without waiting for these.
        const int rowc = min(row, n - 1);
        T l[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) l[j] = ld_off(Lb, (unsigned)(rowc + min(col0 + j, kb - 1) * n));
        double(*Lc)[kThreads + 4] = reinterpret_cast<double(*)[kThreads + 4]>(&S.Ls[0][0]);   // 16 x 260
        double(*red)[16][17] = reinterpret_cast<double(*)[16][17]>(&S.Ls[0][0]);              // 4 x 16 x 17
        const int rc = lane & 15;
        for (int r0 = 0; r0 < nrhs; r0 += kRhsGroup) {
            const int nr = min(kRhsGroup, nrhs - r0);
#pragma unroll
            for (int r = 0; r < kRhsGroup; ++r)
                S.Xs[tid][r] = (live && r < nr) ? ((row < c && !nd.xz) ? Y[row + (int64_t)(r0 + r) * c]
                                                                          : Z[row + (int64_t)(r0 + r) * n])
                                              : 0.0;
#pragma unroll
            for (int j = 0; j < 16; ++j) Lc[j][tid] = (double)l[j];
            __syncthreads();
            acc_t acc = {0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int rr = 64 * wv + 4 * q + (lane >> 4);
                const double a = Lc[rc][rr];
                const double b = rc < nr ? S.Xs[rr][rc] : 0.0;
                acc = Mfma<double>::mma(a, b, acc);
            }
            __syncthreads();
#pragma unroll
            for (int v = 0; v < 4; ++v) red[wv][Mfma<double>::row(lane, v)][rc] = acc[v];
            __syncthreads();
            if (tid < 16 * nr) {
                const int m = tid & 15, r = tid >> 4;
                part[(int64_t)t.slot * 64 * nrhs + (int64_t)(r0 + r) * 64 + col0 + m] =
                    (red[0][m][r] + red[1][m][r]) + (red[2][m][r] + red[3][m][r]);
            }
            __syncthreads();
        }
        return;
    }
}

template <typename T>
__device__ __forceinline__ void bwd_step_task(const T* pool, double* zpool, double* part, double* x, int ldx, int nrhs,
                                              const NodeDev* nodes, const SolveTask t, BwdStepLds& u)
{
    const NodeDev nd = nodes[t.node];
    const int n = nd.n, c = nd.c, kb = t.kb, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double* Z = zpool + nd.z * nrhs;
    double* Y = Z + (int64_t)n * nrhs;
    const T* Lb = pool + nd.front + (int64_t)t.k0 * n;
    if (t.row0 >= 0) {
        bwd_chunk_task<T>(pool, zpool, part, nrhs, nodes, t, u.c);
        return;
    }
    // block task: every global read first
    const int kn = t.pad;  // rows of block t+1 (0: none)
    const T* inv = pool + nd.inv + (int64_t)(t.k0 / 64) * 4096;
    // entries outside the kb x kb block read entry 0 and are zeroed below
    T li[16], ln[16];
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const int e = tid + kThreads * p, i = e & 63, k = e >> 6;
        li[p] = ld_off(inv, (unsigned)((i < kb && k < kb) ? e : 0));
        // (row k0 itself when there is no next block: always inside the front)
        ln[p] = ld_off(Lb, (unsigned)(t.k0 + (kn > 0 ? 64 + min(i, kn - 1) : 0) + min(k, kb - 1) * n));
    }
    // z = y_t - the partial sums, and x of block t+1, for one group of
    // right-hand sides (two entries of each per thread); the first group's
    // loads go out with the factor loads above
    double zv[2], xv[2];
    // (from clamped in-bounds positions, x of block t+1 from entry 0 when
    // there is none; masked as they are staged)
    auto load_zx = [&](int r0) {
        const int nr = min(kRhsGroup, nrhs - r0);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int e = tid + kThreads * h, k = e & 63, r = e >> 6, kc = min(k, kb - 1);  // r: uniform over the wave
            const int64_t rc = r0 + r;
            double z = 0.0;
            if (r < nr) {
                z = Y[t.k0 + kc + rc * c];
                const double* pp = part + (int64_t)t.slot * 64 * nrhs + kc + rc * 64;
#pragma unroll 4
                for (int ch = 0; ch < t.nchunk; ++ch) z -= pp[(int64_t)ch * 64 * nrhs];
            }
            zv[h] = z;
            xv[h] = (r < nr && kn > 0) ? Y[t.k0 + 64 + min(k, kn - 1) + rc * c] : 0.0;
        }
    };
    load_zx(0);
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const int e = tid + kThreads * p, i = e & 63, k = e >> 6;
        ln[p] = (i < kn && k < kb) ? ln[p] : T(0);
        li[p] = (i < kb && k < kb) ? li[p] : T(0);
    }
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const int e = tid + kThreads * p;
        u.b.LI[e & 63][e >> 6] = li[p];
        u.b.Ln[e >> 6][e & 63] = ln[p];
    }
    for (int r0 = 0; r0 < nrhs; r0 += kRhsGroup) {
        const int nr = min(kRhsGroup, nrhs - r0);
        if (r0 > 0) load_zx(r0);
        // zt, and x of block t+1 staged in u.b.Xn[0..kn)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int e = tid + kThreads * h, k = e & 63, r = e >> 6;
            u.b.zt[k][r] = (k < kb && r < nr) ? zv[h] : 0.0;
            u.b.Xn[k][r] = (k < kn && r < nr) ? xv[h] : 0.0;
        }
        __syncthreads();
        // z_k -= sum_i L[k0 + 64 + i][k0 + k] x_{t+1}[i]; Ln, LI, Xn and zt
        // are zero past kn / kb, so the sums run over all 64 entries,
        // unrolled, with four partial sums
        for (int r = wv; r < nr; r += 4) {
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int i = 0; i < 64; ++i) acc[i & 3] += u.b.Ln[lane][i] * u.b.Xn[i][r];
            if (lane < kb) u.b.zt[lane][r] -= (acc[0] + acc[1]) + (acc[2] + acc[3]);
        }
        __syncthreads();
        for (int r = wv; r < nr; r += 4) {
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int m = 0; m < 64; ++m) acc[m & 3] += u.b.LI[lane][m] * u.b.zt[m][r];
            const double v = (acc[0] + acc[1]) + (acc[2] + acc[3]);
            if (lane < kb) {
                Y[t.k0 + lane + (int64_t)(r0 + r) * c] = v;
                if (t.k0 == 0) x[(int64_t)(nd.first + lane) + (int64_t)(r0 + r) * ldx] = v;
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
template <typename T>
__global__ __launch_bounds__(kThreads) void k_bwd_step(const T* pool, double* zpool, double* part, double* x, int ldx,
                                                       int nrhs, const NodeDev* nodes, const SolveTask* tasks)
{
    __shared__ BwdStepLds u;
    bwd_step_task<T>(pool, zpool, part, x, ldx, nrhs, nodes, tasks[blockIdx.x], u);
}

// A backward step's chunk tasks alone (steps with many of them): without the
// block task's LDS and registers more workgroups share a CU.
template <typename T>
__global__ __launch_bounds__(kThreads, 3) void k_bwd_chunk(const T* pool, const double* zpool, double* part, int nrhs,
                                                        const NodeDev* nodes, const SolveTask* tasks)
{
    __shared__ BwdChunkLds S;
    bwd_chunk_task<T>(pool, zpool, part, nrhs, nodes, tasks[blockIdx.x], S);
}

// A backward step's superblock top tasks (workgroups [0, nx)) and its step
// tasks (the rest) in one launch: the top tasks of superblock t read the
// partial sums of the bank the step tasks do not write and x of superblock
// t + 1, and write x of superblock t, which no step task of the launch reads.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_bwd_pair(const T* pool, const T* sinv, double* zpool, double* part,
                                                       double* x, int ldx, int nrhs, const NodeDev* nodes,
                                                       const SuperTask* xtasks, int nx, const SolveTask* tasks)
{
    __shared__ union {
        SuperLds x;
        BwdStepLds s;
    } u;
    if ((int)blockIdx.x < nx)
        bwd_super_task<T>(pool, sinv, zpool, part, x, ldx, nrhs, nodes, xtasks[blockIdx.x], u.x);
    else
        bwd_step_task<T>(pool, zpool, part, x, ldx, nrhs, nodes, tasks[blockIdx.x - nx], u.s);
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
// r = b - A x over the full symmetric matrix in CSR order (val: the values
// in that order, gathered once per factorization), four right-hand sides per
// pass over a row; each sum runs over the row's entries in order.
__global__ __launch_bounds__(kThreads) void k_residual(double* r, const double* b, const double* x, int n, int nrhs,
                                                       const int* ptr, const int* col, const double* val)
{
    const int row = blockIdx.x * kThreads + threadIdx.x;
    if (row >= n) return;
    const int e0 = ptr[row], e1 = ptr[row + 1];
    for (int k0 = 0; k0 < nrhs; k0 += 4) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int e = e0; e < e1; ++e) {
            const double v = val[e];
            const double* xc = x + col[e] + (int64_t)k0 * n;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (k0 + q < nrhs) acc[q] += v * xc[(int64_t)q * n];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (k0 + q < nrhs) r[row + (int64_t)(k0 + q) * n] = b[row + (int64_t)(k0 + q) * n] - acc[q];
    }
}

// val[e] = a[src[e]]: the caller's values in CSR order
__global__ __launch_bounds__(kThreads) void k_gather_vals(double* val, const double* a, const int64_t* src, int64_t nnz)
{
    const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (e < nnz) val[e] = a[src[e]];
}

// Maximum over the workgroup, then one atomic per workgroup (as IEEE bits:
// order-preserving for non-negative doubles, so the result is exact).
__device__ __forceinline__ void block_max_to(unsigned long long* out, double m)
{
    __shared__ double wmax[kThreads / 64];
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kThreads / 64; ++w) m = fmax(m, wmax[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
    }
}

// The norms a backward error needs, in one pass, per column k (blockIdx.y)
// of the n x nrhs vectors r, x, b and (if not null) d, ld n: the maxima of
// |r|, |x|, |b|, |d| into out[k], out[nrhs + k], out[2 nrhs + k], out[3 nrhs
// + k] (IEEE bits, one atomic per workgroup: order-preserving for
// non-negative doubles, so exact), and the sums of squares of r and x as
// per-workgroup partials part[k parts + w], part[(nrhs + k) parts + w] (parts
// = gridDim.x), which the caller adds in order, so the sums do not depend on
// the timing of the workgroups.
__global__ __launch_bounds__(kThreads) void k_err_norms(unsigned long long* out, double* part, const double* r,
                                                        const double* x, const double* b, const double* d, int64_t n,
                                                        int nrhs)
{
    __shared__ double red[6][kThreads / 64];
    const int k = blockIdx.y;
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // max |r|, |x|, |b|, |d|; sum r^2, x^2
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const int64_t e = i + (int64_t)k * n;
        const double rv = r[e], xv = x[e], bv = b[e], dv = d ? d[e] : 0.0;
        v[0] = fmax(v[0], fabs(rv));
        v[1] = fmax(v[1], fabs(xv));
        v[2] = fmax(v[2], fabs(bv));
        v[3] = fmax(v[3], fabs(dv));
        v[4] += rv * rv;
        v[5] += xv * xv;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const double o = __shfl_xor(v[q], off);
            v[q] = q < 4 ? fmax(v[q], o) : v[q] + o;
        }
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < 6; ++q) red[q][threadIdx.x >> 6] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kThreads / 64; ++w)
#pragma unroll
            for (int q = 0; q < 6; ++q) v[q] = q < 4 ? fmax(v[q], red[q][w]) : v[q] + red[q][w];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q < 3 || d) atomicMax(out + q * nrhs + k, (unsigned long long)__double_as_longlong(v[q]));
        part[(int64_t)k * gridDim.x + blockIdx.x] = v[4];
        part[(int64_t)(nrhs + k) * gridDim.x + blockIdx.x] = v[5];
    }
}

// out[0] = max over rows of sum_e |val[e]|  (the infinity norm of A)
__global__ __launch_bounds__(kThreads) void k_norm_inf(unsigned long long* out, int n, const int* ptr,
                                                       const double* val)
{
    double m = 0.0;
    for (int row = blockIdx.x * kThreads + threadIdx.x; row < n; row += gridDim.x * kThreads) {
        double s = 0.0;
        for (int e = ptr[row]; e < ptr[row + 1]; ++e) s += fabs(val[e]);
        m = fmax(m, s);
    }
    block_max_to(out, m);
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
                     const double* val, hipStream_t st)
{
    if (n > 0)
        hipLaunchKernelGGL(k_residual, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, st, r, b, x, n, nrhs, ptr,
                           col, val);
}
void launch_gather_vals(double* val, const double* a, const int64_t* src, int64_t nnz, hipStream_t st)
{
    if (nnz > 0)
        hipLaunchKernelGGL(k_gather_vals, dim3((unsigned)((nnz + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, val, a,
                           src, nnz);
}
// at least 8 entries per thread: few workgroups per column, so few atomics
// on each maximum and a short read-back of the partial sums
int sumsq_parts(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(n / (8 * kThreads), kSumParts)); }
void launch_err_norms(unsigned long long* out, double* part, const double* r, const double* x, const double* b,
                      const double* d, int64_t n, int nrhs, hipStream_t st)
{
    if (n > 0)
        hipLaunchKernelGGL(k_err_norms, dim3(sumsq_parts(n), nrhs), dim3(kThreads), 0, st, out, part, r, x, b, d, n, nrhs);
}
void launch_norm_inf(unsigned long long* out, int n, const int* ptr, const double* val, hipStream_t st)
{
    const int g = std::min((n + kThreads - 1) / kThreads, 4096);
    if (n > 0) hipLaunchKernelGGL(k_norm_inf, dim3(g), dim3(kThreads), 0, st, out, n, ptr, val);
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
void launch_fill_lower(T* pool, const NodeDev* nodes, const int2* tasks, int count, hipStream_t st)
{
    if (count > 0) hipLaunchKernelGGL(k_fill_lower<T>, dim3(count), dim3(kThreads), 0, st, pool, nodes, tasks);
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
void launch_front_wg(T* pool, const NodeDev* nodes, const int* tasks, int count, int* info, long long* clocks,
                     hipStream_t st, int lowp)
{
    if (count <= 0) return;
    if constexpr (sizeof(T) == 4) {
        if (lowp == kF16) {
            hipLaunchKernelGGL((k_front_wg<T, kF16>), dim3(count), dim3(kThreads), 0, st, pool, nodes, tasks, info, clocks);
            return;
        }
        if (lowp == kBF16) {
            hipLaunchKernelGGL((k_front_wg<T, kBF16>), dim3(count), dim3(kThreads), 0, st, pool, nodes, tasks, info,
                               clocks);
            return;
        }
    }
    hipLaunchKernelGGL((k_front_wg<T, kNative>), dim3(count), dim3(kThreads), 0, st, pool, nodes, tasks, info, clocks);
}
template <typename T>
void launch_diag(T* pool, const NodeDev* nodes, const DiagTask* tasks, int count, int* info, hipStream_t st)
{
    if (count > 0) hipLaunchKernelGGL(k_diag<T>, dim3(count), dim3(kThreads), 0, st, pool, nodes, tasks, info);
}
template <typename T>
void launch_upd_diag(T* pool, const NodeDev* nodes, const DiagTask* diags, int ndiag, const GemmTask* tiles, int ntiles,
                     int* info, hipStream_t st)
{
    if (ndiag + ntiles > 0)
        hipLaunchKernelGGL(k_upd_diag<T>, dim3(ndiag + ntiles), dim3(kThreads), 0, st, pool, nodes, diags, ndiag, tiles,
                           info);
}
template <typename T>
void launch_gemm(T* pool, const GemmTask* tasks, int count, int kind, const int* lidx, hipStream_t st, int lowp)
{
    if (count <= 0) return;
    if constexpr (sizeof(T) == 4) {
        if (lowp == kF16 && kind != 0) {
            if (kind == 1)
                hipLaunchKernelGGL((k_gemm<T, 1, kF16>), dim3(count), dim3(kThreads), 0, st, pool, tasks, lidx);
            else
                hipLaunchKernelGGL((k_gemm<T, 2, kF16>), dim3(count), dim3(kThreads), 0, st, pool, tasks, lidx);
            return;
        }
        if (lowp == kBF16 && kind != 0) {
            if (kind == 1)
                hipLaunchKernelGGL((k_gemm<T, 1, kBF16>), dim3(count), dim3(kThreads), 0, st, pool, tasks, lidx);
            else
                hipLaunchKernelGGL((k_gemm<T, 2, kBF16>), dim3(count), dim3(kThreads), 0, st, pool, tasks, lidx);
            return;
        }
    }
    if (kind == 0)
        hipLaunchKernelGGL((k_gemm<T, 0>), dim3(count), dim3(kThreads), 0, st, pool, tasks, lidx);
    else if (kind == 1)
        hipLaunchKernelGGL((k_gemm<T, 1>), dim3(count), dim3(kThreads), 0, st, pool, tasks, lidx);
    else
        hipLaunchKernelGGL((k_gemm<T, 2>), dim3(count), dim3(kThreads), 0, st, pool, tasks, lidx);
}
void launch_diag_scale(double* s, const double* a, const int64_t* diag_src, int n, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(k_diag_scale, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, st, s, a, diag_src, n);
}
template <typename T>
void launch_scatter_a_scaled(T* pool, const double* a, const int64_t* dst, const int64_t* src, const int2* rc,
                             const double* s, int64_t nnz, hipStream_t st)
{
    if (nnz > 0)
        hipLaunchKernelGGL(k_scatter_a_scaled<T>, dim3((unsigned)((nnz + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                           st, pool, a, dst, src, rc, s, nnz);
}
void launch_scale_rows(double* out, const double* in, const double* s, int n, int nrhs, hipStream_t st)
{
    const int64_t m = (int64_t)n * nrhs;
    if (m > 0)
        hipLaunchKernelGGL(k_scale_rows, dim3((unsigned)((m + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, out, in,
                           s, n, nrhs);
}
template <typename T>
void launch_gemm_seg(T* pool, const GemmTask* tasks, const SegTask* segs, const int* seg_ptr, int groups,
                     const FixTask* fixes, int nfix, T* part, const int* lidx, hipStream_t st, int lowp)
{
    if (groups <= 0) return;
    if constexpr (sizeof(T) == 4) {
        if (lowp == kF16)
            hipLaunchKernelGGL((k_gemm_seg<T, kF16>), dim3(groups), dim3(kThreads), 0, st, pool, tasks, segs, seg_ptr,
                               part, lidx);
        else if (lowp == kBF16)
            hipLaunchKernelGGL((k_gemm_seg<T, kBF16>), dim3(groups), dim3(kThreads), 0, st, pool, tasks, segs, seg_ptr,
                               part, lidx);
        else
            hipLaunchKernelGGL((k_gemm_seg<T, kNative>), dim3(groups), dim3(kThreads), 0, st, pool, tasks, segs,
                               seg_ptr, part, lidx);
    } else {
        hipLaunchKernelGGL((k_gemm_seg<T, kNative>), dim3(groups), dim3(kThreads), 0, st, pool, tasks, segs, seg_ptr,
                           part, lidx);
    }
    if (nfix > 0)
        hipLaunchKernelGGL(k_gemm_fix<T>, dim3(nfix), dim3(kThreads), 0, st, pool, tasks, fixes, part, lidx);
}
void launch_fwd_init(double* zpool, const double* b, int ldb, int nrhs, const NodeDev* nodes, const int2* tasks,
                     int count, hipStream_t st)
{
    if (count > 0) hipLaunchKernelGGL(k_fwd_init, dim3(count), dim3(kThreads), 0, st, zpool, b, ldb, nrhs, nodes, tasks);
}
void launch_fwd_add(double* zpool, int nrhs, const NodeDev* nodes, const int* lidx, const int2* tasks, int count,
                    hipStream_t st)
{
    if (count > 0) hipLaunchKernelGGL(k_fwd_add, dim3(count), dim3(kThreads), 0, st, zpool, nrhs, nodes, lidx, tasks);
}
template <typename T>
void launch_fwd_step(const T* pool, double* zpool, int nrhs, const NodeDev* nodes, const SolveTask* tasks, int count,
                     hipStream_t st)
{
    if (count > 0) hipLaunchKernelGGL(k_fwd_step<T>, dim3(count), dim3(kThreads), 0, st, pool, zpool, nrhs, nodes, tasks);
}
template <typename T>
void launch_fwd_chunk(const T* pool, double* zpool, int nrhs, const NodeDev* nodes, const SolveTask* tasks, int count,
                      hipStream_t st)
{
    if (count > 0)
        hipLaunchKernelGGL(k_fwd_chunk<T>, dim3(count), dim3(kThreads), 0, st, pool, zpool, nrhs, nodes, tasks);
}
template <typename T>
void launch_bwd_chunk(const T* pool, const double* zpool, double* part, int nrhs, const NodeDev* nodes,
                      const SolveTask* tasks, int count, hipStream_t st)
{
    if (count > 0)
        hipLaunchKernelGGL(k_bwd_chunk<T>, dim3(count), dim3(kThreads), 0, st, pool, zpool, part, nrhs, nodes, tasks);
}
template <typename T>
void launch_superinv(const T* pool, T* sinv, const NodeDev* nodes, const SuperInvTask* tasks, int count, bool wv,
                     hipStream_t st)
{
    if (count <= 0) return;
    if (wv)
        hipLaunchKernelGGL(k_superwv<T>, dim3(count), dim3(kThreads), 0, st, pool, sinv, nodes, tasks);
    else
        hipLaunchKernelGGL(k_superinv<T>, dim3(count), dim3(kThreads), 0, st, pool, sinv, nodes, tasks);
}
template <typename T>
void launch_fwd_super(const T* pool, const T* sinv, double* zpool, int nrhs, const NodeDev* nodes,
                      const SuperTask* tasks, int count, hipStream_t st)
{
    if (count > 0)
        hipLaunchKernelGGL(k_fwd_super<T>, dim3(count), dim3(kThreads), 0, st, pool, sinv, zpool, nrhs, nodes, tasks);
}
template <typename T>
void launch_bwd_super(const T* pool, const T* sinv, double* zpool, const double* part, double* x, int ldx, int nrhs,
                      const NodeDev* nodes, const SuperTask* tasks, int count, hipStream_t st)
{
    if (count > 0)
        hipLaunchKernelGGL(k_bwd_super<T>, dim3(count), dim3(kThreads), 0, st, pool, sinv, zpool, part, x, ldx, nrhs,
                           nodes, tasks);
}
template <typename T>
void launch_fwd_pair(const T* pool, const T* sinv, double* zpool, int nrhs, const NodeDev* nodes,
                     const SuperTask* xtasks, int nx, const SolveTask* tasks, int count, hipStream_t st)
{
    if (nx + count > 0)
        hipLaunchKernelGGL(k_fwd_pair<T>, dim3(nx + count), dim3(kThreads), 0, st, pool, sinv, zpool, nrhs, nodes, xtasks,
                           nx, tasks);
}
template <typename T>
void launch_bwd_pair(const T* pool, const T* sinv, double* zpool, double* part, double* x, int ldx, int nrhs,
                     const NodeDev* nodes, const SuperTask* xtasks, int nx, const SolveTask* tasks, int count,
                     hipStream_t st)
{
    if (nx + count > 0)
        hipLaunchKernelGGL(k_bwd_pair<T>, dim3(nx + count), dim3(kThreads), 0, st, pool, sinv, zpool, part, x, ldx, nrhs,
                           nodes, xtasks, nx, tasks);
}
void launch_bwd_gather(double* zpool, int nrhs, const NodeDev* nodes, const int* lidx, const int2* tasks, int count,
                       hipStream_t st)
{
    if (count > 0) hipLaunchKernelGGL(k_bwd_gather, dim3(count), dim3(kThreads), 0, st, zpool, nrhs, nodes, lidx, tasks);
}
template <typename T>
void launch_bwd_step(const T* pool, double* zpool, double* part, double* x, int ldx, int nrhs, const NodeDev* nodes,
                     const SolveTask* tasks, int count, hipStream_t st)
{
    if (count > 0)
        hipLaunchKernelGGL(k_bwd_step<T>, dim3(count), dim3(kThreads), 0, st, pool, zpool, part, x, ldx, nrhs, nodes,
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
    template void launch_fill_lower<T>(T*, const NodeDev*, const int2*, int, hipStream_t);                            \
    template void launch_scatter_a<T>(T*, const double*, const int64_t*, const int64_t*, int64_t, hipStream_t);        \
    template void launch_extend_add<T>(T*, const int*, const NodeDev*, const ExtTask*, int, hipStream_t);              \
    template void launch_small_front<T>(T*, const NodeDev*, const int*, int, int*, hipStream_t);                       \
    template void launch_front_wg<T>(T*, const NodeDev*, const int*, int, int*, long long*, hipStream_t, int);           \
    template void launch_diag<T>(T*, const NodeDev*, const DiagTask*, int, int*, hipStream_t);                        \
    template void launch_upd_diag<T>(T*, const NodeDev*, const DiagTask*, int, const GemmTask*, int, int*, hipStream_t); \
    template void launch_gemm<T>(T*, const GemmTask*, int, int, const int*, hipStream_t, int);                         \
    template void launch_scatter_a_scaled<T>(T*, const double*, const int64_t*, const int64_t*, const int2*,           \
                                             const double*, int64_t, hipStream_t);                                    \
    template void launch_fwd_step<T>(const T*, double*, int, const NodeDev*, const SolveTask*, int, hipStream_t);       \
    template void launch_fwd_chunk<T>(const T*, double*, int, const NodeDev*, const SolveTask*, int, hipStream_t);      \
    template void launch_bwd_chunk<T>(const T*, const double*, double*, int, const NodeDev*, const SolveTask*, int,    \
                                      hipStream_t);                                                                    \
    template void launch_superinv<T>(const T*, T*, const NodeDev*, const SuperInvTask*, int, bool, hipStream_t);      \
    template void launch_fwd_super<T>(const T*, const T*, double*, int, const NodeDev*, const SuperTask*, int,         \
                                      hipStream_t);                                                                    \
    template void launch_gemm_seg<T>(T*, const GemmTask*, const SegTask*, const int*, int, const FixTask*, int, T*,  \
                                     const int*, hipStream_t, int);                                                    \
    template void launch_bwd_super<T>(const T*, const T*, double*, const double*, double*, int, int, const NodeDev*,    \
                                      const SuperTask*, int, hipStream_t);                                             \
    template void launch_fwd_pair<T>(const T*, const T*, double*, int, const NodeDev*, const SuperTask*, int,          \
                                     const SolveTask*, int, hipStream_t);                                              \
    template void launch_bwd_pair<T>(const T*, const T*, double*, double*, double*, int, int, const NodeDev*,          \
                                     const SuperTask*, int, const SolveTask*, int, hipStream_t);                       \
    template void launch_bwd_step<T>(const T*, double*, double*, double*, int, int, const NodeDev*, const SolveTask*,  \
                                     int, hipStream_t);
MFD_INSTANTIATE(double)
MFD_INSTANTIATE(float)

}  // namespace mfd
