/*
 * mfd_oracle.c -- CPU restatement of the numeric phase of the reference
 * supernodal multifrontal Cholesky (to-mo-ki/multifrontal-DAG), kept in the
 * reference's own tile storage so that every array can be compared entry by
 * entry.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load it,
 * and only as the checker.  The product path (multifrontal-dag_amd/) never
 * links, loads or calls it.
 *
 * Indexing follows the reference: nodes, tiles and in-tile positions are
 * 1-based in every formula below; element (r, q) of a column-major (ld, *)
 * array is stored at [(q-1)*ld + (r-1)].  Sizes and offsets are int64 (the
 * reference uses int32 intra-node offsets, integer_function_m.f90:14-30).
 *
 * Tile BLAS are textbook loops (no BLAS/LAPACK is installed in the image);
 * their calling conventions are exactly the reference's calls:
 *   mydpotrf  DPOTRF("U")              factorize_kernel_m.f90:4-9
 *   mydtrsm   DTRSM("L","U","T","N")   factorize_kernel_m.f90:11-16
 *   mydsyrk   DSYRK("U","T", -1, 1)    factorize_kernel_m.f90:18-23
 *   mydgemm   DGEMM("T","N", -1, 1)    factorize_kernel_m.f90:25-33
 *   mydtrsv_l DTRSV("U","T","N")       solve_kernel_m.f90:4-9
 *   mydtrsv_u DTRSV("U","N","N")       solve_kernel_m.f90:11-16
 *   mydgemv_t DGEMV("T", -1, 1)        solve_kernel_m.f90:18-23
 *   mydgemv_n DGEMV("N", -1, 1)        solve_kernel_m.f90:25-30
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

/* ---------------------------------------------------------------- integers */
/* integer_function_m.f90:14-30 */
static i64 psum1(i64 e) { return e < 1 ? 0 : e * (e + 1) / 2; }
static i64 psum2(i64 s, i64 e) { return e < s ? 0 : psum1(e) - psum1(s - 1); }
/* integer_function_m.f90:45-48: Fortran integer division truncates toward 0 */
static int div_ceiling(int a, int p) { return (a - 1) / p + 1; }
static int mod2(int a, int p) { return (a - 1) % p + 1; }
/* block_size_calculator_m.f90:12-21 */
static int block_size(int idx, int nb, int n) { return (i64)idx * nb > n ? n % nb : nb; }
/* get_block_size2, block_size_calculator_m.f90:23-36: a first block of
   first_block entries (0: none), then blocks of nb */
static int block_size2(int idx, int nb, int n, int first_block)
{
    if (first_block == 0) return block_size(idx, nb, n);
    return idx == 1 ? first_block : block_size(idx - 1, nb, n - first_block);
}

/* ------------------------------------------------------------ tile kernels */
#define CM(a, ld, r, q) (a)[((i64)(q) - 1) * (ld) + ((r) - 1)]

/* A = U^T U on the upper triangle of a column-major n x n array (lda = n). */
void orc_dpotrf_u(int n, double *a, int *info)
{
    *info = 0;
    for (int j = 1; j <= n; ++j) {
        double s = CM(a, n, j, j);
        for (int k = 1; k < j; ++k) s -= CM(a, n, k, j) * CM(a, n, k, j);
        if (!(s > 0.0)) { *info = j; return; }
        double d = sqrt(s);
        CM(a, n, j, j) = d;
        for (int i = j + 1; i <= n; ++i) {
            double t = CM(a, n, j, i);
            for (int k = 1; k < j; ++k) t -= CM(a, n, k, j) * CM(a, n, k, i);
            CM(a, n, j, i) = t / d;
        }
    }
}

/* B(m x nrhs, ldb = m) <- U^{-T} B with U upper (m x m, lda = m). */
void orc_dtrsm_lutn(int m, int nrhs, const double *u, double *b)
{
    for (int c = 1; c <= nrhs; ++c)
        for (int i = 1; i <= m; ++i) {
            double t = CM(b, m, i, c);
            for (int k = 1; k < i; ++k) t -= CM(u, m, k, i) * CM(b, m, k, c);
            CM(b, m, i, c) = t / CM(u, m, i, i);
        }
}

/* C(n x n, upper) -= A^T A with A (k x n, lda = k). */
void orc_dsyrk_ut(int n, int k, const double *a, double *c)
{
    for (int j = 1; j <= n; ++j)
        for (int i = 1; i <= j; ++i) {
            double t = 0.0;
            for (int p = 1; p <= k; ++p) t += CM(a, k, p, i) * CM(a, k, p, j);
            CM(c, n, i, j) -= t;
        }
}

/* C(m x n, ldc = m) -= A^T B with A (k x m), B (k x n), lda = ldb = k. */
void orc_dgemm_tn(int m, int n, int k, const double *a, const double *b, double *c)
{
    for (int j = 1; j <= n; ++j)
        for (int i = 1; i <= m; ++i) {
            double t = 0.0;
            for (int p = 1; p <= k; ++p) t += CM(a, k, p, i) * CM(b, k, p, j);
            CM(c, m, i, j) -= t;
        }
}

/* x <- U^{-T} x */
void orc_dtrsv_ut(int n, const double *u, double *x)
{
    for (int i = 1; i <= n; ++i) {
        double t = x[i - 1];
        for (int k = 1; k < i; ++k) t -= CM(u, n, k, i) * x[k - 1];
        x[i - 1] = t / CM(u, n, i, i);
    }
}

/* x <- U^{-1} x */
void orc_dtrsv_un(int n, const double *u, double *x)
{
    for (int i = n; i >= 1; --i) {
        double t = x[i - 1];
        for (int k = i + 1; k <= n; ++k) t -= CM(u, n, i, k) * x[k - 1];
        x[i - 1] = t / CM(u, n, i, i);
    }
}

/* y(n) -= A^T x with A (m x n, lda = m) */
void orc_dgemv_t(int m, int n, const double *a, const double *x, double *y)
{
    for (int j = 1; j <= n; ++j) {
        double t = 0.0;
        for (int i = 1; i <= m; ++i) t += CM(a, m, i, j) * x[i - 1];
        y[j - 1] -= t;
    }
}

/* y(m) -= A x with A (m x n, lda = m) */
void orc_dgemv_n(int m, int n, const double *a, const double *x, double *y)
{
    for (int j = 1; j <= n; ++j)
        for (int i = 1; i <= m; ++i) y[i - 1] -= CM(a, m, i, j) * x[j - 1];
}

/* extend_add_kernel_m.f90:6-19: to(col_local(j), row_local(i)) += from(coffset+j, roffset+i) */
void orc_extend_add_rect(const double *from, double *to, const int *col_local, int ncol,
                         const int *row_local, int nrow, int ld_from, int ld_to, int coffset, int roffset)
{
    for (int i = 1; i <= nrow; ++i)
        for (int j = 1; j <= ncol; ++j)
            CM(to, ld_to, col_local[j - 1], row_local[i - 1]) += CM(from, ld_from, coffset + j, roffset + i);
}

/* extend_add_kernel_m.f90:21-34: lower triangle of one index group */
void orc_extend_add_tri(const double *from, double *to, const int *local, int nloc,
                        int ld_from, int ld_to, int offset)
{
    for (int i = 1; i <= nloc; ++i)
        for (int j = 1; j <= i; ++j)
            CM(to, ld_to, local[j - 1], local[i - 1]) += CM(from, ld_from, offset + j, offset + i);
}

/* matrix_splitter_m.f90:5-29 */
void orc_split_tri(const double *origin, double *left, double *right, int lsize, int rsize)
{
    int lo = lsize + rsize;
    for (int i = 1; i <= lsize; ++i)
        for (int j = 1; j <= i; ++j) CM(left, lsize, j, i) = CM(origin, lo, j, i);
    for (int i = lsize + 1; i <= lo; ++i)
        for (int j = 1; j <= lsize; ++j) CM(left, lsize, j, i) = CM(origin, lo, j, i);
    for (int i = 1; i <= rsize; ++i)
        for (int j = 1; j <= i; ++j) CM(right, rsize, j, i) = CM(origin, lo, j + lsize, i + lsize);
}

/* matrix_splitter_m.f90:31-46 */
void orc_split_rect(const double *origin, double *left, double *right, int lsize, int rsize, int nrow)
{
    int lo = lsize + rsize;
    for (int i = 1; i <= nrow; ++i)
        for (int j = 1; j <= lsize; ++j) CM(left, lsize, j, i) = CM(origin, lo, j, i);
    for (int i = 1; i <= nrow; ++i)
        for (int j = 1; j <= rsize; ++j) CM(right, rsize, j, i) = CM(origin, lo, lsize + j, i);
}

/* ------------------------------------------------------- BLAS/LAPACK hooks */
/* The drivers below call the tile operations through these wrappers.  With
   no hooks set they are the textbook kernels above (the checker's own
   arithmetic); orc_set_blas installs a host BLAS/LAPACK (Fortran calling
   convention, hidden string lengths last) for the CPU comparator, calling
   the routines the reference calls at factorize_kernel_m.f90:8,15,22,32,
   border_kernel_m.f90:9-30 and solve_kernel_m.f90:8,15,22,29. */
typedef void (*potrf_f)(const char *, const int *, double *, const int *, int *, size_t);
typedef void (*trsm_f)(const char *, const char *, const char *, const char *, const int *, const int *,
                       const double *, const double *, const int *, double *, const int *, size_t, size_t, size_t,
                       size_t);
typedef void (*syrk_f)(const char *, const char *, const int *, const int *, const double *, const double *,
                       const int *, const double *, double *, const int *, size_t, size_t);
typedef void (*gemm_f)(const char *, const char *, const int *, const int *, const int *, const double *,
                       const double *, const int *, const double *, const int *, const double *, double *,
                       const int *, size_t, size_t);
typedef void (*trsv_f)(const char *, const char *, const char *, const int *, const double *, const int *, double *,
                       const int *, size_t, size_t, size_t);
typedef void (*gemv_f)(const char *, const int *, const int *, const double *, const double *, const int *,
                       const double *, const int *, const double *, double *, const int *, size_t);
static struct {
    potrf_f potrf;
    trsm_f trsm;
    syrk_f syrk;
    gemm_f gemm;
    trsv_f trsv;
    gemv_f gemv;
} blas;

/* All six or none (NULL). */
int orc_set_blas(void *potrf, void *trsm, void *syrk, void *gemm, void *trsv, void *gemv)
{
    int set = potrf && trsm && syrk && gemm && trsv && gemv;
    blas.potrf = set ? (potrf_f)potrf : 0;
    blas.trsm = set ? (trsm_f)trsm : 0;
    blas.syrk = set ? (syrk_f)syrk : 0;
    blas.gemm = set ? (gemm_f)gemm : 0;
    blas.trsv = set ? (trsv_f)trsv : 0;
    blas.gemv = set ? (gemv_f)gemv : 0;
    return set;
}

static const double ONE = 1.0, MONE = -1.0;
static const int INC1 = 1;
/* dpotrf("U") */
static void t_potrf(int n, double *a, int *info)
{
    *info = 0;
    if (n <= 0) return;
    if (blas.potrf) blas.potrf("U", &n, a, &n, info, 1);
    else orc_dpotrf_u(n, a, info);
}
/* dtrsm("L", "U", "T", "N"): B <- U^{-T} B */
static void t_trsm(int m, int nrhs, const double *u, double *b)
{
    if (m <= 0 || nrhs <= 0) return;
    if (blas.trsm) blas.trsm("L", "U", "T", "N", &m, &nrhs, &ONE, u, &m, b, &m, 1, 1, 1, 1);
    else orc_dtrsm_lutn(m, nrhs, u, b);
}
/* dsyrk("U", "T", alpha = -1, beta = 1) */
static void t_syrk(int n, int k, const double *a, double *c)
{
    if (n <= 0 || k <= 0) return;
    if (blas.syrk) blas.syrk("U", "T", &n, &k, &MONE, a, &k, &ONE, c, &n, 1, 1);
    else orc_dsyrk_ut(n, k, a, c);
}
/* dgemm("T", "N", alpha = -1, beta = 1) */
static void t_gemm(int m, int n, int k, const double *a, const double *b, double *c)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    if (blas.gemm) blas.gemm("T", "N", &m, &n, &k, &MONE, a, &k, b, &k, &ONE, c, &m, 1, 1);
    else orc_dgemm_tn(m, n, k, a, b, c);
}
/* dtrsv("U", "T", "N") and dtrsv("U", "N", "N") */
static void t_trsv_ut(int n, const double *u, double *x)
{
    if (n <= 0) return;
    if (blas.trsv) blas.trsv("U", "T", "N", &n, u, &n, x, &INC1, 1, 1, 1);
    else orc_dtrsv_ut(n, u, x);
}
static void t_trsv_un(int n, const double *u, double *x)
{
    if (n <= 0) return;
    if (blas.trsv) blas.trsv("U", "N", "N", &n, u, &n, x, &INC1, 1, 1, 1);
    else orc_dtrsv_un(n, u, x);
}
/* dgemv("T" / "N", alpha = -1, beta = 1) */
static void t_gemv_t(int m, int n, const double *a, const double *x, double *y)
{
    if (m <= 0 || n <= 0) return;
    if (blas.gemv) blas.gemv("T", &m, &n, &MONE, a, &m, x, &INC1, &ONE, y, &INC1, 1);
    else orc_dgemv_t(m, n, a, x, y);
}
static void t_gemv_n(int m, int n, const double *a, const double *x, double *y)
{
    if (m <= 0 || n <= 0) return;
    if (blas.gemv) blas.gemv("N", &m, &n, &MONE, a, &m, x, &INC1, &ONE, y, &INC1, 1);
    else orc_dgemv_n(m, n, a, x, y);
}

/* ------------------------------------------------------------- node data */
/* node_data_m.f90:8-184 */
typedef struct {
    int num_node, nb;
    int *c, *w, *parent;          /* supernode_size, work_size, parent (1-based; 0 = root) */
    int *ss, *wsz;                /* border_supernode_size, border_work_size */
    i64 *first_col;               /* first global (postordered) column of each node, 1-based */
    /* factors: three packed arrays (factors_m.f90:26-36) */
    i64 *sn_off, *wk_off, *bd_off;  /* per-node start (0-based) into each array, num_node+1 */
    double *sn, *wk, *bd;
    /* right-hand side arrays (right_hand_m.f90:24-35) */
    i64 *rsn_off, *rwk_off, *rbd_off;
    double *rsn, *rwk, *rbd;
    /* relative indices (block_local_index_info_m.f90) */
    i64 *li_off;                  /* per-node start into lidx, num_node+1 */
    int *lidx;                    /* local_index values (parent front positions) */
    int *blidx;                   /* same values after rebuild_val (mod2(.,nb)) */
    i64 *grp_off;                 /* per-node start into the group arrays, num_node+1 */
    int *grp_start;               /* group start row (1-based) within the node's list */
    int *grp_len, *grp_parent, *grp_child;
} orc_t;

static int n_of(const orc_t *h, int s) { return h->c[s] + h->w[s]; }
static int num_matrix_block(const orc_t *h, int s) { return div_ceiling(n_of(h, s), h->nb); }
static int work_start(const orc_t *h, int s)
{
    int ws = h->c[s] / h->nb + 1;
    if (h->w[s] == 0) ws = div_ceiling(h->c[s], h->nb) + 1;
    return ws;
}
static int matrix_block_size(const orc_t *h, int idx, int s) { return block_size(idx, h->nb, n_of(h, s)); }
static int supernode_block_size(const orc_t *h, int idx, int s) { return block_size(idx, h->nb, h->c[s]); }
static int work_block_size(const orc_t *h, int idx, int s)
{
    return idx == work_start(h, s) ? h->wsz[s] : block_size(idx, h->nb, n_of(h, s));
}
static int divisible(const orc_t *h, int s) { return h->ss[s] == 0; }
/* node_data_m.f90 get_matrix_num: the tile holding front position idx */
static int matrix_num(const orc_t *h, int idx) { return div_ceiling(idx, h->nb); }
/* node_data_m.f90 get_work_num: the tile holding update row idx of node s */
static int work_num(const orc_t *h, int idx, int s) { return div_ceiling(h->c[s] + idx, h->nb); }
static int exist_border(const orc_t *h, int s) { return !divisible(h, s) && h->w[s] != 0; }
enum { EX_SUPERNODE = 1, EX_BORDER = 2, EX_WORK = 3 };
static int extractor_type(const orc_t *h, int idx, int s)
{
    int ws = work_start(h, s);
    if (divisible(h, s)) return idx <= ws - 1 ? EX_SUPERNODE : EX_WORK;
    if (idx < ws) return EX_SUPERNODE;
    if (idx > ws) return EX_WORK;
    return EX_BORDER;
}

/* ----------------------------------------------------- factor extractors */
/* supernode_matrix_extractor_m.f90:14-54 */
static i64 sn_node_size(const orc_t *h, int s)
{
    int nb = h->nb, c = h->c[s], w = h->w[s];
    return psum2(w + 1, c + w) + psum1(nb - 1) * (c / nb) + psum1(h->ss[s] - 1);
}
static i64 sn_tile_pos(const orc_t *h, int s, int i, int j)
{
    int n = n_of(h, s), nb = h->nb, col = supernode_block_size(h, j, s);
    i64 left = psum2((i64)n - (i64)(j - 1) * nb + 1, n) + psum1(nb - 1) * (j - 1);
    return left + (i64)(i - j) * nb * col;   /* 0-based */
}
/* work_matrix_extractor_m.f90:14-74 */
static i64 wk_node_size(const orc_t *h, int s)
{
    int nb = h->nb, w = h->w[s];
    int last = num_matrix_block(h, s);
    int lower_rows = matrix_block_size(h, last, s);
    int fw = h->wsz[s];
    int wn = (num_matrix_block(h, s) - work_start(h, s) + 1) - 2;
    return (i64)w * fw + psum1(wn) * nb * nb + (i64)(w - fw) * lower_rows;
}
static i64 wk_tile_pos(const orc_t *h, int s, int i, int j)
{
    int nb = h->nb, c = h->c[s], w = h->w[s], fw = h->wsz[s];
    int ws = work_start(h, s);
    int colw = work_block_size(h, j, s);
    i64 up_rows, left;
    if (j == ws) {
        i64 t = (i64)nb * (i - 1) - c;
        up_rows = t > 0 ? t : 0;
        left = 0;
    } else {
        up_rows = (i64)(i - j) * nb;
        int wn = j - ws - 1;
        left = psum2((i64)c + w - (i64)(j - 1) * nb + 1, w) + psum1(nb - 1) * wn + psum1(fw - 1);
    }
    return left + colw * up_rows;
}
/* border_matrix_extractor_m.f90:13-54 */
static i64 bd_node_size(const orc_t *h, int s)
{
    if (!exist_border(h, s)) return 0;
    int nb = h->nb, c = h->c[s], w = h->w[s];
    int ws = work_start(h, s);
    return (i64)(c + w - (c / nb) * nb) * matrix_block_size(h, ws, s);
}
static i64 bd_tile_pos(const orc_t *h, int s, int i, int j) { return (i64)(i - j) * h->nb * h->nb; }

static double *get_supernode(orc_t *h, int s, int i, int j) { return h->sn + h->sn_off[s] + sn_tile_pos(h, s, i, j); }
static double *get_work(orc_t *h, int s, int i, int j) { return h->wk + h->wk_off[s] + wk_tile_pos(h, s, i, j); }
static double *get_border(orc_t *h, int s, int i, int j) { return h->bd + h->bd_off[s] + bd_tile_pos(h, s, i, j); }
/* factors_m.f90:38-55 */
static double *get_matrix(orc_t *h, int s, int i, int j)
{
    switch (extractor_type(h, j, s)) {
    case EX_SUPERNODE: return get_supernode(h, s, i, j);
    case EX_BORDER: return get_border(h, s, i, j);
    default: return get_work(h, s, i, j);
    }
}

/* --------------------------------------------------------- rhs extractors */
/* supernode_array_extractor_m.f90, work_array_extractor_m.f90, border_array_extractor_m.f90 */
static double *rh_supernode(orc_t *h, int s, int idx) { return h->rsn + h->rsn_off[s] + (i64)(idx - 1) * h->nb; }
static double *rh_work(orc_t *h, int s, int idx)
{
    int ws = work_start(h, s);
    i64 pos = idx == ws ? 0 : h->wsz[s] + (i64)(idx - ws - 1) * h->nb;
    return h->rwk + h->rwk_off[s] + pos;
}
static double *rh_border(orc_t *h, int s, int idx) { (void)idx; return h->rbd + h->rbd_off[s]; }
/* right_hand_m.f90:47-64 */
static double *rh_array(orc_t *h, int s, int idx)
{
    switch (extractor_type(h, idx, s)) {
    case EX_SUPERNODE: return rh_supernode(h, s, idx);
    case EX_BORDER: return rh_border(h, s, idx);
    default: return rh_work(h, s, idx);
    }
}

/* ------------------------------------------------------------ construction */
static i64 *prefix(int n, i64 (*f)(const orc_t *, int), const orc_t *h)
{
    i64 *off = (i64 *)calloc((size_t)n + 2, sizeof(i64));
    for (int s = 1; s <= n; ++s) off[s + 1] = off[s] + f(h, s);
    return off;
}
static i64 rsn_size(const orc_t *h, int s) { return h->c[s]; }
static i64 rwk_size(const orc_t *h, int s) { return h->w[s]; }
static i64 rbd_size(const orc_t *h, int s) { return exist_border(h, s) ? matrix_block_size(h, work_start(h, s), s) : 0; }
static i64 li_size(const orc_t *h, int s) { return h->w[s]; }

/*
 * Build node data, the three factor arrays, the rhs arrays and the
 * block_local_index structures from the front structure.
 *   c, w, parent : num_node entries (parent 1-based, 0 for the root)
 *   lidx         : sum(w) entries, node by node; entry i of node s is the
 *                  1-based position of update row i of s inside its parent
 *                  front (create_local_index_m.f90:10-45).  Ignored for roots.
 * node_data_m.f90:31-52; factors_m.f90:26-36; right_hand_m.f90:24-35;
 * block_local_index_info_m.f90:29-56; block_local_index_creator_m.f90:13-122.
 */
orc_t *orc_create(int num_node, const int *c, const int *w, const int *parent, int nb, const int *lidx)
{
    orc_t *h = (orc_t *)calloc(1, sizeof(orc_t));
    h->num_node = num_node;
    h->nb = nb;
    h->c = (int *)calloc((size_t)num_node + 1, sizeof(int));
    h->w = (int *)calloc((size_t)num_node + 1, sizeof(int));
    h->parent = (int *)calloc((size_t)num_node + 1, sizeof(int));
    h->ss = (int *)calloc((size_t)num_node + 1, sizeof(int));
    h->wsz = (int *)calloc((size_t)num_node + 1, sizeof(int));
    h->first_col = (i64 *)calloc((size_t)num_node + 2, sizeof(i64));
    h->first_col[1] = 1;
    for (int s = 1; s <= num_node; ++s) {
        h->c[s] = c[s - 1];
        h->w[s] = w[s - 1];
        h->parent[s] = parent[s - 1];
        h->ss[s] = h->c[s] % nb;
        int r = nb - h->ss[s];
        h->wsz[s] = r < h->w[s] ? r : h->w[s];
        h->first_col[s + 1] = h->first_col[s] + h->c[s];
    }
    h->sn_off = prefix(num_node, sn_node_size, h);
    h->wk_off = prefix(num_node, wk_node_size, h);
    h->bd_off = prefix(num_node, bd_node_size, h);
    h->sn = (double *)calloc((size_t)h->sn_off[num_node + 1] + 1, sizeof(double));
    h->wk = (double *)calloc((size_t)h->wk_off[num_node + 1] + 1, sizeof(double));
    h->bd = (double *)calloc((size_t)h->bd_off[num_node + 1] + 1, sizeof(double));
    h->rsn_off = prefix(num_node, rsn_size, h);
    h->rwk_off = prefix(num_node, rwk_size, h);
    h->rbd_off = prefix(num_node, rbd_size, h);
    h->rsn = (double *)calloc((size_t)h->rsn_off[num_node + 1] + 1, sizeof(double));
    h->rwk = (double *)calloc((size_t)h->rwk_off[num_node + 1] + 1, sizeof(double));
    h->rbd = (double *)calloc((size_t)h->rbd_off[num_node + 1] + 1, sizeof(double));

    /* local index and its tile-aligned groups */
    h->li_off = prefix(num_node, li_size, h);
    i64 nli = h->li_off[num_node + 1];
    h->lidx = (int *)calloc((size_t)nli + 1, sizeof(int));
    h->blidx = (int *)calloc((size_t)nli + 1, sizeof(int));
    memcpy(h->lidx + 1, lidx, (size_t)nli * sizeof(int));
    h->grp_off = (i64 *)calloc((size_t)num_node + 2, sizeof(i64));
    /* count groups: create_num_blocks, block_local_index_creator_m.f90:13-44
       (the last node is the root and is skipped, as in the reference) */
    for (int s = 1; s <= num_node; ++s) {
        i64 ng = 0;
        if (s < num_node) {
            int prev_p = 0, prev_c = 0;
            for (int i = 1; i <= h->w[s]; ++i) {
                int pn = matrix_num(h, h->lidx[h->li_off[s] + i]);
                int cn = work_num(h, i, s);
                if (pn != prev_p || cn != prev_c) { prev_p = pn; prev_c = cn; ++ng; }
            }
        }
        h->grp_off[s + 1] = h->grp_off[s] + ng;
    }
    i64 ngt = h->grp_off[num_node + 1];
    h->grp_start = (int *)calloc((size_t)ngt + 1, sizeof(int));
    h->grp_len = (int *)calloc((size_t)ngt + 1, sizeof(int));
    h->grp_parent = (int *)calloc((size_t)ngt + 1, sizeof(int));
    h->grp_child = (int *)calloc((size_t)ngt + 1, sizeof(int));
    /* create_num_indices / create_block_nums, block_local_index_creator_m.f90:46-122 */
    for (int s = 1; s < num_node; ++s) {
        int prev_p = 0, prev_c = 0;
        i64 g = h->grp_off[s];
        for (int i = 1; i <= h->w[s]; ++i) {
            int pn = matrix_num(h, h->lidx[h->li_off[s] + i]);
            int cn = work_num(h, i, s);
            if (pn != prev_p || cn != prev_c) {
                prev_p = pn; prev_c = cn; ++g;
                h->grp_start[g] = i;
                h->grp_parent[g] = pn;
                h->grp_child[g] = cn;
            }
            h->grp_len[g]++;
        }
    }
    /* rebuild_val: in-place mod2 (block_local_index_creator_m.f90:80-89) */
    for (i64 k = 1; k <= nli; ++k) h->blidx[k] = mod2(h->lidx[k], nb);
    return h;
}

void orc_destroy(orc_t *h)
{
    if (!h) return;
    free(h->c); free(h->w); free(h->parent); free(h->ss); free(h->wsz); free(h->first_col);
    free(h->sn_off); free(h->wk_off); free(h->bd_off); free(h->sn); free(h->wk); free(h->bd);
    free(h->rsn_off); free(h->rwk_off); free(h->rbd_off); free(h->rsn); free(h->rwk); free(h->rbd);
    free(h->li_off); free(h->lidx); free(h->blidx); free(h->grp_off);
    free(h->grp_start); free(h->grp_len); free(h->grp_parent); free(h->grp_child);
    free(h);
}

/* block_local_index_info_m.f90:71-85 */
static int block_offset(const orc_t *h, int s, int g)
{
    int r = h->grp_start[h->grp_off[s] + g];
    if (r <= h->wsz[s]) return r - 1;
    return (h->ss[s] + r - 1) % h->nb;
}
static const int *group_local(const orc_t *h, int s, int g)
{
    return h->blidx + h->li_off[s] + h->grp_start[h->grp_off[s] + g];
}

/* ------------------------------------------------------------- assembly */
/* zero_setter_m.f90:12-31 with set_zero_kernel_m.f90:6-25: zero every tile's
   lower (row-major) part; coefficient_setter_m.f90:15-50: scatter A.
   colptr (N+1, 0-based offsets), rows (front-relative 1-based rows =
   supernodal_index), vals: the postordered lower CCS of A, column-major over
   the global postordered columns. */
void orc_assemble(orc_t *h, const i64 *colptr, const int *rows, const double *vals)
{
    int nb = h->nb;
    for (int s = 1; s <= h->num_node; ++s) {
        int B = num_matrix_block(h, s);
        for (int j = 1; j <= B; ++j) {
            int ncol = matrix_block_size(h, j, s);
            double *a = get_matrix(h, s, j, j);
            for (int q = 1; q <= ncol; ++q)
                for (int r = 1; r <= q; ++r) CM(a, ncol, r, q) = 0.0;
            for (int i = j + 1; i <= B; ++i) {
                int nrow = matrix_block_size(h, i, s);
                a = get_matrix(h, s, i, j);
                memset(a, 0, sizeof(double) * (size_t)ncol * nrow);
            }
        }
    }
    for (int s = 1; s <= h->num_node; ++s) {
        for (i64 jg = h->first_col[s]; jg < h->first_col[s + 1]; ++jg) {
            int col = (int)(jg - h->first_col[s] + 1);
            int bcol = div_ceiling(col, nb), bj = mod2(col, nb);
            for (i64 p = colptr[jg - 1]; p < colptr[jg]; ++p) {
                int row = rows[p];
                int brow = div_ceiling(row, nb), bi = mod2(row, nb);
                int ld = matrix_block_size(h, bcol, s);
                double *m = get_matrix(h, s, brow, bcol);
                CM(m, ld, bj, bi) = vals[p];
            }
        }
    }
}

/* ----------------------------------------------------------- factorize */
/* factorize_subroutines_m.f90:9-66 */
static int st_factorize(orc_t *h, int s, int j)
{
    int info;
    t_potrf(matrix_block_size(h, j, s), get_matrix(h, s, j, j), &info);
    return info;
}
static void st_solve(orc_t *h, int s, int i, int j)
{
    t_trsm(matrix_block_size(h, j, s), matrix_block_size(h, i, s), get_matrix(h, s, j, j), get_matrix(h, s, i, j));
}
static void st_sym_update(orc_t *h, int s, int i, int j)
{
    t_syrk(matrix_block_size(h, i, s), matrix_block_size(h, j, s), get_matrix(h, s, i, j), get_matrix(h, s, i, i));
}
/* factorize_subroutines_m.f90:51-65; the panel width is passed explicitly
   (the reference passes upper_n, which equals it whenever k > i) */
static void st_update(orc_t *h, int s, int up, int lo, int col)
{
    int upper_n = matrix_block_size(h, up, s), lower_n = matrix_block_size(h, lo, s);
    int kw = matrix_block_size(h, col, s);
    t_gemm(upper_n, lower_n, kw, get_matrix(h, s, up, col), get_matrix(h, s, lo, col), get_matrix(h, s, lo, up));
}

/* seq_factorize_m.f90:31-52 */
static int supernode_factorize(orc_t *h, int s)
{
    int B = num_matrix_block(h, s), ws = work_start(h, s);
    for (int j = 1; j <= ws - 1; ++j) {
        int info = st_factorize(h, s, j);
        if (info) return (int)(h->first_col[s] - 1) + (j - 1) * h->nb + info;
        for (int i = j + 1; i <= B; ++i) st_solve(h, s, i, j);
        for (int i = j + 1; i <= B; ++i) {
            st_sym_update(h, s, i, j);
            for (int k = i + 1; k <= B; ++k) st_update(h, s, i, k, j);
        }
    }
    return 0;
}

/* seq_factorize_m.f90:54-79 with rearrange_subroutines_m.f90:12-43 and
   border_subroutines_m.f90:9-73 / border_kernel_m.f90:9-30 */
static int border_factorize(orc_t *h, int s)
{
    int j = work_start(h, s), B = num_matrix_block(h, s);
    int ss = h->ss[s], wsz = h->wsz[s];
    orc_split_tri(get_matrix(h, s, j, j), get_supernode(h, s, j, j), get_work(h, s, j, j), ss, wsz);
    for (int i = j + 1; i <= B; ++i)
        orc_split_rect(get_matrix(h, s, i, j), get_supernode(h, s, i, j), get_work(h, s, i, j), ss, wsz,
                       matrix_block_size(h, i, s));
    /* border_potrf */
    double *sup = get_supernode(h, s, j, j);
    int info;
    t_potrf(ss, sup, &info);
    if (info) return (int)(h->first_col[s] - 1) + (j - 1) * h->nb + info;
    t_trsm(ss, wsz, sup, sup + (i64)ss * ss);
    t_syrk(wsz, ss, sup + (i64)ss * ss, get_work(h, s, j, j));
    /* border_trsm */
    for (int i = j + 1; i <= B; ++i) {
        int nrow = matrix_block_size(h, i, s);
        double *sol = get_supernode(h, s, i, j);
        t_trsm(ss, nrow, sup, sol);
        t_gemm(wsz, nrow, ss, sup + (i64)ss * ss, sol, get_work(h, s, i, j));
    }
    for (int i = j + 1; i <= B; ++i) {
        int nrow = matrix_block_size(h, i, s);
        t_syrk(nrow, ss, get_supernode(h, s, i, j), get_matrix(h, s, i, i));
        for (int k = i + 1; k <= B; ++k)
            t_gemm(matrix_block_size(h, i, s), matrix_block_size(h, k, s), ss, get_supernode(h, s, i, j),
                         get_supernode(h, s, k, j), get_matrix(h, s, k, i));
    }
    return 0;
}

/* seq_factorize_m.f90:81-98 with extend_add_subroutines_m.f90:13-59 */
static void extend_add(orc_t *h, int s, int p)
{
    int G = (int)(h->grp_off[s + 1] - h->grp_off[s]);
    for (int j = 1; j <= G; ++j) {
        i64 gj = h->grp_off[s] + j;
        int cj = h->grp_child[gj], pj = h->grp_parent[gj];
        int ldp = matrix_block_size(h, pj, p), ldc = work_block_size(h, cj, s);
        orc_extend_add_tri(get_work(h, s, cj, cj), get_matrix(h, p, pj, pj), group_local(h, s, j), h->grp_len[gj], ldc,
                           ldp, block_offset(h, s, j));
        for (int i = j + 1; i <= G; ++i) {
            i64 gi = h->grp_off[s] + i;
            int ci = h->grp_child[gi], pi = h->grp_parent[gi];
            orc_extend_add_rect(get_work(h, s, ci, cj), get_matrix(h, p, pi, pj), group_local(h, s, j), h->grp_len[gj],
                                group_local(h, s, i), h->grp_len[gi], ldc, ldp, block_offset(h, s, j),
                                block_offset(h, s, i));
        }
    }
}

/* seq_factorize_m.f90:13-29.  Returns 0, or the 1-based global (postordered)
   column at which a pivot was not positive (the reference ignores DPOTRF's
   info, factorize_subroutines_m.f90:18). */
int orc_factorize(orc_t *h)
{
    for (int s = 1; s <= h->num_node; ++s) {
        int info = supernode_factorize(h, s);
        if (info) return info;
        if (exist_border(h, s)) {
            info = border_factorize(h, s);
            if (info) return info;
        }
        if (h->parent[s] > 0) extend_add(h, s, h->parent[s]);
    }
    return 0;
}

/* ------------------------------------------------------------ task DAG */
/* The reference's StarPU task graph (starpu_factorize_m.f90:13-100) as
   OpenMP tasks, for the CPU comparator: one task per tile operation, with
   the access modes of the reference's codelets on the tiles it touches
   (factorize_tasks_m.f90:12-15: POTRF RW, TRSM R/RW, SYRK R/RW|COMMUTE,
   GEMM R/R/RW|COMMUTE; rearrange_tasks_m.f90:12-13: split R/RW/RW;
   border_tasks_m.f90:12-13: border POTRF RW/RW, border TRSM R/RW/RW;
   extend_add_tasks_m.f90:12-13: child R, parent RW).  R is depend(in), RW
   depend(inout) and RW|COMMUTE depend(mutexinoutset), each on the tile's
   first entry (the tile is the StarPU data handle).  Tasks are submitted in
   the reference's order (nodes in postorder), so the depend clauses order
   them as StarPU's sequential consistency does.  The tile operations are
   the same wrappers as the sequential driver's (host BLAS when hooks are
   set).  Returns 0 or the first failing global column. */
static void dag_fail(int *fail, int col)
{
#pragma omp critical(orc_dag_fail)
    if (*fail == 0 || col < *fail) *fail = col;
}

int orc_factorize_dag(orc_t *h, int nthreads)
{
    int fail = 0;
    int *failp = &fail;
    const int nb = h->nb;
#pragma omp parallel num_threads(nthreads)
#pragma omp single
    for (int s = 1; s <= h->num_node; ++s) {
        const int B = num_matrix_block(h, s), ws = work_start(h, s);
        const int col0 = (int)(h->first_col[s] - 1);
        for (int j = 1; j <= ws - 1; ++j) {
            double *d = get_matrix(h, s, j, j);
            const int nj = matrix_block_size(h, j, s);
#pragma omp task depend(inout : d[0]) firstprivate(d, nj, j)
            {
                int info;
                t_potrf(nj, d, &info);
                if (info) dag_fail(failp, col0 + (j - 1) * nb + info);
            }
            for (int i = j + 1; i <= B; ++i) {
                double *a = get_matrix(h, s, i, j);
                const int ni = matrix_block_size(h, i, s);
#pragma omp task depend(in : d[0]) depend(inout : a[0]) firstprivate(a, d, ni, nj)
                t_trsm(nj, ni, d, a);
            }
            for (int i = j + 1; i <= B; ++i) {
                double *a = get_matrix(h, s, i, j), *c = get_matrix(h, s, i, i);
                const int ni = matrix_block_size(h, i, s);
#pragma omp task depend(in : a[0]) depend(mutexinoutset : c[0]) firstprivate(a, c, ni, nj)
                t_syrk(ni, nj, a, c);
                for (int k = i + 1; k <= B; ++k) {
                    double *b = get_matrix(h, s, k, j), *u = get_matrix(h, s, k, i);
                    const int nk = matrix_block_size(h, k, s);
#pragma omp task depend(in : a[0], b[0]) depend(mutexinoutset : u[0]) firstprivate(a, b, u, ni, nk, nj)
                    t_gemm(ni, nk, nj, a, b, u);
                }
            }
        }
        if (exist_border(h, s)) {
            /* seq_factorize_m.f90:54-79 as tasks */
            const int j = ws, ss = h->ss[s], wsz = h->wsz[s];
            double *bd = get_matrix(h, s, j, j), *sup = get_supernode(h, s, j, j), *wk = get_work(h, s, j, j);
#pragma omp task depend(in : bd[0]) depend(inout : sup[0], wk[0]) firstprivate(bd, sup, wk, ss, wsz)
            orc_split_tri(bd, sup, wk, ss, wsz);
            for (int i = j + 1; i <= B; ++i) {
                double *bi = get_matrix(h, s, i, j), *si = get_supernode(h, s, i, j), *wi = get_work(h, s, i, j);
                const int ni = matrix_block_size(h, i, s);
#pragma omp task depend(in : bi[0]) depend(inout : si[0], wi[0]) firstprivate(bi, si, wi, ss, wsz, ni)
                orc_split_rect(bi, si, wi, ss, wsz, ni);
            }
#pragma omp task depend(inout : sup[0], wk[0]) firstprivate(sup, wk, ss, wsz, j)
            {
                /* border_potrf, border_kernel_m.f90:9-19 */
                int info;
                t_potrf(ss, sup, &info);
                if (info) dag_fail(failp, col0 + (j - 1) * nb + info);
                t_trsm(ss, wsz, sup, sup + (i64)ss * ss);
                t_syrk(wsz, ss, sup + (i64)ss * ss, wk);
            }
            for (int i = j + 1; i <= B; ++i) {
                double *si = get_supernode(h, s, i, j), *wi = get_work(h, s, i, j);
                const int ni = matrix_block_size(h, i, s);
#pragma omp task depend(in : sup[0]) depend(inout : si[0], wi[0]) firstprivate(sup, si, wi, ss, wsz, ni)
                {
                    /* border_trsm, border_kernel_m.f90:21-30 */
                    t_trsm(ss, ni, sup, si);
                    t_gemm(wsz, ni, ss, sup + (i64)ss * ss, si, wi);
                }
            }
            for (int i = j + 1; i <= B; ++i) {
                double *si = get_supernode(h, s, i, j), *c = get_matrix(h, s, i, i);
                const int ni = matrix_block_size(h, i, s);
#pragma omp task depend(in : si[0]) depend(mutexinoutset : c[0]) firstprivate(si, c, ni, ss)
                t_syrk(ni, ss, si, c);
                for (int k = i + 1; k <= B; ++k) {
                    double *sk = get_supernode(h, s, k, j), *u = get_matrix(h, s, k, i);
                    const int nk = matrix_block_size(h, k, s);
#pragma omp task depend(in : si[0], sk[0]) depend(mutexinoutset : u[0]) firstprivate(si, sk, u, ni, nk, ss)
                    t_gemm(ni, nk, ss, si, sk, u);
                }
            }
        }
        const int p = h->parent[s];
        if (p > 0) {
            /* seq_factorize_m.f90:81-98 as tasks (extend_add_subroutines_m.f90:13-59) */
            const int G = (int)(h->grp_off[s + 1] - h->grp_off[s]);
            for (int g = 1; g <= G; ++g) {
                const i64 gj = h->grp_off[s] + g;
                const int cj = h->grp_child[gj], pj = h->grp_parent[gj];
                const int ldp = matrix_block_size(h, pj, p), ldc = work_block_size(h, cj, s);
                const int *lj = group_local(h, s, g), nlj = h->grp_len[gj], oj = block_offset(h, s, g);
                double *from = get_work(h, s, cj, cj), *to = get_matrix(h, p, pj, pj);
#pragma omp task depend(in : from[0]) depend(inout : to[0]) firstprivate(from, to, lj, nlj, ldc, ldp, oj)
                orc_extend_add_tri(from, to, lj, nlj, ldc, ldp, oj);
                for (int i = g + 1; i <= G; ++i) {
                    const i64 gi = h->grp_off[s] + i;
                    const int ci = h->grp_child[gi], pi = h->grp_parent[gi];
                    const int *li = group_local(h, s, i), nli = h->grp_len[gi], oi = block_offset(h, s, i);
                    double *f2 = get_work(h, s, ci, cj), *t2 = get_matrix(h, p, pi, pj);
#pragma omp task depend(in : f2[0]) depend(inout : t2[0]) firstprivate(f2, t2, lj, nlj, li, nli, ldc, ldp, oj, oi)
                    orc_extend_add_rect(f2, t2, lj, nlj, li, nli, ldc, ldp, oj, oi);
                }
            }
        }
    }
    return fail;
}

/* orc_assemble with the nodes spread over nthreads (the reference's
   omp parallel do over nodes, zero_setter_m.f90:18, coefficient_setter_m.f90:25). */
void orc_assemble_par(orc_t *h, const i64 *colptr, const int *rows, const double *vals, int nthreads)
{
    int nb = h->nb;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int s = 1; s <= h->num_node; ++s) {
        int B = num_matrix_block(h, s);
        for (int j = 1; j <= B; ++j) {
            int ncol = matrix_block_size(h, j, s);
            double *a = get_matrix(h, s, j, j);
            for (int q = 1; q <= ncol; ++q)
                for (int r = 1; r <= q; ++r) CM(a, ncol, r, q) = 0.0;
            for (int i = j + 1; i <= B; ++i) {
                int nrow = matrix_block_size(h, i, s);
                a = get_matrix(h, s, i, j);
                memset(a, 0, sizeof(double) * (size_t)ncol * nrow);
            }
        }
        for (i64 jg = h->first_col[s]; jg < h->first_col[s + 1]; ++jg) {
            int col = (int)(jg - h->first_col[s] + 1);
            int bcol = div_ceiling(col, nb), bj = mod2(col, nb);
            for (i64 p = colptr[jg - 1]; p < colptr[jg]; ++p) {
                int row = rows[p];
                int brow = div_ceiling(row, nb), bi = mod2(row, nb);
                int ld = matrix_block_size(h, bcol, s);
                double *m = get_matrix(h, s, brow, bcol);
                CM(m, ld, bj, bi) = vals[p];
            }
        }
    }
}

/* ---------------------------------------------------------------- solve */
/* seq_forward_m.f90:14-81 */
static void forward_sweep(orc_t *h)
{
    for (int s = 1; s <= h->num_node; ++s) {
        int B = num_matrix_block(h, s), ws = work_start(h, s);
        for (int j = 1; j <= ws - 1; ++j) {
            t_trsv_ut(matrix_block_size(h, j, s), get_supernode(h, s, j, j), rh_array(h, s, j));
            for (int i = j + 1; i <= B; ++i)
                t_gemv_t(supernode_block_size(h, j, s), matrix_block_size(h, i, s), get_supernode(h, s, i, j),
                            rh_array(h, s, j), rh_array(h, s, i));
        }
        if (exist_border(h, s)) {
            int j = ws, ncol = supernode_block_size(h, j, s), nrow = matrix_block_size(h, j, s);
            /* split_array, array_rearrange_kernel_m.f90:8-15 */
            double *bd = rh_border(h, s, j), *sp = rh_supernode(h, s, j), *wk = rh_work(h, s, j);
            for (int k = 0; k < ncol; ++k) sp[k] += bd[k];
            for (int k = 0; k < nrow - ncol; ++k) wk[k] = bd[ncol + k];
            /* border_dtrsv_l, border_solve_kernel_m.f90:9-16 */
            double *a = get_supernode(h, s, j, j);
            t_trsv_ut(ncol, a, sp);
            t_gemv_t(ncol, nrow - ncol, a + (i64)ncol * ncol, sp, wk);
            for (int i = j + 1; i <= B; ++i)
                t_gemv_t(ncol, matrix_block_size(h, i, s), get_supernode(h, s, i, j), sp, rh_array(h, s, i));
        }
        int p = h->parent[s];
        if (p > 0) {
            /* scatter_add, sparse_add_subroutines_m.f90:13-30 */
            int G = (int)(h->grp_off[s + 1] - h->grp_off[s]);
            for (int g = 1; g <= G; ++g) {
                i64 gg = h->grp_off[s] + g;
                const double *x = rh_work(h, s, h->grp_child[gg]) + block_offset(h, s, g);
                double *y = rh_array(h, p, h->grp_parent[gg]);
                const int *loc = group_local(h, s, g);
                for (int k = 0; k < h->grp_len[gg]; ++k) y[loc[k] - 1] += x[k];
            }
        }
    }
}

/* seq_backward_m.f90:14-131 */
static void backward_sweep(orc_t *h)
{
    for (int s = h->num_node; s >= 1; --s) {
        int p = h->parent[s];
        if (p > 0) {
            /* gather, sparse_add_subroutines_m.f90:32-49 */
            int G = (int)(h->grp_off[s + 1] - h->grp_off[s]);
            for (int g = 1; g <= G; ++g) {
                i64 gg = h->grp_off[s] + g;
                double *y = rh_work(h, s, h->grp_child[gg]) + block_offset(h, s, g);
                const double *x = rh_array(h, p, h->grp_parent[gg]);
                const int *loc = group_local(h, s, g);
                for (int k = 0; k < h->grp_len[gg]; ++k) y[k] = x[loc[k] - 1];
            }
        }
        int B = num_matrix_block(h, s), ws = work_start(h, s);
        if (exist_border(h, s)) {
            int j = ws, ncol = supernode_block_size(h, j, s), nrow = matrix_block_size(h, j, s);
            double *sp = rh_supernode(h, s, j), *wk = rh_work(h, s, j), *bd = rh_border(h, s, j);
            for (int i = j + 1; i <= B; ++i)
                t_gemv_n(ncol, matrix_block_size(h, i, s), get_supernode(h, s, i, j), rh_array(h, s, i), sp);
            double *a = get_supernode(h, s, j, j);
            t_gemv_n(ncol, nrow - ncol, a + (i64)ncol * ncol, wk, sp);
            t_trsv_un(ncol, a, sp);
            /* join_array, array_rearrange_kernel_m.f90:17-24 */
            for (int k = 0; k < ncol; ++k) bd[k] = sp[k];
            for (int k = 0; k < nrow - ncol; ++k) bd[ncol + k] = wk[k];
        }
        for (int j = ws - 1; j >= 1; --j) {
            for (int i = j + 1; i <= B; ++i)
                t_gemv_n(supernode_block_size(h, j, s), matrix_block_size(h, i, s), get_supernode(h, s, i, j),
                            rh_array(h, s, i), rh_array(h, s, j));
            t_trsv_un(matrix_block_size(h, j, s), get_supernode(h, s, j, j), rh_array(h, s, j));
        }
    }
}

/* hppllt.f90:152-173 between the two permutations: bperm (length N, in the
   postordered supernodal numbering) is overwritten with the solution.
   right_hand_m.f90:37-45: work and border arrays are zeroed, the supernode
   array is the permuted vector itself. */
void orc_solve(orc_t *h, double *bperm)
{
    i64 N = h->rsn_off[h->num_node + 1];
    memcpy(h->rsn, bperm, sizeof(double) * (size_t)N);
    memset(h->rwk, 0, sizeof(double) * (size_t)h->rwk_off[h->num_node + 1]);
    memset(h->rbd, 0, sizeof(double) * (size_t)h->rbd_off[h->num_node + 1]);
    forward_sweep(h);
    backward_sweep(h);
    memcpy(bperm, h->rsn, sizeof(double) * (size_t)N);
}
/* forward sweep only (seq_forward_test.f90 checks the result after it) */
void orc_forward(orc_t *h, double *bperm)
{
    i64 N = h->rsn_off[h->num_node + 1];
    memcpy(h->rsn, bperm, sizeof(double) * (size_t)N);
    memset(h->rwk, 0, sizeof(double) * (size_t)h->rwk_off[h->num_node + 1]);
    memset(h->rbd, 0, sizeof(double) * (size_t)h->rbd_off[h->num_node + 1]);
    forward_sweep(h);
    memcpy(bperm, h->rsn, sizeof(double) * (size_t)N);
}
/* backward sweep only (seq_backward_test.f90) */
void orc_backward(orc_t *h, double *bperm)
{
    i64 N = h->rsn_off[h->num_node + 1];
    memcpy(h->rsn, bperm, sizeof(double) * (size_t)N);
    memset(h->rwk, 0, sizeof(double) * (size_t)h->rwk_off[h->num_node + 1]);
    memset(h->rbd, 0, sizeof(double) * (size_t)h->rbd_off[h->num_node + 1]);
    backward_sweep(h);
    memcpy(bperm, h->rsn, sizeof(double) * (size_t)N);
}

/* -------------------------------------------------------------- access */
i64 orc_array_size(orc_t *h, int which)
{
    switch (which) {
    case 1: return h->sn_off[h->num_node + 1];
    case 2: return h->wk_off[h->num_node + 1];
    case 3: return h->bd_off[h->num_node + 1];
    default: return -1;
    }
}
double *orc_array(orc_t *h, int which)
{
    switch (which) {
    case 1: return h->sn;
    case 2: return h->wk;
    case 3: return h->bd;
    default: return 0;
    }
}
/* 0-based start of tile (i,j) of node s inside array `which`, and its
   (rows, cols) in the row-major tile sense */
i64 orc_tile(orc_t *h, int which, int s, int i, int j, int *rows, int *cols)
{
    switch (which) {
    case 1:
        *rows = matrix_block_size(h, i, s);
        *cols = supernode_block_size(h, j, s);
        return h->sn_off[s] + sn_tile_pos(h, s, i, j);
    case 2:
        *rows = work_block_size(h, i, s);
        *cols = work_block_size(h, j, s);
        return h->wk_off[s] + wk_tile_pos(h, s, i, j);
    default:
        *rows = matrix_block_size(h, i, s);
        *cols = matrix_block_size(h, j, s);
        return h->bd_off[s] + bd_tile_pos(h, s, i, j);
    }
}
int orc_num_matrix_block(orc_t *h, int s) { return num_matrix_block(h, s); }
/* The layout quantities the reference's layout tests assert, for the
   checker's known-answer tests (node s and indices 1-based):
    0 divisible(s)              1 border_supernode_size(s)   2 border_work_size(s)
    3 num_matrix_block(s)       4 work_start_index(s)        5 matrix_block_size(i, s)
    6 supernode_block_size(i,s) 7 work_block_size(i, s)      8 matrix_num(i)
    9 work_num(i, s)
   10/11/12 entries of node s in the supernode / work / border factor array
   13/14/15 0-based start of tile (i, j) of node s within its part of that array
   16/17/18 entries of node s in the supernode / work / border rhs array
   19/20/21 0-based start of block i of node s within its part of that rhs array
   (get_block_size with a first block is orc_block_size) */
i64 orc_layout(orc_t *h, int what, int s, int i, int j)
{
    switch (what) {
    case 0: return divisible(h, s);
    case 1: return h->ss[s];
    case 2: return h->wsz[s];
    case 3: return num_matrix_block(h, s);
    case 4: return work_start(h, s);
    case 5: return matrix_block_size(h, i, s);
    case 6: return supernode_block_size(h, i, s);
    case 7: return work_block_size(h, i, s);
    case 8: return matrix_num(h, i);
    case 9: return work_num(h, i, s);
    case 10: return sn_node_size(h, s);
    case 11: return wk_node_size(h, s);
    case 12: return bd_node_size(h, s);
    case 13: return sn_tile_pos(h, s, i, j);
    case 14: return wk_tile_pos(h, s, i, j);
    case 15: return bd_tile_pos(h, s, i, j);
    case 16: return rsn_size(h, s);
    case 17: return rwk_size(h, s);
    case 18: return rbd_size(h, s);
    case 19: return rh_supernode(h, s, i) - (h->rsn + h->rsn_off[s]);
    case 20: return rh_work(h, s, i) - (h->rwk + h->rwk_off[s]);
    case 21: return rh_border(h, s, i) - (h->rbd + h->rbd_off[s]);
    default: return -1;
    }
}
int orc_block_size(int idx, int nb, int n, int first_block) { return block_size2(idx, nb, n, first_block); }
int orc_work_start(orc_t *h, int s) { return work_start(h, s); }
int orc_num_groups(orc_t *h, int s) { return (int)(h->grp_off[s + 1] - h->grp_off[s]); }
/* group g of node s: fills start_row, parent tile, child tile, block offset,
   and copies the tile-local indices into out (must hold len entries) */
int orc_group(orc_t *h, int s, int g, int *start_row, int *parent_num, int *child_num, int *offset, int *out)
{
    i64 gg = h->grp_off[s] + g;
    *start_row = h->grp_start[gg];
    *parent_num = h->grp_parent[gg];
    *child_num = h->grp_child[gg];
    *offset = block_offset(h, s, g);
    const int *loc = group_local(h, s, g);
    for (int k = 0; k < h->grp_len[gg]; ++k) out[k] = loc[k];
    return h->grp_len[gg];
}

/* Dense front view of node s after factorization, column-major n x n
   (ld = n, 0-based front positions): the first c columns hold L, the
   trailing w x w block holds the update matrix; only r >= q is written. */
void orc_front(orc_t *h, int s, double *out)
{
    int n = n_of(h, s), nb = h->nb, c = h->c[s];
    int B = num_matrix_block(h, s), ws = work_start(h, s);
    for (int j = 1; j <= B; ++j) {
        int et = extractor_type(h, j, s);
        for (int i = j; i <= B; ++i) {
            int rows = matrix_block_size(h, i, s);
            for (int ri = 1; ri <= rows; ++ri) {
                int r = (i - 1) * nb + ri - 1;
                if (et == EX_SUPERNODE || et == EX_BORDER) {
                    int cols = supernode_block_size(h, j, s);
                    double *t = get_supernode(h, s, i, j);
                    for (int qj = 1; qj <= cols; ++qj) {
                        int q = (j - 1) * nb + qj - 1;
                        if (q > r) break;
                        out[(i64)q * n + r] = t[(i64)(ri - 1) * cols + (qj - 1)];
                    }
                }
                if (et == EX_WORK || et == EX_BORDER) {
                    /* work tile (i, j): rows of work block i, cols of work block j */
                    int wcols = work_block_size(h, j, s);
                    int wrow0 = (i == ws) ? c : (i - 1) * nb;   /* first front row of work block i */
                    int wcol0 = (j == ws) ? c : (j - 1) * nb;
                    if (r < wrow0) continue;
                    double *t = get_work(h, s, i, j);
                    int wr = r - wrow0;
                    for (int qj = 0; qj < wcols; ++qj) {
                        int q = wcol0 + qj;
                        if (q > r) break;
                        out[(i64)q * n + r] = t[(i64)wr * wcols + qj];
                    }
                }
            }
        }
    }
}
