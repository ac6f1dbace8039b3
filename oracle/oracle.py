"""ctypes front end of the C oracle (mfd_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the GPU path.  Only tests/, the
smoke check and bench.py's cpu_baseline leg may import it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libmfd_oracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = C.c_void_p
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        lp = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        i = C.c_int
        L.orc_create.restype = P
        L.orc_create.argtypes = [i, ip, ip, ip, i, ip]
        L.orc_destroy.argtypes = [P]
        L.orc_assemble.argtypes = [P, lp, ip, dp]
        L.orc_factorize.argtypes = [P]
        L.orc_factorize.restype = i
        for f in (L.orc_solve, L.orc_forward, L.orc_backward):
            f.argtypes = [P, dp]
        L.orc_array_size.restype = C.c_int64
        L.orc_array_size.argtypes = [P, i]
        L.orc_array.restype = C.POINTER(C.c_double)
        L.orc_array.argtypes = [P, i]
        L.orc_tile.restype = C.c_int64
        L.orc_tile.argtypes = [P, i, i, i, i, C.POINTER(i), C.POINTER(i)]
        for f in (L.orc_num_matrix_block, L.orc_work_start, L.orc_num_groups):
            f.restype = i
            f.argtypes = [P, i]
        L.orc_group.restype = i
        L.orc_group.argtypes = [P, i, i] + [C.POINTER(i)] * 4 + [ip]
        L.orc_front.argtypes = [P, i, dp]
        L.orc_layout.restype = C.c_int64
        L.orc_layout.argtypes = [P, i, i, i, i]
        L.orc_block_size.restype = i
        L.orc_block_size.argtypes = [i, i, i, i]
        # tile kernels
        L.orc_dpotrf_u.argtypes = [i, dp, C.POINTER(i)]
        L.orc_dtrsm_lutn.argtypes = [i, i, dp, dp]
        L.orc_dsyrk_ut.argtypes = [i, i, dp, dp]
        L.orc_dgemm_tn.argtypes = [i, i, i, dp, dp, dp]
        L.orc_dtrsv_ut.argtypes = [i, dp, dp]
        L.orc_dtrsv_un.argtypes = [i, dp, dp]
        L.orc_dgemv_t.argtypes = [i, i, dp, dp, dp]
        L.orc_dgemv_n.argtypes = [i, i, dp, dp, dp]
        L.orc_extend_add_rect.argtypes = [dp, dp, ip, i, ip, i, i, i, i, i]
        L.orc_extend_add_tri.argtypes = [dp, dp, ip, i, i, i, i]
        L.orc_split_tri.argtypes = [dp, dp, dp, i, i]
        L.orc_split_rect.argtypes = [dp, dp, dp, i, i, i]
        L.orc_set_blas.argtypes = [P] * 6
        L.orc_set_blas.restype = i
        L.orc_factorize_dag.argtypes = [P, i]
        L.orc_factorize_dag.restype = i
        L.orc_assemble_par.argtypes = [P, lp, ip, dp, i]
        _lib = L
    return _lib


_blas = None


def host_blas():
    """The host BLAS/LAPACK the CPU comparator calls for its tile operations:
    scipy's bundled OpenBLAS (LP64), one thread per call (the reference links
    sequential MKL, camphor2.cmake:2), installed into the oracle's tile
    wrappers.  Returns a description, or None (the oracle's own textbook
    kernels stay) when it is not found."""
    global _blas
    if _blas is None:
        import glob
        try:
            import scipy
        except ImportError:
            return None
        libs = sorted(glob.glob(os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs",
                                             "libscipy_openblas*.so")))
        if not libs:
            return None
        B = C.CDLL(libs[0])
        B.scipy_openblas_set_num_threads.argtypes = [C.c_int]
        B.scipy_openblas_set_num_threads(1)
        names = ["dpotrf_", "dtrsm_", "dsyrk_", "dgemm_", "dtrsv_", "dgemv_"]
        ptrs = [C.cast(getattr(B, "scipy_" + n), C.c_void_p) for n in names]
        if not lib().orc_set_blas(*ptrs):
            return None
        cfg = B.scipy_openblas_get_config
        cfg.restype = C.c_char_p
        _blas = (B, f"scipy-bundled OpenBLAS ({os.path.basename(libs[0])}; {cfg().decode().strip()}), 1 thread per call")
    return _blas[1]


def textbook_kernels():
    """Back to the oracle's own tile kernels (after host_blas)."""
    global _blas
    lib().orc_set_blas(None, None, None, None, None, None)
    _blas = None


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


class OracleFactor:
    """The reference's numeric phase on one problem, in the reference layout.

    c, w, parent: per-node supernode size, work size and parent (1-based, 0
    for a root); lidx: concatenated local_index (parent front positions)."""

    def __init__(self, c, w, parent, nb, lidx):
        L = lib()
        self.c = _i32(c)
        self.w = _i32(w)
        self.parent = _i32(parent)
        self.nb = int(nb)
        self.num_node = len(self.c)
        lid = _i32(lidx) if len(lidx) else np.zeros(1, np.int32)
        self.h = L.orc_create(self.num_node, self.c, self.w, self.parent, self.nb, lid)
        self.N = int(self.c.sum())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def assemble(self, colptr, rows, vals):
        """colptr: N+1 0-based offsets over the postordered columns; rows:
        front-relative 1-based rows (supernodal_index); vals: values."""
        lib().orc_assemble(self.h, np.ascontiguousarray(colptr, np.int64), _i32(rows),
                           np.ascontiguousarray(vals, np.float64))

    def factorize(self):
        return lib().orc_factorize(self.h)

    def assemble_par(self, colptr, rows, vals, nthreads):
        """assemble with the nodes spread over nthreads threads."""
        lib().orc_assemble_par(self.h, np.ascontiguousarray(colptr, np.int64), _i32(rows),
                               np.ascontiguousarray(vals, np.float64), int(nthreads))

    def factorize_dag(self, nthreads):
        """The reference's StarPU task graph as OpenMP tasks on nthreads
        threads (orc_factorize_dag)."""
        return lib().orc_factorize_dag(self.h, int(nthreads))

    def solve(self, bperm):
        x = np.array(bperm, dtype=np.float64)
        lib().orc_solve(self.h, x)
        return x

    def forward(self, bperm):
        x = np.array(bperm, dtype=np.float64)
        lib().orc_forward(self.h, x)
        return x

    def backward(self, bperm):
        x = np.array(bperm, dtype=np.float64)
        lib().orc_backward(self.h, x)
        return x

    def array(self, which):
        """1 = supernode, 2 = work, 3 = border (a view; keep self alive)."""
        n = lib().orc_array_size(self.h, which)
        if n == 0:
            return np.zeros(0)
        p = lib().orc_array(self.h, which)
        return np.ctypeslib.as_array(p, shape=(n,))

    def tile(self, which, node, i, j):
        r, c = C.c_int(), C.c_int()
        off = lib().orc_tile(self.h, which, node, i, j, C.byref(r), C.byref(c))
        return off, r.value, c.value

    def tile_view(self, which, node, i, j):
        off, r, c = self.tile(which, node, i, j)
        return self.array(which)[off:off + r * c]

    def groups(self, node):
        L = lib()
        out = []
        buf = np.zeros(max(int(self.w[node - 1]), 1), np.int32)
        for g in range(1, L.orc_num_groups(self.h, node) + 1):
            a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            n = L.orc_group(self.h, node, g, C.byref(a), C.byref(b), C.byref(c), C.byref(d), buf)
            out.append((buf[:n].tolist(), a.value, d.value, b.value, c.value))
        return out

    def layout(self, what, node, i=0, j=0):
        """One layout quantity of node_data / the extractors (orc_layout's
        table in mfd_oracle.c)."""
        return int(lib().orc_layout(self.h, int(what), int(node), int(i), int(j)))

    def front(self, node):
        """Dense n x n (column-major, lower part) front of `node` after
        factorization: L in the first c columns, update matrix below-right."""
        n = int(self.c[node - 1] + self.w[node - 1])
        out = np.zeros(n * n)
        lib().orc_front(self.h, node, out)
        return out.reshape(n, n, order="F")
