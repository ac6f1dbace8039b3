"""Synthetic SPD test problems and their nested-dissection orderings.

Every matrix is returned as the reference's input format: the lower triangle
in 1-based compressed columns with the diagonal first in each column
(hppllt.f90:169-176), values in the same order.  Orderings are new -> old,
1-based (the convention of METIS's perm, ordering_m.f90:11-35).  Values come
from numpy's PCG64 with the given seed, generated in lexicographic node order
(x fastest) and (axis, node) edge order.
"""
from __future__ import annotations

import numpy as np


class Problem:
    def __init__(self, n, ccs_col, ccs_row, val, perm=None, name="", coords=None):
        self.n = n
        self.ccs_col = ccs_col
        self.ccs_row = ccs_row
        self.val = val
        self.perm = perm
        self.name = name
        self.coords = coords

    @property
    def nnz(self):
        return len(self.ccs_row)

    def dense(self):
        A = np.zeros((self.n, self.n))
        for j in range(self.n):
            for p in range(self.ccs_col[j] - 1, self.ccs_col[j + 1] - 1):
                r = self.ccs_row[p] - 1
                A[r, j] = self.val[p]
                A[j, r] = self.val[p]
        return A

    def scipy(self):
        import scipy.sparse as sp
        cols = np.repeat(np.arange(self.n), np.diff(self.ccs_col))
        rows = self.ccs_row - 1
        L = sp.csc_matrix((self.val, (rows, cols)), shape=(self.n, self.n))
        return (L + sp.tril(L, -1).T).tocsr()

    def matvec(self, x):
        """y = A x with the symmetric matrix (x: (n,) or (n, k))."""
        return self.scipy() @ x


def lower_ccs(n, rows, cols, vals):
    """COO of the lower triangle (row >= col, 0-based, diagonal included) ->
    1-based CCS with the diagonal first and rows ascending."""
    order = np.lexsort((rows, cols))
    rows, cols, vals = rows[order], cols[order], vals[order]
    ccs_col = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ccs_col, cols + 1, 1)
    ccs_col = np.cumsum(ccs_col) + 1
    return ccs_col.astype(np.int32), (rows + 1).astype(np.int32), vals.astype(np.float64)


def grid_edges(shape, offsets):
    """Edges (i < j) of a structured grid for the given neighbour offsets
    (each a tuple over the axes, lexicographically positive), in (offset,
    node) order; node index = x + nx*(y + ny*z)."""
    dims = len(shape)
    idx = np.arange(int(np.prod(shape))).reshape(shape[::-1])  # [z][y][x]
    out = []
    for off in offsets:
        sl_a, sl_b = [], []
        for ax in range(dims):
            o = off[ax]
            # array axis order is reversed (x is the last array axis)
            if o > 0:
                sl_a.append(slice(0, shape[ax] - o))
                sl_b.append(slice(o, shape[ax]))
            elif o < 0:
                sl_a.append(slice(-o, shape[ax]))
                sl_b.append(slice(0, shape[ax] + o))
            else:
                sl_a.append(slice(None))
                sl_b.append(slice(None))
        a = idx[tuple(sl_a[::-1])].ravel()
        b = idx[tuple(sl_b[::-1])].ravel()
        out.append((np.minimum(a, b), np.maximum(a, b)))
    return out


def nested_dissection(shape, leaf=64, coords_sep=None):
    """Recursive coordinate bisection of a structured grid: split the longest
    extent with a one-node-thick plane separator, order the two halves first
    and the separator last; boxes of <= `leaf` nodes keep natural order.
    Returns perm (new -> old, 1-based)."""
    dims = len(shape)
    strides = [1]
    for d in range(dims - 1):
        strides.append(strides[-1] * shape[d])
    order = []

    def box_nodes(lo, hi):
        axes = [np.arange(lo[d], hi[d]) for d in range(dims)]
        grid = np.meshgrid(*axes[::-1], indexing="ij")  # z, y, x
        idx = np.zeros(grid[0].shape, dtype=np.int64)
        for d in range(dims):
            idx += grid[dims - 1 - d] * strides[d]
        return idx.ravel()

    stack = [(tuple(0 for _ in shape), tuple(shape), False)]
    # iterative post-order: (box, expanded) entries
    out_stack = []
    while stack:
        lo, hi, done = stack.pop()
        ext = [hi[d] - lo[d] for d in range(dims)]
        size = int(np.prod(ext))
        if size == 0:
            continue
        if done or size <= leaf or max(ext) < 3:
            out_stack.append(box_nodes(lo, hi))
            continue
        d = int(np.argmax(ext[::-1]))
        d = dims - 1 - d  # ties: prefer the slowest axis (z, then y, then x)
        mid = lo[d] + ext[d] // 2
        left_hi = list(hi)
        left_hi[d] = mid
        right_lo = list(lo)
        right_lo[d] = mid + 1
        sep_lo = list(lo)
        sep_lo[d] = mid
        sep_hi = list(hi)
        sep_hi[d] = mid + 1
        # pushed in reverse so that left, right, separator come out in order
        stack.append((tuple(sep_lo), tuple(sep_hi), True))
        stack.append((tuple(right_lo), hi, False))
        stack.append((lo, tuple(left_hi), False))
    order = np.concatenate(out_stack) if out_stack else np.zeros(0, np.int64)
    return (order + 1).astype(np.int32)


def laplacian_2d(nx, ny=None, leaf=64):
    """2D 5-point Laplacian, Dirichlet: A_ii = 4, A_ij = -1."""
    ny = ny or nx
    n = nx * ny
    (a1, b1), (a2, b2) = grid_edges((nx, ny), [(1, 0), (0, 1)])
    r = np.concatenate([np.arange(n), b1, b2])
    c = np.concatenate([np.arange(n), a1, a2])
    v = np.concatenate([np.full(n, 4.0), np.full(len(a1) + len(a2), -1.0)])
    cc, rr, vv = lower_ccs(n, r, c, v)
    return Problem(n, cc, rr, vv, nested_dissection((nx, ny), leaf), f"lap2d_{nx}x{ny}")


def weighted_laplacian_3d(nx, ny=None, nz=None, seed=0, lo=1e-2, hi=1e2, shift=1e-3, leaf=64):
    """3D 7-point Laplacian with edge weights log-uniform in [lo, hi];
    A_ij = -w, A_ii = sum_j |A_ij| + shift."""
    ny = ny or nx
    nz = nz or nx
    n = nx * ny * nz
    rng = np.random.Generator(np.random.PCG64(seed))
    edges = grid_edges((nx, ny, nz), [(1, 0, 0), (0, 1, 0), (0, 0, 1)])
    a = np.concatenate([e[0] for e in edges])
    b = np.concatenate([e[1] for e in edges])
    w = np.exp(rng.uniform(np.log(lo), np.log(hi), size=len(a)))
    diag = np.full(n, shift)
    np.add.at(diag, a, w)
    np.add.at(diag, b, w)
    cc, rr, vv = lower_ccs(n, np.concatenate([np.arange(n), b]), np.concatenate([np.arange(n), a]),
                           np.concatenate([diag, -w]))
    return Problem(n, cc, rr, vv, nested_dissection((nx, ny, nz), leaf), f"wlap3d_{nx}x{ny}x{nz}")


def stencil27_3d(nx, seed=0, lo=1e-1, hi=1e1, shift=1e-4, leaf=64):
    """3D 27-point stencil, weights log-uniform in [lo, hi], A_ii = sum|A_ij| + shift."""
    n = nx ** 3
    rng = np.random.Generator(np.random.PCG64(seed))
    offs = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
            if (dz, dy, dx) > (0, 0, 0)]
    edges = grid_edges((nx, nx, nx), offs)
    a = np.concatenate([e[0] for e in edges])
    b = np.concatenate([e[1] for e in edges])
    w = np.exp(rng.uniform(np.log(lo), np.log(hi), size=len(a)))
    diag = np.full(n, shift)
    np.add.at(diag, a, w)
    np.add.at(diag, b, w)
    cc, rr, vv = lower_ccs(n, np.concatenate([np.arange(n), b]), np.concatenate([np.arange(n), a]),
                           np.concatenate([diag, -w]))
    return Problem(n, cc, rr, vv, nested_dissection((nx, nx, nx), leaf), f"st27_{nx}")


def block_stencil_3dof(nx, seed=0, leaf=64):
    """3 unknowns per node of an nx^3 grid, 27-point block stencil: for every
    neighbour pair (a, b), A_ab = -w (I + 0.3 R R^T) with R ~ N(0,1)^{3x3} and
    w ~ U(0.5, 2); diagonal blocks D_a = sum_b ||A_ab||_inf I + sum_b w (I +
    0.3 R R^T) (block diagonally dominant, hence SPD).  Unknown 3 p + d is
    dof d of node p; the nested-dissection ordering keeps a node's three
    unknowns together."""
    nn = nx ** 3
    n = 3 * nn
    rng = np.random.Generator(np.random.PCG64(seed))
    offs = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
            if (dz, dy, dx) > (0, 0, 0)]
    edges = grid_edges((nx, nx, nx), offs)
    a = np.concatenate([e[0] for e in edges])
    b = np.concatenate([e[1] for e in edges])
    ne = len(a)
    w = rng.uniform(0.5, 2.0, size=ne)
    R = rng.standard_normal((ne, 3, 3))
    M = np.eye(3)[None] + 0.3 * np.einsum("eij,ekj->eik", R, R)
    blk = -w[:, None, None] * M                      # A_ab, symmetric 3 x 3
    rown = np.abs(blk).sum(axis=2).max(axis=1)       # ||A_ab||_inf
    D = np.zeros((nn, 3, 3))
    add = rown[:, None, None] * np.eye(3)[None] + w[:, None, None] * M
    np.add.at(D, a, add)
    np.add.at(D, b, add)
    # lower triangle: diagonal blocks (i >= j) and the full block at (b, a)
    di, dj = np.tril_indices(3)
    rows = [(3 * np.arange(nn)[:, None] + di[None]).ravel()]
    cols = [(3 * np.arange(nn)[:, None] + dj[None]).ravel()]
    vals = [D[:, di, dj].ravel()]
    ii, jj = np.meshgrid(np.arange(3), np.arange(3), indexing="ij")
    lo_, hi_ = np.minimum(a, b), np.maximum(a, b)
    rows.append((3 * hi_[:, None] + ii.ravel()[None]).ravel())
    cols.append((3 * lo_[:, None] + jj.ravel()[None]).ravel())
    vals.append(blk[:, ii.ravel(), jj.ravel()].ravel())
    cc, rr, vv = lower_ccs(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))
    pn = nested_dissection((nx, nx, nx), leaf) - 1
    perm = (3 * pn[:, None] + np.arange(3)[None]).ravel() + 1
    return Problem(n, cc, rr, vv, perm.astype(np.int32), f"blk3_{nx}")


def anisotropic_laplacian_2d(L, cy, leaf=64):
    """2D 5-point anisotropic Laplacian on an L x L grid, Dirichlet:
    x-coupling 1, y-coupling cy, A_ii = 2 + 2 cy."""
    n = L * L
    (a1, b1), (a2, b2) = grid_edges((L, L), [(1, 0), (0, 1)])
    r = np.concatenate([np.arange(n), b1, b2])
    c = np.concatenate([np.arange(n), a1, a2])
    v = np.concatenate([np.full(n, 2.0 + 2.0 * cy), np.full(len(a1), -1.0), np.full(len(a2), -cy)])
    cc, rr, vv = lower_ccs(n, r, c, v)
    return Problem(n, cc, rr, vv, nested_dissection((L, L), leaf), f"alap2d_{L}_{cy:.3g}")


def batch_anisotropic(nsys=512, seed=0, lo=16, hi=160, leaf=64):
    """The batching workload: nsys independent anisotropic Laplacians, system
    k on an L_k x L_k grid with L_k ~ U{lo..hi} and y-coupling c_y
    log-uniform in [1e-3, 1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Ls = rng.integers(lo, hi + 1, size=nsys)
    cys = np.exp(rng.uniform(np.log(1e-3), 0.0, size=nsys))
    return [anisotropic_laplacian_2d(int(L), float(cy), leaf) for L, cy in zip(Ls, cys)]


def batch_problem(systems):
    """The block-diagonal system of a batch, as one Problem (for checks):
    values and right-hand sides are the systems' concatenated, in the
    order mfd_analyze_batch joins them."""
    n = sum(p.n for p in systems)
    cols, rows, vals = [np.zeros(1, np.int64)], [], []
    off = 0
    eoff = 0
    for p in systems:
        cols.append(np.asarray(p.ccs_col[1:], np.int64) - 1 + eoff)
        rows.append(np.asarray(p.ccs_row, np.int64) + off)
        vals.append(np.asarray(p.val))
        off += p.n
        eoff += len(p.ccs_row)
    cc = (np.concatenate(cols) + 1).astype(np.int32)
    return Problem(n, cc, np.concatenate(rows).astype(np.int32), np.concatenate(vals), None,
                   f"batch_{len(systems)}")


def batch_backward_error(systems, x, b):
    """max over the systems of a batch of each system's backward error
    (x, b: the systems' solutions and right-hand sides stacked)."""
    worst, off = 0.0, 0
    for p in systems:
        worst = max(worst, backward_error(p, x[off:off + p.n], b[off:off + p.n]))
        off += p.n
    return worst


def rhs(problem, nrhs=1, seed=0):
    """x_true ~ U(-1, 1) (n x nrhs) and b = A x_true."""
    rng = np.random.Generator(np.random.PCG64(seed + 12345))
    x = rng.uniform(-1.0, 1.0, size=(problem.n, nrhs))
    return x, problem.matvec(x)


def backward_error(problem, x, b):
    """max over columns of ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)."""
    A = problem.scipy()
    x2 = x.reshape(problem.n, -1)
    b2 = b.reshape(problem.n, -1)
    r = b2 - A @ x2
    an = abs(A).sum(axis=1).max()
    return max(np.abs(r[:, k]).max() / (an * np.abs(x2[:, k]).max() + np.abs(b2[:, k]).max())
               for k in range(x2.shape[1]))
