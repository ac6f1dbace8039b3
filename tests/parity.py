"""Shared checker: runs the oracle on the structure the product's analysis
produced and compares fronts, reference-layout tiles and solutions."""
import numpy as np

from oracle import oracle as O


def oracle_for(solver, values, nb=None):
    """Oracle factorization of the same problem (same structure, same values)."""
    st = solver.structure()
    f = O.OracleFactor(st["supernode_size"], st["work_size"], st["parent"], nb or solver.nb, st["local_index"])
    vals = np.asarray(values, dtype=np.float64)[st["ccs_perm"] - 1]
    f.assemble(st["a_colptr"], st["supernodal_index"], vals)
    info = f.factorize()
    return f, st, info


def front_errors(solver, f, st):
    """Per front: ||F_gpu - F_ref||_F / ||F_ref||_F over the stored lower part."""
    errs = []
    for s in range(1, len(st["supernode_size"]) + 1):
        ref = np.tril(f.front(s))
        got = np.tril(solver.front(s))
        nrm = np.linalg.norm(ref)
        errs.append(np.linalg.norm(got - ref) / (nrm if nrm > 0 else 1.0))
    return np.array(errs)


def tile_mask_compare(solver, f):
    """Max relative difference of the exported reference-layout arrays against
    the oracle's, over the entries both define (the lower part of diagonal
    tiles, all of off-diagonal ones).  Returns (max_rel_diff, entries)."""
    sn, wk = solver.export_tiles()
    worst, count = 0.0, 0
    st = solver.structure()
    for which, arr in ((1, sn), (2, wk)):
        ref = f.array(which)
        mask = np.zeros(len(ref), bool)
        for s in range(1, len(st["supernode_size"]) + 1):
            B = O.lib().orc_num_matrix_block(f.h, s)
            ws = O.lib().orc_work_start(f.h, s)
            c = int(st["supernode_size"][s - 1])
            nbk = f.nb
            divisible = c % nbk == 0
            for j in range(1, B + 1):
                in_sn = j <= ws - 1 if divisible else j <= ws
                in_wk = j > ws - 1 if divisible else j >= ws
                if (which == 1 and not in_sn) or (which == 2 and not in_wk):
                    continue
                for i in range(j, B + 1):
                    off, rows, cols = f.tile(which, s, i, j)
                    if rows <= 0 or cols <= 0:
                        continue
                    m = np.ones((rows, cols), bool)
                    if i == j:
                        if which == 1:
                            m = np.tril(m)
                        else:
                            m = np.tril(m)
                    mask[off:off + rows * cols] |= m.ravel()
        if mask.any():
            scale = max(np.abs(ref[mask]).max(), 1e-300)
            worst = max(worst, np.abs(arr[mask] - ref[mask]).max() / scale)
            count += int(mask.sum())
    return worst, count
