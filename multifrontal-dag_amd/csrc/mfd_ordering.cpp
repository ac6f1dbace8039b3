// Fill-reducing ordering for use_metis = 1 without a caller ordering: graph
// nested dissection in place of METIS_NodeND (called at hppllt.f90:47-51
// through ordering_m.f90:10-38, which returns perm new -> old).
//
// The adjacency graph of A (both triangles, no self loops) is split
// recursively by vertex separators taken from breadth-first level
// structures:
//   * a pseudo-peripheral root (George and Liu, "Computer Solution of Large
//     Sparse Positive Definite Systems", 1981, sec. 4.3.2): repeated
//     breadth-first searches from a least-degree vertex of the last level
//     while the eccentricity grows;
//   * the separator is one level of the root's level structure, the
//     smallest among the levels that leave both sides at least 30% of the
//     vertices (the median level when none does); every edge joins equal or
//     adjacent levels, so a level separates the levels before it from those
//     after it;
//   * the separator is thinned: a separator vertex with no neighbour on one
//     side joins the other side;
//   * the two sides are ordered first (recursively), the separator last;
//     disconnected pieces are ordered one after another, and pieces of at
//     most kLeaf vertices keep their natural order.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "mfdag.h"

namespace {

constexpr int kLeaf = 64;

struct Dissector {
    int n;
    std::vector<int64_t> xadj;
    std::vector<int> adj;
    std::vector<int> part;   // id of the piece each vertex currently belongs to
    std::vector<int> level;  // scratch: BFS level (-1: not reached)
    std::vector<int> queue;
    int next_id = 1;
    int* perm;               // output, 1-based new -> old

    // Breadth-first search from r over vertices with part == id; fills
    // queue (in visiting order) and level, returns the number of levels.
    int bfs(int r, int id, int& count)
    {
        count = 0;
        queue[count++] = r;
        level[r] = 0;
        int depth = 0;
        for (int h = 0; h < count; ++h) {
            const int v = queue[h];
            depth = level[v];
            for (int64_t e = xadj[v]; e < xadj[v + 1]; ++e) {
                const int u = adj[e];
                if (part[u] == id && level[u] < 0) {
                    level[u] = level[v] + 1;
                    queue[count++] = u;
                }
            }
        }
        return depth + 1;
    }

    void clear_levels(int count)
    {
        for (int h = 0; h < count; ++h) level[queue[h]] = -1;
    }

    int degree_in(int v, int id) const
    {
        int d = 0;
        for (int64_t e = xadj[v]; e < xadj[v + 1]; ++e) d += part[adj[e]] == id;
        return d;
    }

    // Writes verts (one piece, all with part == id) to perm[first ..) with
    // its separators after the parts they separate.
    void dissect(std::vector<int>& verts, int id, int first)
    {
        const int m = (int)verts.size();
        if (m <= kLeaf) {
            std::sort(verts.begin(), verts.end());
            for (int k = 0; k < m; ++k) perm[first + k] = verts[k] + 1;
            return;
        }
        // connected pieces, each ordered on its own
        int count = 0;
        bfs(verts[0], id, count);
        if (count < m) {
            clear_levels(count);
            std::vector<std::vector<int>> pieces;
            for (int v : verts) {
                if (part[v] != id) continue;
                bfs(v, id, count);
                pieces.emplace_back(queue.begin(), queue.begin() + count);
                clear_levels(count);
                const int pid = next_id++;
                for (int u : pieces.back()) part[u] = pid;
            }
            std::vector<int>().swap(verts);
            int at = first;
            for (auto& piece : pieces) {
                const int sz = (int)piece.size();
                dissect(piece, part[piece[0]], at);
                at += sz;
            }
            return;
        }
        clear_levels(count);
        // pseudo-peripheral root
        int root = verts[0], best = degree_in(root, id);
        for (int v : verts) {
            const int d = degree_in(v, id);
            if (d < best) best = d, root = v;
        }
        int depth = bfs(root, id, count);
        for (int it = 0; it < 8; ++it) {
            const int last = level[queue[count - 1]];
            int cand = -1, cdeg = 0;
            for (int h = count - 1; h >= 0 && level[queue[h]] == last; --h) {
                const int d = degree_in(queue[h], id);
                if (cand < 0 || d < cdeg) cand = queue[h], cdeg = d;
            }
            clear_levels(count);
            const int d2 = bfs(cand, id, count);
            if (d2 <= depth) {
                clear_levels(count);
                depth = bfs(root, id, count);
                break;
            }
            root = cand;
            depth = d2;
        }
        if (depth < 3) {  // no level leaves two non-empty sides
            clear_levels(count);
            std::sort(verts.begin(), verts.end());
            for (int k = 0; k < m; ++k) perm[first + k] = verts[k] + 1;
            return;
        }
        std::vector<int> width(depth, 0);
        for (int h = 0; h < count; ++h) ++width[level[queue[h]]];
        int sep = -1, median = -1;
        for (int k = 0, cum = 0; k < depth; ++k) {
            if (median < 0 && cum + width[k] > m / 2) median = k;
            if (k >= 1 && k <= depth - 2) {
                const int after = m - cum - width[k];
                if (10 * cum >= 3 * m && 10 * after >= 3 * m && (sep < 0 || width[k] < width[sep])) sep = k;
            }
            cum += width[k];
        }
        if (sep < 0) sep = std::min(std::max(median, 1), depth - 2);
        int before = 0;
        for (int k = 0; k < sep; ++k) before += width[k];
        if (10 * std::max(before, m - before - width[sep]) > 9 * m) {  // no useful split
            clear_levels(count);
            std::sort(verts.begin(), verts.end());
            for (int k = 0; k < m; ++k) perm[first + k] = verts[k] + 1;
            return;
        }
        // sides: a = levels before sep, b = after, s = the separator level
        const int ida = next_id++, idb = next_id++, ids = next_id++;
        std::vector<int> va, vb, vs;
        for (int h = 0; h < count; ++h) {
            const int v = queue[h], l = level[v];
            if (l < sep)
                part[v] = ida, va.push_back(v);
            else if (l > sep)
                part[v] = idb, vb.push_back(v);
            else
                part[v] = ids, vs.push_back(v);
        }
        clear_levels(count);
        std::vector<int>().swap(verts);
        // thinning
        std::vector<int> keep;
        for (int v : vs) {
            bool nb_b = false, nb_a = false;
            for (int64_t e = xadj[v]; e < xadj[v + 1]; ++e) {
                nb_b |= part[adj[e]] == idb;
                nb_a |= part[adj[e]] == ida;
            }
            if (!nb_b)
                part[v] = ida, va.push_back(v);
            else if (!nb_a)
                part[v] = idb, vb.push_back(v);
            else
                keep.push_back(v);
        }
        const int na = (int)va.size(), nb = (int)vb.size();
        dissect(va, ida, first);
        dissect(vb, idb, first + na);
        std::sort(keep.begin(), keep.end());
        for (size_t k = 0; k < keep.size(); ++k) perm[first + na + nb + (int)k] = keep[k] + 1;
    }
};

}  // namespace

extern "C" int mfd_order_nd(int n, const int* ccs_col, const int* ccs_row, int* perm)
{
    if (n < 1 || !ccs_col || !ccs_row || !perm) return MFD_ERR_ARG;
    if (ccs_col[0] != 1) return MFD_ERR_ARG;
    for (int j = 0; j < n; ++j)
        if (ccs_col[j + 1] < ccs_col[j]) return MFD_ERR_ARG;
    Dissector d;
    d.n = n;
    d.xadj.assign(n + 1, 0);
    for (int j = 0; j < n; ++j)
        for (int p = ccs_col[j] - 1; p < ccs_col[j + 1] - 1; ++p) {
            const int i = ccs_row[p] - 1;
            if (i < 0 || i >= n) return MFD_ERR_ARG;
            if (i == j) continue;
            ++d.xadj[i + 1];
            ++d.xadj[j + 1];
        }
    for (int v = 0; v < n; ++v) d.xadj[v + 1] += d.xadj[v];
    d.adj.resize(d.xadj[n]);
    std::vector<int64_t> fill(d.xadj.begin(), d.xadj.end() - 1);
    for (int j = 0; j < n; ++j)
        for (int p = ccs_col[j] - 1; p < ccs_col[j + 1] - 1; ++p) {
            const int i = ccs_row[p] - 1;
            if (i == j) continue;
            d.adj[fill[i]++] = j;
            d.adj[fill[j]++] = i;
        }
    d.part.assign(n, 0);
    d.level.assign(n, -1);
    d.queue.assign(n, 0);
    d.perm = perm;
    std::vector<int> all(n);
    for (int v = 0; v < n; ++v) all[v] = v;
    d.dissect(all, 0, 0);
    return MFD_OK;
}
