// Fill-reducing ordering for use_metis = 1 without a caller ordering: graph
// nested dissection in place of METIS_NodeND (called at hppllt.f90:47-51
// through ordering_m.f90:10-38, which returns perm new -> old).
//
// The adjacency graph of A (both triangles, no self loops) is split
// recursively by vertex separators; the two sides are ordered first
// (recursively), the separator last; disconnected pieces are ordered one
// after another, and pieces of at most kLeaf vertices keep their natural
// order.  A connected piece gets several candidate separators (pieces of
// fewer than kMultilevelMin vertices only the first) and keeps the smallest
// of those that leave both sides at least a fifth of the piece:
//   * a breadth-first level: the level structure of a pseudo-peripheral
//     root (George and Liu, "Computer Solution of Large Sparse Positive
//     Definite Systems", 1981, sec. 4.3.2: repeated searches from a
//     least-degree vertex of the last level while the eccentricity grows),
//     the smallest level that leaves both sides at least 30% of the
//     vertices (the median level when none does); every edge joins equal or
//     adjacent levels, so a level separates the levels before it from those
//     after it;
//   * multilevel edge bisections (Karypis and Kumar, "A fast and high
//     quality multilevel scheme for partitioning irregular graphs", 1998),
//     three, each from its own pseudo-random sequence: heavy-edge matching
//     down to about 120 vertices, a bisection grown breadth-first from
//     several starts, Fiduccia-Mattheyses passes at every level on the way
//     back (sides of at most 55% of the weight), and a minimum vertex cover
//     of the cut edges (Koenig's theorem from a maximum matching, Hopcroft
//     and Karp) as the separator.
// Every candidate is thinned first: a separator vertex with no neighbour on
// one side joins the other side.  All choices are deterministic.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <queue>
#include <vector>

#include "mfdag.h"

namespace {

constexpr int kLeaf = 64;
constexpr int kCoarsest = 120;
constexpr int kTrials = 3;          // multilevel bisections tried per piece
constexpr int kMultilevelMin = 400;  // smaller pieces take the level separator

// A weighted graph (CSR, 0-based) local to one piece.
struct WGraph {
    int n = 0;
    std::vector<int64_t> xadj;
    std::vector<int> adj, ew, vw;
};

struct Lcg {
    uint32_t s;
    uint32_t next()
    {
        s = s * 1664525u + 1013904223u;
        return s >> 8;
    }
};

// Heavy-edge matching: c is g with matched pairs merged (weights summed),
// cmap maps g's vertices to c's.  False when it would shrink g by < 15%.
bool coarsen(const WGraph& g, WGraph& c, std::vector<int>& cmap, Lcg& rng)
{
    const int n = g.n;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    for (int i = n - 1; i > 0; --i) std::swap(order[i], order[rng.next() % (uint32_t)(i + 1)]);
    std::vector<int> match(n, -1);
    for (int v : order) {
        if (match[v] >= 0) continue;
        int best = -1, bw = -1;
        for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
            const int u = g.adj[e];
            if (u != v && match[u] < 0 && g.ew[e] > bw) best = u, bw = g.ew[e];
        }
        if (best >= 0) {
            match[v] = best;
            match[best] = v;
        } else {
            match[v] = v;
        }
    }
    cmap.assign(n, -1);
    int nc = 0;
    std::vector<int> m1, m2;
    m1.reserve(n);
    m2.reserve(n);
    for (int v = 0; v < n; ++v) {
        if (cmap[v] >= 0) continue;
        cmap[v] = cmap[match[v]] = nc++;
        m1.push_back(v);
        m2.push_back(match[v] != v ? match[v] : -1);
    }
    if (20 * (int64_t)nc > 17 * (int64_t)n) return false;
    c.n = nc;
    c.vw.assign(nc, 0);
    c.xadj.assign(nc + 1, 0);
    c.adj.clear();
    c.ew.clear();
    c.adj.reserve(g.adj.size());
    c.ew.reserve(g.adj.size());
    std::vector<int64_t> pos(nc, -1);
    for (int k = 0; k < nc; ++k) {
        const int64_t start = (int64_t)c.adj.size();
        for (int v : {m1[k], m2[k]}) {
            if (v < 0) continue;
            c.vw[k] += g.vw[v];
            for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
                const int ck = cmap[g.adj[e]];
                if (ck == k) continue;
                if (pos[ck] >= start) {
                    c.ew[pos[ck]] += g.ew[e];
                } else {
                    pos[ck] = (int64_t)c.adj.size();
                    c.adj.push_back(ck);
                    c.ew.push_back(g.ew[e]);
                }
            }
        }
        c.xadj[k + 1] = (int64_t)c.adj.size();
    }
    return true;
}

int64_t cut_of(const WGraph& g, const std::vector<char>& side)
{
    int64_t cut = 0;
    for (int v = 0; v < g.n; ++v)
        for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
            if (side[g.adj[e]] != side[v]) cut += g.ew[e];
    return cut / 2;
}

// Fiduccia-Mattheyses passes on a two-way edge cut: moves of the best gain
// that keep each side at most maxw, the best prefix of each pass kept.  The
// external and internal edge weights of every vertex are kept up to date
// across moves and passes, so a pass costs its moves, not the graph.
void fm_refine(const WGraph& g, std::vector<char>& side, int passes)
{
    const int n = g.n;
    int64_t w[2] = {0, 0}, maxvw = 0;
    for (int v = 0; v < n; ++v) w[(int)side[v]] += g.vw[v], maxvw = std::max<int64_t>(maxvw, g.vw[v]);
    const int64_t total = w[0] + w[1], maxw = std::max(total * 55 / 100, total / 2 + maxvw);
    std::vector<int64_t> ext(n, 0), in(n, 0);
    for (int v = 0; v < n; ++v)
        for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) (side[g.adj[e]] != side[v] ? ext[v] : in[v]) += g.ew[e];
    std::vector<char> locked(n, 0);
    std::vector<int> moved;
    // v changes side; its neighbours' weights follow
    auto flip = [&](int v) {
        const int from = side[v], to = 1 - from;
        side[v] = (char)to;
        w[from] -= g.vw[v];
        w[to] += g.vw[v];
        std::swap(ext[v], in[v]);
        for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
            const int u = g.adj[e];
            if (u == v) continue;
            if (side[u] == to)
                ext[u] -= g.ew[e], in[u] += g.ew[e];
            else
                ext[u] += g.ew[e], in[u] -= g.ew[e];
        }
    };
    for (int pass = 0; pass < passes; ++pass) {
        std::priority_queue<std::pair<int64_t, int>> heap;
        for (int v = 0; v < n; ++v)
            if (ext[v] > 0) heap.push({ext[v] - in[v], -v});
        moved.clear();
        int64_t cum = 0, best = 0;
        size_t best_k = 0;
        int since = 0;
        while (!heap.empty() && since < 200) {
            const auto top = heap.top();
            heap.pop();
            const int v = -top.second;
            if (locked[v] || top.first != ext[v] - in[v]) continue;
            locked[v] = 1;
            moved.push_back(v);
            if (w[1 - side[v]] + g.vw[v] > maxw) {
                moved.pop_back();
                continue;
            }
            cum += ext[v] - in[v];
            flip(v);
            if (cum > best) {
                best = cum;
                best_k = moved.size();
                since = 0;
            } else {
                ++since;
            }
            for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
                const int u = g.adj[e];
                if (!locked[u]) heap.push({ext[u] - in[u], -u});
            }
        }
        for (size_t k = moved.size(); k > best_k; --k) flip(moved[k - 1]);
        for (int v = 0; v < n; ++v) locked[v] = 0;
        if (best <= 0) break;
    }
}

// A bisection of a small graph: side 0 grown breadth-first from a start
// vertex to half the weight, refined; the smallest cut of a few starts.
void grow_bisection(const WGraph& g, std::vector<char>& side, Lcg& rng)
{
    const int n = g.n;
    const int64_t total = std::accumulate(g.vw.begin(), g.vw.end(), (int64_t)0);
    std::vector<char> trial(n);
    std::vector<int> queue;
    int64_t best = -1;
    const int tries = std::min(n, 3);
    for (int t = 0; t < tries; ++t) {
        const int start = t == 0 ? 0 : (int)(rng.next() % (uint32_t)n);
        std::fill(trial.begin(), trial.end(), 1);
        int64_t w0 = g.vw[start];
        queue.assign(1, start);
        trial[start] = 0;
        for (size_t h = 0; w0 < total / 2;) {
            if (h == queue.size()) {  // piece exhausted: continue from any vertex on side 1
                int v = 0;
                while (v < n && trial[v] == 0) ++v;
                if (v == n) break;
                trial[v] = 0;
                w0 += g.vw[v];
                queue.push_back(v);
                continue;
            }
            const int v = queue[h++];
            for (int64_t e = g.xadj[v]; e < g.xadj[v + 1] && w0 < total / 2; ++e) {
                const int u = g.adj[e];
                if (trial[u] == 0) continue;
                trial[u] = 0;
                w0 += g.vw[u];
                queue.push_back(u);
            }
        }
        fm_refine(g, trial, 4);
        const int64_t cut = cut_of(g, trial);
        if (best < 0 || cut < best) best = cut, side = trial;
    }
}

// Minimum vertex cover of the bipartite graph of the cut edges (left: the
// boundary of side 0, right: that of side 1) by Koenig's theorem from a
// maximum matching (Hopcroft and Karp).  adj[i]: right neighbours of left
// vertex i.  Returns the cover as in_cover flags for left and right.
struct CutCover {
    std::vector<std::vector<int>> adj;
    int nl = 0, nr = 0;
    std::vector<int> match_l, match_r, dist;

    bool bfs()
    {
        std::vector<int> q;
        bool found = false;
        dist.assign(nl, -1);
        for (int u = 0; u < nl; ++u)
            if (match_l[u] < 0) dist[u] = 0, q.push_back(u);
        for (size_t h = 0; h < q.size(); ++h) {
            const int u = q[h];
            for (int v : adj[u]) {
                const int w = match_r[v];
                if (w < 0)
                    found = true;
                else if (dist[w] < 0)
                    dist[w] = dist[u] + 1, q.push_back(w);
            }
        }
        return found;
    }
    bool dfs(int u)
    {
        for (int v : adj[u]) {
            const int w = match_r[v];
            if (w < 0 || (dist[w] == dist[u] + 1 && dfs(w))) {
                match_l[u] = v;
                match_r[v] = u;
                return true;
            }
        }
        dist[u] = -1;
        return false;
    }
    void cover(std::vector<char>& cl, std::vector<char>& cr)
    {
        match_l.assign(nl, -1);
        match_r.assign(nr, -1);
        while (bfs())
            for (int u = 0; u < nl; ++u)
                if (match_l[u] < 0) dfs(u);
        // alternating reachability from the free left vertices
        std::vector<char> zl(nl, 0), zr(nr, 0);
        std::vector<int> q;
        for (int u = 0; u < nl; ++u)
            if (match_l[u] < 0) zl[u] = 1, q.push_back(u);
        for (size_t h = 0; h < q.size(); ++h)
            for (int v : adj[q[h]]) {
                if (zr[v]) continue;
                zr[v] = 1;
                const int w = match_r[v];
                if (w >= 0 && !zl[w]) zl[w] = 1, q.push_back(w);
            }
        cl.assign(nl, 0);
        cr.assign(nr, 0);
        for (int u = 0; u < nl; ++u) cl[u] = !zl[u];
        for (int v = 0; v < nr; ++v) cr[v] = zr[v];
    }
};

// Multilevel edge bisection of g (side 0 / 1 per vertex).
void ml_bisect(const WGraph& g0, std::vector<char>& side, Lcg& rng)
{
    std::vector<WGraph> levels(1, g0);
    std::vector<std::vector<int>> maps;
    while (levels.back().n > kCoarsest) {
        WGraph c;
        std::vector<int> cmap;
        if (!coarsen(levels.back(), c, cmap, rng)) break;
        maps.push_back(std::move(cmap));
        levels.push_back(std::move(c));
    }
    grow_bisection(levels.back(), side, rng);
    for (int l = (int)maps.size() - 1; l >= 0; --l) {
        std::vector<char> fine(levels[l].n);
        for (int v = 0; v < levels[l].n; ++v) fine[v] = side[maps[l][v]];
        side.swap(fine);
        fm_refine(levels[l], side, 8);
    }
}

struct Dissector {
    int n;
    std::vector<int64_t> xadj;
    std::vector<int> adj;
    std::vector<int> part;   // id of the piece each vertex currently belongs to
    std::vector<int> level;  // scratch: BFS level (-1: not reached)
    std::vector<int> queue;
    std::vector<int> loc;    // scratch: vertex -> index in the current piece (-1: none)
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

    void natural(std::vector<int>& verts, int first)
    {
        std::sort(verts.begin(), verts.end());
        for (size_t k = 0; k < verts.size(); ++k) perm[first + (int)k] = verts[k] + 1;
    }

    // Breadth-first level separator of a connected piece: lab[i] = 0 / 1
    // (sides) / 2 (separator) for verts[i]; false when no level splits it.
    bool level_separator(const std::vector<int>& verts, int id, std::vector<char>& lab)
    {
        const int m = (int)verts.size();
        int count = 0;
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
        bool ok = depth >= 3;
        int sep = -1;
        if (ok) {
            std::vector<int> width(depth, 0);
            for (int h = 0; h < count; ++h) ++width[level[queue[h]]];
            int median = -1;
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
            ok = 10 * std::max(before, m - before - width[sep]) <= 9 * m;
        }
        if (ok)
            for (int h = 0; h < count; ++h) {
                const int l = level[queue[h]];
                lab[loc[queue[h]]] = l < sep ? 0 : (l > sep ? 1 : 2);
            }
        clear_levels(count);
        return ok;
    }

    // The piece as a weighted graph of local indices (loc must be set).
    void local_graph(const std::vector<int>& verts, int id, WGraph& g)
    {
        const int m = (int)verts.size();
        g.n = m;
        g.vw.assign(m, 1);
        g.xadj.assign(m + 1, 0);
        int64_t edges = 0;
        for (int v : verts) edges += xadj[v + 1] - xadj[v];
        g.adj.clear();
        g.ew.clear();
        g.adj.reserve(edges);
        g.ew.reserve(edges);
        for (int i = 0; i < m; ++i) {
            const int v = verts[i];
            for (int64_t e = xadj[v]; e < xadj[v + 1]; ++e)
                if (part[adj[e]] == id) {
                    g.adj.push_back(loc[adj[e]]);
                    g.ew.push_back(1);
                }
            g.xadj[i + 1] = (int64_t)g.adj.size();
        }
    }

    // Multilevel separator of a connected piece given as g (labels as
    // above); trial selects the pseudo-random sequence of the matching and
    // the initial starts.
    bool multilevel_separator(const WGraph& g, int trial, std::vector<char>& lab)
    {
        const int m = g.n;
        Lcg rng{(uint32_t)m * 2654435761u + 12345u + 977u * (uint32_t)trial};
        std::vector<char> side;
        ml_bisect(g, side, rng);
        // the separator: a minimum vertex cover of the cut edges
        std::vector<int> bidx(m, -1), bl, br;
        for (int i = 0; i < m; ++i)
            for (int64_t e = g.xadj[i]; e < g.xadj[i + 1]; ++e)
                if (side[g.adj[e]] != side[i]) {
                    std::vector<int>& b = side[i] == 0 ? bl : br;
                    bidx[i] = (int)b.size();
                    b.push_back(i);
                    break;
                }
        if (bl.empty()) return false;
        CutCover cc;
        cc.nl = (int)bl.size();
        cc.nr = (int)br.size();
        cc.adj.resize(cc.nl);
        for (int k = 0; k < cc.nl; ++k) {
            const int i = bl[k];
            for (int64_t e = g.xadj[i]; e < g.xadj[i + 1]; ++e)
                if (side[g.adj[e]] != side[i]) cc.adj[k].push_back(bidx[g.adj[e]]);
        }
        std::vector<char> cl, cr;
        cc.cover(cl, cr);
        for (int i = 0; i < m; ++i) lab[i] = side[i];
        for (int k = 0; k < cc.nl; ++k)
            if (cl[k]) lab[bl[k]] = 2;
        for (int k = 0; k < cc.nr; ++k)
            if (cr[k]) lab[br[k]] = 2;
        return true;
    }

    // A separator vertex with no neighbour on one side joins the other.
    void thin(const std::vector<int>& verts, std::vector<char>& lab)
    {
        for (size_t i = 0; i < verts.size(); ++i) {
            if (lab[i] != 2) continue;
            bool nb_a = false, nb_b = false;
            const int v = verts[i];
            for (int64_t e = xadj[v]; e < xadj[v + 1]; ++e) {
                const int j = loc[adj[e]];
                if (j < 0) continue;
                nb_a |= lab[j] == 0;
                nb_b |= lab[j] == 1;
            }
            if (!nb_b)
                lab[i] = 0;
            else if (!nb_a)
                lab[i] = 1;
        }
    }

    // Writes verts (one piece, all with part == id) to perm[first ..) with
    // its separators after the parts they separate.
    void dissect(std::vector<int>& verts, int id, int first)
    {
        const int m = (int)verts.size();
        if (m <= kLeaf) return natural(verts, first);
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
        for (int i = 0; i < m; ++i) loc[verts[i]] = i;
        // candidates: the level separator and kTrials multilevel ones; kept:
        // a balanced one (smaller side >= m / 5) before an unbalanced one,
        // then the smallest separator
        std::vector<char> lab, cand(m);
        int best_ns = 0;
        bool best_bal = false;
        const int trials = m >= kMultilevelMin ? kTrials : 0;
        WGraph g;
        if (trials > 0) local_graph(verts, id, g);
        for (int t = -1; t < trials; ++t) {
            const bool ok = t < 0 ? level_separator(verts, id, cand) : multilevel_separator(g, t, cand);
            if (!ok) continue;
            thin(verts, cand);
            int na = 0, nb = 0, ns = 0;
            for (char l : cand) na += l == 0, nb += l == 1, ns += l == 2;
            // a split leaving one side over 90% would let the recursion run
            // as deep as the piece is large
            if (na == 0 || nb == 0 || 10 * std::max(na, nb) > 9 * m) continue;
            const bool bal = 5 * std::min(na, nb) >= m;
            if (lab.empty() || (bal != best_bal ? bal : ns < best_ns)) lab = cand, best_ns = ns, best_bal = bal;
        }
        for (int v : verts) loc[v] = -1;
        if (lab.empty()) return natural(verts, first);
        const int ida = next_id++, idb = next_id++, ids = next_id++;
        std::vector<int> va, vb, vs;
        for (int i = 0; i < m; ++i) {
            const int v = verts[i];
            if (lab[i] == 0)
                part[v] = ida, va.push_back(v);
            else if (lab[i] == 1)
                part[v] = idb, vb.push_back(v);
            else
                part[v] = ids, vs.push_back(v);
        }
        std::vector<int>().swap(verts);
        const int na = (int)va.size(), nb = (int)vb.size();
        dissect(va, ida, first);
        dissect(vb, idb, first + na);
        natural(vs, first + na + nb);
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
    d.loc.assign(n, -1);
    d.perm = perm;
    std::vector<int> all(n);
    for (int v = 0; v < n; ++v) all[v] = v;
    d.dissect(all, 0, 0);
    return MFD_OK;
}
