// Host symbolic analysis, giving the reference's integer structure bit for bit.
//
// Each step below names the reference routine whose result it reproduces
// (paths relative to the reference root, src/main/).  The algorithms are
// written for linear or near-linear time (the reference's bubble sorts and
// recursive helpers are replaced), but wherever the reference's output depends
// on an iteration order -- postorder stack, linked-list merge order, the `<=`
// tie-break of choose_merge_node, the stack order of create_perm -- that order
// is kept exactly.
#include "mfd_analysis.h"

#include <algorithm>
#include <climits>
#include <numeric>

#include "mfdag.h"

namespace mfd {
namespace {

// A jagged array of 1-based integers: list k (0-based) is val[ptr[k] .. ptr[k+1]).
struct Jag {
    std::vector<int64_t> ptr{0};
    std::vector<int> val;
    int lists() const { return (int)ptr.size() - 1; }
    int len(int k) const { return (int)(ptr[k + 1] - ptr[k]); }
    const int* at(int k) const { return val.data() + ptr[k]; }
    int* at(int k) { return val.data() + ptr[k]; }
    static Jag with_lengths(const std::vector<int>& lens)
    {
        Jag j;
        j.ptr.resize(lens.size() + 1);
        j.ptr[0] = 0;
        for (size_t k = 0; k < lens.size(); ++k) j.ptr[k + 1] = j.ptr[k] + lens[k];
        j.val.assign(j.ptr.back(), 0);
        return j;
    }
};

std::vector<int> inverse(const std::vector<int>& perm)  // perm_m.f90:11-23
{
    std::vector<int> ip(perm.size());
    for (size_t i = 0; i < perm.size(); ++i) ip[perm[i] - 1] = (int)i + 1;
    return ip;
}

// ccs_to_crs_m.f90:9-43: row lists in ascending column order
Jag transpose(const Jag& ccs)
{
    int n = ccs.lists();
    std::vector<int> cnt(n, 0);
    for (int v : ccs.val) cnt[v - 1]++;
    Jag crs = Jag::with_lengths(cnt);
    std::vector<int64_t> pos(crs.ptr.begin(), crs.ptr.end() - 1);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < ccs.len(j); ++k) {
            int r = ccs.at(j)[k];
            crs.val[pos[r - 1]++] = j + 1;
        }
    return crs;
}

// compute_tree_m.f90:10-35.  The elimination tree is unique; it is computed
// here with the path-halving ancestor array of Liu's algorithm.  Row i lists
// its columns k < i (the diagonal is last and skipped, as there).
std::vector<int> etree(const Jag& crs)
{
    int n = crs.lists();
    std::vector<int> parent(n + 1, 0), anc(n + 1, 0);
    for (int i = 1; i <= n; ++i) {
        const int* cols = crs.at(i - 1);
        int m = crs.len(i - 1) - 1;
        for (int t = 0; t < m; ++t) {
            int k = cols[t];
            while (k != 0 && k < i) {
                int next = anc[k];
                anc[k] = i;
                if (next == 0) parent[k] = i;
                k = next;
            }
        }
    }
    return std::vector<int>(parent.begin() + 1, parent.end());
}

// tree_m.f90:15-37 and :39-55: child lists in ascending order
Jag children_of(const std::vector<int>& parent)
{
    int n = (int)parent.size();
    std::vector<int> cnt(n, 0);
    for (int i = 0; i < n; ++i)
        if (parent[i] > 0) cnt[parent[i] - 1]++;
    Jag tc = Jag::with_lengths(cnt);
    std::vector<int64_t> pos(tc.ptr.begin(), tc.ptr.end() - 1);
    for (int i = 0; i < n; ++i)
        if (parent[i] > 0) tc.val[pos[parent[i] - 1]++] = i + 1;
    return tc;
}

// tree_m.f90:78-99 (tree_traverse_postordering): stack from the last node,
// perm filled from the end; child subtrees therefore appear in stored order.
std::vector<int> postorder(const Jag& tc)
{
    int n = tc.lists();
    std::vector<int> perm(n), stack;
    stack.reserve(n);
    stack.push_back(n);
    int ptr = n;
    while (!stack.empty()) {
        int node = stack.back();
        stack.pop_back();
        perm[--ptr] = node;
        for (int k = 0; k < tc.len(node - 1); ++k) stack.push_back(tc.at(node - 1)[k]);
    }
    return perm;
}

// tree_m.f90:57-76 (create_parent_in_postordering_tree)
std::vector<int> parent_in_postorder(const std::vector<int>& num_child)
{
    int n = (int)num_child.size();
    std::vector<int> parent(n, 0), stack;
    for (int i = 1; i <= n; ++i) {
        for (int j = 0; j < num_child[i - 1]; ++j) {
            parent[stack.back() - 1] = i;
            stack.pop_back();
        }
        stack.push_back(i);
    }
    return parent;
}

// reordering_m.f90:31-96 (repostordering_ccs1/2): column perm(c) becomes c,
// rows renumbered in place (not re-sorted).  src (optional) receives, per new
// entry, the 1-based flat index of its source entry.
Jag repostorder(const Jag& ccs, const std::vector<int>& perm, const std::vector<int>& iperm,
                std::vector<int64_t>* src)
{
    int n = ccs.lists();
    std::vector<int> lens(n);
    for (int c = 0; c < n; ++c) lens[c] = ccs.len(perm[c] - 1);
    Jag out = Jag::with_lengths(lens);
    if (src) src->assign(out.val.size(), 0);
    for (int co = 0; co < n; ++co) {
        int cr = iperm[co] - 1;
        for (int k = 0; k < ccs.len(co); ++k) {
            out.at(cr)[k] = iperm[ccs.at(co)[k] - 1];
            if (src) (*src)[out.ptr[cr] + k] = ccs.ptr[co] + k + 1;
        }
    }
    return out;
}

// reordering_m.f90:98-148 (reordering_ccs): symmetric permutation of a lower
// CCS; each new column's rows are sorted ascending with their sources.
Jag reorder(const Jag& ccs, const std::vector<int>& iperm, std::vector<int64_t>& src)
{
    int n = ccs.lists();
    std::vector<int> cnt(n, 0);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < ccs.len(j); ++k)
            cnt[std::min(iperm[j], iperm[ccs.at(j)[k] - 1]) - 1]++;
    Jag out = Jag::with_lengths(cnt);
    src.assign(out.val.size(), 0);
    std::vector<int64_t> pos(out.ptr.begin(), out.ptr.end() - 1);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < ccs.len(j); ++k) {
            int a = iperm[j], b = iperm[ccs.at(j)[k] - 1];
            int col = std::min(a, b) - 1;
            out.val[pos[col]] = std::max(a, b);
            src[pos[col]] = ccs.ptr[j] + k + 1;
            pos[col]++;
        }
    std::vector<std::pair<int, int64_t>> tmp;
    for (int c = 0; c < n; ++c) {
        tmp.clear();
        for (int64_t p = out.ptr[c]; p < out.ptr[c + 1]; ++p) tmp.emplace_back(out.val[p], src[p]);
        std::sort(tmp.begin(), tmp.end(), [](auto& x, auto& y) { return x.first < y.first; });
        for (size_t t = 0; t < tmp.size(); ++t) {
            out.val[out.ptr[c] + t] = tmp[t].first;
            src[out.ptr[c] + t] = tmp[t].second;
        }
    }
    return out;
}

// Tarjan-style set labels: find() returns the most recent node a set was
// linked into, which is what disjoint_set_m.f90:32-69 returns.
struct Labels {
    std::vector<int> up;
    explicit Labels(int n) : up(n + 1) { std::iota(up.begin(), up.end(), 0); }
    int find(int x)
    {
        int r = x;
        while (up[r] != r) r = up[r];
        while (up[x] != r) {
            int nx = up[x];
            up[x] = r;
            x = nx;
        }
        return r;
    }
};

// column_count_m.f90:9-72 (postordered matrix and tree)
std::vector<int> column_counts(const Jag& ccs, const Jag& tc, const std::vector<int>& parent)
{
    int n = (int)parent.size();
    Labels lab(n);
    std::vector<int> lca(ccs.val.size(), 0), prev(n + 1, 0);
    for (int j = 1; j <= n - 1; ++j) {
        for (int k = 1; k < ccs.len(j - 1); ++k) {
            int r = ccs.at(j - 1)[k];
            int pp = prev[r];
            lca[ccs.ptr[j - 1] + k] = pp ? lab.find(pp) : 0;
            prev[r] = j;
        }
        lab.up[j] = parent[j - 1];
    }
    std::vector<int> wt(n + 1, 0);
    for (int i = 1; i <= n; ++i) wt[i] = tc.len(i - 1) == 0 ? 1 : 0;
    for (int j = 1; j <= n; ++j) {
        if (j != n) wt[parent[j - 1]]--;
        for (int k = 1; k < ccs.len(j - 1); ++k) {
            wt[j]++;
            int l = lca[ccs.ptr[j - 1] + k];
            if (l) wt[l]--;
        }
    }
    std::vector<int> cc(wt.begin() + 1, wt.end());
    for (int i = 1; i <= n - 1; ++i) cc[parent[i - 1] - 1] += cc[i - 1];
    return cc;
}

// Array-backed linked lists with the exact add / merge / remove semantics of
// doubly_linked_lists_m.f90:34-91 (merge appends without fixing prev, as there).
struct Lists {
    std::vector<int> head, tail, next, prev;
    explicit Lists(int n) : head(n + 1, 0), tail(n + 1, 0), next(n + 1, 0), prev(n + 1, 0) {}
    void add(int node, int idx)
    {
        if (head[idx] == 0) {
            head[idx] = tail[idx] = node;
        } else {
            prev[node] = tail[idx];
            next[tail[idx]] = node;
            tail[idx] = node;
        }
    }
    void merge(int from, int to)
    {
        if (head[from] == 0) return;
        if (head[to] == 0) {
            head[to] = head[from];
            tail[to] = tail[from];
        } else {
            next[tail[to]] = head[from];
            tail[to] = tail[from];
        }
        head[from] = tail[from] = 0;
    }
    void remove(int node, int idx)
    {
        int nx = next[node], pv = prev[node];
        if (node == head[idx]) head[idx] = nx; else next[pv] = nx;
        if (node == tail[idx]) tail[idx] = pv; else prev[nx] = pv;
        next[node] = prev[node] = 0;
    }
    template <class F> void each(int idx, F f) const
    {
        for (int x = head[idx]; x != 0; x = next[x]) f(x);
    }
    int length(int idx) const
    {
        int c = 0;
        each(idx, [&](int) { ++c; });
        return c;
    }
};

}  // namespace

int analyze(int n, const int* ccs_col, const int* ccs_row, int max_zero, const int* perm_in, Structure& S)
{
    if (n < 1 || !ccs_col || !ccs_row) return MFD_ERR_ARG;
    if (ccs_col[0] != 1) return MFD_ERR_ARG;
    Jag A;
    A.ptr.resize(n + 1);
    for (int j = 0; j <= n; ++j) A.ptr[j] = (int64_t)ccs_col[j] - 1;
    for (int j = 0; j < n; ++j)
        if (A.ptr[j + 1] <= A.ptr[j]) return MFD_ERR_ARG;  // every column holds its diagonal
    A.val.assign(ccs_row, ccs_row + A.ptr[n]);
    for (int j = 0; j < n; ++j) {
        if (A.at(j)[0] != j + 1) return MFD_ERR_ARG;  // diagonal first
        for (int k = 1; k < A.len(j); ++k)
            if (A.at(j)[k] <= j + 1 || A.at(j)[k] > n) return MFD_ERR_ARG;  // strictly lower
    }

    // optional ordering in place of METIS (hppllt.f90:192-199)
    std::vector<int> rperm;
    std::vector<int64_t> reorder_src;
    Jag R;
    if (perm_in) {
        rperm.assign(perm_in, perm_in + n);
        std::vector<int> seen(n + 1, 0);
        for (int v : rperm) {
            if (v < 1 || v > n || seen[v]) return MFD_ERR_ARG;
            seen[v] = 1;
        }
        R = reorder(A, inverse(rperm), reorder_src);
    } else {
        R = A;
    }

    // ---- analyze_phase (analyse_phase_m.f90:22-71)
    Jag crs = transpose(R);
    std::vector<int> et = etree(crs);
    for (int i = 0; i < n - 1; ++i)
        if (et[i] == 0) return MFD_ERR_REDUCIBLE;  // the reference assumes one tree rooted at n
    Jag tc0 = children_of(et);
    std::vector<int> pperm = postorder(tc0);
    std::vector<int> piperm = inverse(pperm);
    Jag P = repostorder(R, pperm, piperm, nullptr);
    std::vector<int> pparent(n, 0);  // reordering_m.f90:17-29
    for (int i = 1; i < n; ++i) pparent[i - 1] = piperm[et[pperm[i - 1] - 1] - 1];
    Jag ptc = children_of(pparent);
    std::vector<int> cc_node = column_counts(P, ptc, pparent);

    // finding_supernode (analyse_phase_m.f90:73-89)
    std::vector<int> sub(n, 1);  // tree_m.f90:101-116
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < ptc.len(i); ++k) sub[i] += sub[ptc.at(i)[k] - 1];
    std::vector<int> isleaf(n, 0), prev_nz(n + 1, 0);  // finding_leaves_m.f90:6-29
    for (int j = 1; j <= n; ++j)
        for (int k = 1; k < P.len(j - 1); ++k) {
            int r = P.at(j - 1)[k];
            if (prev_nz[r] < j - sub[j - 1] + 1) isleaf[j - 1] = 1;
            prev_nz[r] = j;
        }
    std::vector<int> flen;  // fundamental_supernode_m.f90:11-38
    for (int i = 0; i < n; ++i) {
        if (isleaf[i] > 0 || ptc.len(i) > 1 || flen.empty()) flen.push_back(0);
        flen.back()++;
    }
    int nf = (int)flen.size();
    std::vector<int> ffirst(nf + 1, 1);
    for (int s = 0; s < nf; ++s) ffirst[s + 1] = ffirst[s] + flen[s];
    std::vector<int> fnum_child(nf), fcc(nf);
    for (int s = 0; s < nf; ++s) {
        fnum_child[s] = ptc.len(ffirst[s] - 1);         // fundamental_supernode_m.f90:55-67
        fcc[s] = cc_node[ffirst[s + 1] - 2] - 1;        // fundamental_supernode_m.f90:40-53
    }
    std::vector<int> fparent = parent_in_postorder(fnum_child);
    Jag ftc = children_of(fparent);

    // relax_supernode (analyse_phase_m.f90:91-108; relax_supernode_m.f90:13-90)
    Lists sons(nf), merge(nf);
    for (int i = 1; i <= nf; ++i) merge.add(i, i);
    std::vector<int> num_cols(nf + 1, 0), num_zeros(nf + 1, 0), stack;
    for (int i = 1; i <= nf; ++i) num_cols[i] = flen[i - 1];
    for (int k = 1; k <= nf; ++k) {
        if (ftc.len(k - 1) == 0) continue;
        for (int t = 0; t < ftc.len(k - 1); ++t) sons.add(ftc.at(k - 1)[t], k);
        for (;;) {
            int min_zero = max_zero, pick = 0;
            sons.each(k, [&](int c) {
                int64_t z = (int64_t)(fcc[k - 1] + num_cols[k] - fcc[c - 1]) * num_cols[c] + num_zeros[k] + num_zeros[c];
                if (z <= min_zero && z <= max_zero) {
                    pick = c;
                    min_zero = (int)z;
                }
            });
            if (pick) {
                merge.merge(pick, k);
                sons.remove(pick, k);
                stack.push_back(pick);
                num_zeros[k] = min_zero;
                num_cols[k] += num_cols[pick];
            } else {
                while (!stack.empty()) {
                    sons.merge(stack.back(), k);
                    stack.pop_back();
                }
                break;
            }
        }
    }
    // relaxed_supernode_m.f90:15-112
    std::vector<int> map;
    for (int i = 1; i <= nf; ++i)
        if (merge.head[i] != 0) map.push_back(i);
    int nr = (int)map.size();
    std::vector<int> rlen(nr, 0), rnum_child(nr), rcc(nr), relax_perm;
    relax_perm.reserve(n);
    for (int i = 0; i < nr; ++i) {
        stack.clear();
        merge.each(map[i], [&](int x) {
            rlen[i] += flen[x - 1];
            stack.push_back(x);
        });
        while (!stack.empty()) {
            int x = stack.back();
            stack.pop_back();
            for (int j = ffirst[x - 1]; j < ffirst[x]; ++j) relax_perm.push_back(j);
        }
        rnum_child[i] = sons.length(map[i]);
        rcc[i] = fcc[map[i] - 1];
    }
    std::vector<int> rparent = parent_in_postorder(rnum_child);
    Jag rtc = children_of(rparent);
    std::vector<int> riperm = inverse(relax_perm);
    Jag RC = repostorder(P, relax_perm, riperm, nullptr);
    std::vector<int> rfirst(nr + 1, 1);
    for (int s = 0; s < nr; ++s) rfirst[s + 1] = rfirst[s] + rlen[s];

    // create_supernodal_ccs (fundamental_supernode_m.f90:70-117) and
    // symbolic_factorize (symbolic_factorize_m.f90:10-62): the row list of a
    // supernode is the set union of its columns' rows and its children's rows
    // below it, then sorted (sort_m.f90:9-21).
    Jag L = Jag::with_lengths(rcc);
    std::vector<int> mark(n + 1, 0);
    for (int s = 1; s <= nr; ++s) {
        int last = rfirst[s] - 1, cnt = 0;
        int* out = L.at(s - 1);
        for (int j = rfirst[s - 1]; j <= last; ++j)
            for (int k = 0; k < RC.len(j - 1); ++k) {
                int r = RC.at(j - 1)[k];
                if (r <= last || mark[r] == s) continue;
                if (cnt >= rcc[s - 1]) return MFD_ERR_INTERNAL;
                mark[r] = s;
                out[cnt++] = r;
            }
        for (int t = 0; t < rtc.len(s - 1); ++t) {
            int c = rtc.at(s - 1)[t];
            for (int k = 0; k < L.len(c - 1); ++k) {
                int r = L.at(c - 1)[k];
                if (r <= last || mark[r] == s) continue;
                if (cnt >= rcc[s - 1]) return MFD_ERR_INTERNAL;
                mark[r] = s;
                out[cnt++] = r;
            }
        }
        if (cnt != rcc[s - 1]) return MFD_ERR_INTERNAL;
        std::sort(out, out + cnt);
    }
    std::vector<int> aperm(n);  // perm_product (perm_m.f90:25-36)
    for (int i = 0; i < n; ++i) aperm[i] = pperm[relax_perm[i] - 1];

    // ---- hppllt_analyze (hppllt.f90:201-216)
    std::vector<int> aiperm = inverse(aperm);
    std::vector<int64_t> asrc;
    Jag AS = repostorder(R, aperm, aiperm, &asrc);
    S = Structure();
    S.n = n;
    S.num_node = nr;
    if (perm_in) {
        S.perm.resize(n);
        for (int i = 0; i < n; ++i) S.perm[i] = rperm[aperm[i] - 1];
        S.ccs_perm.resize(asrc.size());
        for (size_t k = 0; k < asrc.size(); ++k) S.ccs_perm[k] = reorder_src[asrc[k] - 1];
    } else {
        S.perm = aperm;
        S.ccs_perm = asrc;
    }
    S.a_colptr = AS.ptr;
    S.c = rlen;
    S.w.resize(nr);
    for (int s = 0; s < nr; ++s) S.w[s] = L.len(s);
    S.parent.assign(nr, 0);  // tree_m.f90:118-132 (create_parent)
    S.children.assign(nr, {});
    for (int s = 0; s < nr; ++s)
        for (int t = 0; t < rtc.len(s); ++t) {
            S.parent[rtc.at(s)[t] - 1] = s + 1;
            S.children[s].push_back(rtc.at(s)[t]);
        }
    S.first = rfirst;
    S.l_ptr = L.ptr;
    S.l_rows = L.val;

    // create_supernodal_index (create_supernodal_index_m.f90:10-40)
    std::vector<int> pos(n + 1, 0);
    S.snidx.assign(AS.val.size(), 0);
    for (int s = 1; s <= nr; ++s) {
        int f = rfirst[s - 1], c = rlen[s - 1];
        for (int j = 0; j < c; ++j) pos[f + j] = j + 1;
        for (int k = 0; k < L.len(s - 1); ++k) pos[L.at(s - 1)[k]] = c + k + 1;
        for (int j = f; j < f + c; ++j)
            for (int64_t p = AS.ptr[j - 1]; p < AS.ptr[j]; ++p) S.snidx[p] = pos[AS.val[p]];
    }
    // create_local_index (create_local_index_m.f90:10-45)
    S.li_ptr = L.ptr;
    S.lidx.assign(L.val.size(), 0);
    for (int s = 1; s <= nr; ++s) {
        if (S.children[s - 1].empty()) continue;
        int f = rfirst[s - 1], c = rlen[s - 1];
        for (int j = 0; j < c; ++j) pos[f + j] = j + 1;
        for (int k = 0; k < L.len(s - 1); ++k) pos[L.at(s - 1)[k]] = c + k + 1;
        for (int ch : S.children[s - 1])
            for (int k = 0; k < L.len(ch - 1); ++k) S.lidx[L.ptr[ch - 1] + k] = pos[L.at(ch - 1)[k]];
    }
    return MFD_OK;
}

int structure_from_parts(int num_node, const int* c, const int* w, const int* parent, const int* lidx,
                         const int64_t* a_colptr, const int* snidx, Structure& S)
{
    if (num_node < 1 || !c || !w || !parent) return MFD_ERR_ARG;
    S = Structure();
    S.num_node = num_node;
    S.c.assign(c, c + num_node);
    S.w.assign(w, w + num_node);
    S.parent.assign(parent, parent + num_node);
    S.children.assign(num_node, {});
    S.first.assign(num_node + 1, 1);
    S.li_ptr.assign(num_node + 1, 0);
    for (int s = 0; s < num_node; ++s) {
        if (S.c[s] < 1 || S.w[s] < 0) return MFD_ERR_ARG;
        int p = S.parent[s];
        if (p != 0 && (p <= s + 1 || p > num_node)) return MFD_ERR_ARG;  // postorder: parent after child
        if (p != 0) S.children[p - 1].push_back(s + 1);
        if (p == 0 && S.w[s] != 0) return MFD_ERR_ARG;
        S.first[s + 1] = S.first[s] + S.c[s];
        S.li_ptr[s + 1] = S.li_ptr[s] + S.w[s];
    }
    S.n = S.first[num_node] - 1;
    S.l_ptr = S.li_ptr;
    S.lidx.assign(lidx, lidx + S.li_ptr[num_node]);
    for (int s = 0; s < num_node; ++s) {
        int p = S.parent[s];
        if (!p) continue;
        int np = S.c[p - 1] + S.w[p - 1], last = 0;
        for (int64_t k = S.li_ptr[s]; k < S.li_ptr[s + 1]; ++k) {
            int v = S.lidx[k];
            if (v <= last || v > np) return MFD_ERR_ARG;  // strictly increasing, inside the parent
            last = v;
        }
    }
    // row lists: update row k of s is the global row of its parent-front
    // position (parents come after children, so walk down from the roots)
    S.l_rows.assign(S.li_ptr[num_node], 0);
    for (int s = num_node - 1; s >= 0; --s) {
        int p = S.parent[s];
        if (!p) continue;
        for (int64_t k = S.li_ptr[s]; k < S.li_ptr[s + 1]; ++k) {
            int pos = S.lidx[k], cp = S.c[p - 1];
            S.l_rows[k] = pos <= cp ? S.first[p - 1] + pos - 1 : S.l_rows[S.li_ptr[p - 1] + (pos - cp - 1)];
        }
    }
    S.perm.resize(S.n);
    std::iota(S.perm.begin(), S.perm.end(), 1);
    if (a_colptr && snidx) {
        S.a_colptr.assign(a_colptr, a_colptr + S.n + 1);
        S.snidx.assign(snidx, snidx + S.a_colptr[S.n]);
        S.ccs_perm.resize(S.snidx.size());
        std::iota(S.ccs_perm.begin(), S.ccs_perm.end(), 1);
        for (int s = 0; s < num_node; ++s) {
            int n = S.c[s] + S.w[s];
            for (int j = S.first[s]; j < S.first[s + 1]; ++j)
                for (int64_t q = S.a_colptr[j - 1]; q < S.a_colptr[j]; ++q) {
                    int r = S.snidx[q];
                    if (r < j - S.first[s] + 1 || r > n) return MFD_ERR_ARG;  // lower part of the front
                }
        }
    }
    return MFD_OK;
}

int analyze_batch(int nsys, const int* n, const int* const* ccs_col, const int* const* ccs_row, int max_zero,
                  const int* const* perm, Structure& out)
{
    if (nsys < 1 || !n || !ccs_col || !ccs_row) return MFD_ERR_ARG;
    out = Structure();
    int64_t entries = 0;  // caller entries of the systems before this one
    for (int k = 0; k < nsys; ++k) {
        if (!ccs_col[k] || !ccs_row[k]) return MFD_ERR_ARG;
        Structure S;
        int rc = analyze(n[k], ccs_col[k], ccs_row[k], max_zero, perm ? perm[k] : nullptr, S);
        if (rc != MFD_OK) return rc;
        if ((int64_t)out.n + S.n > INT32_MAX - 1) return MFD_ERR_ARG;
        const int col_off = out.n, node_off = out.num_node;
        const int64_t a_off = (int64_t)out.snidx.size(), l_off = (int64_t)out.l_rows.size(),
                      li_off = (int64_t)out.lidx.size();
        for (int v : S.perm) out.perm.push_back(v + col_off);
        for (int64_t v : S.ccs_perm) out.ccs_perm.push_back(v + entries);
        for (int j = 0; j < S.n; ++j) out.a_colptr.push_back(S.a_colptr[j] + a_off);
        out.snidx.insert(out.snidx.end(), S.snidx.begin(), S.snidx.end());
        out.c.insert(out.c.end(), S.c.begin(), S.c.end());
        out.w.insert(out.w.end(), S.w.begin(), S.w.end());
        for (int v : S.parent) out.parent.push_back(v ? v + node_off : 0);
        for (int s = 0; s < S.num_node; ++s) {
            out.first.push_back(S.first[s] + col_off);
            out.l_ptr.push_back(S.l_ptr[s] + l_off);
            out.li_ptr.push_back(S.li_ptr[s] + li_off);
            std::vector<int> ch(S.children[s]);
            for (int& v : ch) v += node_off;
            out.children.push_back(std::move(ch));
        }
        for (int v : S.l_rows) out.l_rows.push_back(v + col_off);
        out.lidx.insert(out.lidx.end(), S.lidx.begin(), S.lidx.end());
        out.n += S.n;
        out.num_node += S.num_node;
        entries += (int64_t)ccs_col[k][n[k]] - 1;
    }
    out.a_colptr.push_back((int64_t)out.snidx.size());
    out.first.push_back(out.n + 1);
    out.l_ptr.push_back((int64_t)out.l_rows.size());
    out.li_ptr.push_back((int64_t)out.lidx.size());
    return MFD_OK;
}

int64_t cost(const Structure& s)
{
    int64_t total = 0;
    for (int k = 0; k < s.num_node; ++k)
        for (int64_t m = s.w[k] + 1; m <= (int64_t)s.w[k] + s.c[k]; ++m) total += m * m;
    return total;
}

int64_t nonzero(const Structure& s)
{
    int64_t total = 0;
    for (int k = 0; k < s.num_node; ++k)
        for (int64_t m = s.w[k] + 1; m <= (int64_t)s.w[k] + s.c[k]; ++m) total += m;
    return total;
}

}  // namespace mfd
