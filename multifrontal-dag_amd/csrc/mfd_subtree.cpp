// Subtree partition of the assembly tree (tree_pruning_m.f90:13-155): the
// tree is cut into subtrees whose work can be spread over p workers with
// the lightest worker carrying more than tol of the heaviest's work.
//
// Work of a front (count_op_count, :76-88): sum over m = w+1 .. w+c of m^2,
// the reference's cost formula; a subtree's work is the sum over its fronts
// (generate_tree_weights, :90-110).  Starting from the root, the heaviest
// listed subtree (the first of equal weight in list order) is replaced by
// its children, each added at the head of the list (partition_tree,
// :112-134; doubly_linked_list_m's add), until the list holds ip subtrees;
// the subtrees are then dealt out largest first to the least loaded of ip
// workers, and ip grows from p until the lightest/heaviest load ratio
// exceeds tol (:37-60).  When the list runs out (every subtree a single
// leaf), the partition for ip - 1 workers with tol = 0 is returned
// (:43-47).  Weights are 64-bit here (the reference's are default
// integers).  A forest starts from all of its roots, the last root at the
// head of the list; with one root that is the reference's start from node n.
#include <algorithm>
#include <cstdint>
#include <list>
#include <queue>
#include <vector>

#include "mfdag.h"

namespace {

struct Pruner {
    int n;
    std::vector<int64_t> wt;                // subtree work
    std::vector<std::vector<int>> child;    // ascending
    std::vector<int> roots;

    // The partition for ip workers grown from the current list; false when
    // the list runs out before it holds ip subtrees.
    bool partition(std::list<int>& lst, int ip) const
    {
        while ((int)lst.size() < ip) {
            auto best = lst.end();
            int64_t bw = 0;
            for (auto it = lst.begin(); it != lst.end(); ++it)
                if (wt[*it] > bw) bw = wt[*it], best = it;
            if (best == lst.end()) return false;
            const int node = *best;
            for (int ch : child[node]) lst.push_front(ch);
            lst.erase(best);
        }
        return true;
    }

    // lightest / heaviest load of ip workers filled largest subtree first
    static double balance(const std::list<int>& lst, const std::vector<int64_t>& wt, int ip, std::vector<int>* bin_of)
    {
        std::vector<std::pair<int64_t, int>> ops;  // (weight, position in the list)
        int pos = 0;
        for (int v : lst) ops.push_back({wt[v], pos++});
        std::stable_sort(ops.begin(), ops.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        std::priority_queue<std::pair<int64_t, int>, std::vector<std::pair<int64_t, int>>, std::greater<>> bins;
        for (int b = 0; b < ip; ++b) bins.push({0, b});
        if (bin_of) bin_of->assign(lst.size(), 0);
        for (const auto& op : ops) {
            auto top = bins.top();
            bins.pop();
            if (bin_of) (*bin_of)[op.second] = top.second;
            bins.push({top.first + op.first, top.second});
        }
        int64_t lo = bins.top().first, hi = 0;
        while (!bins.empty()) hi = std::max(hi, bins.top().first), bins.pop();
        return hi > 0 ? (double)lo / (double)hi : 0.0;
    }

    // tree_pruning (:13-74)
    std::list<int> prune(double tol, int p, int& workers) const
    {
        std::list<int> lst;
        for (int r : roots) lst.push_front(r);
        double ratio = 0.0;
        int ip = p - 1;
        while (ratio <= tol) {
            ++ip;
            if (!partition(lst, ip)) return prune(0.0, ip - 1, workers);
            ratio = balance(lst, wt, ip, nullptr);
        }
        workers = ip;
        return lst;
    }
};

}  // namespace

extern "C" int mfd_tree_pruning(int num_node, const int* c, const int* w, const int* parent, double tol, int p,
                                int* subtrees, int* bins, int* count)
{
    if (num_node < 1 || !c || !w || !parent || !subtrees || !count || p < 1 || !(tol >= 0.0) || !(tol < 1.0))
        return MFD_ERR_ARG;
    Pruner pr;
    pr.n = num_node;
    pr.wt.assign(num_node, 0);
    pr.child.assign(num_node, {});
    for (int s = 0; s < num_node; ++s) {
        if (c[s] < 1 || w[s] < 0) return MFD_ERR_ARG;
        const int pa = parent[s];
        if (pa == 0)
            pr.roots.push_back(s);
        else if (pa <= s + 1 || pa > num_node)  // postorder: parents after children
            return MFD_ERR_ARG;
        else
            pr.child[pa - 1].push_back(s);
    }
    auto sq = [](int64_t x) { return x * (x + 1) * (2 * x + 1) / 6; };
    for (int s = 0; s < num_node; ++s) {
        pr.wt[s] += sq((int64_t)w[s] + c[s]) - sq(w[s]);
        if (parent[s] > 0) pr.wt[parent[s] - 1] += pr.wt[s];
    }
    if (pr.roots.empty()) return MFD_ERR_ARG;
    int workers = 0;
    const std::list<int> lst = pr.prune(tol, p, workers);
    std::vector<int> bin_of;
    Pruner::balance(lst, pr.wt, workers, &bin_of);
    int k = 0;
    for (int v : lst) {
        subtrees[k] = v + 1;
        if (bins) bins[k] = bin_of[k];
        ++k;
    }
    *count = k;
    return MFD_OK;
}
