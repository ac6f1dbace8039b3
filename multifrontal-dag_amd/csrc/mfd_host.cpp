// Host numeric factorization of whole subtrees (see mfd_host.h).
#include "mfd_host.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>

#include "mfdag.h"

namespace mfd {

namespace {

int64_t slot_of(int n) { return ((int64_t)n * n + 31) / 32 * 32; }

// Columns [jj, jj + 8) of the trailing block (rows from jj) minus the panel
// [k0, j0) times its rows' transpose, 64 rows at a time.
void trailing_group(double* A, int n, int k0, int j0, int jj)
{
    const int nq = std::min(8, n - jj);
    for (int i0 = jj; i0 < n; i0 += 64) {
        const int ni = std::min(64, n - i0);
        double acc[8][64];
        for (int q = 0; q < 8; ++q)
            for (int i = 0; i < 64; ++i) acc[q][i] = 0.0;
        for (int p = k0; p < j0; ++p) {
            const double* ap = A + (int64_t)p * n;
            const double* ai = ap + i0;
            double b[8];
            for (int q = 0; q < 8; ++q) b[q] = q < nq ? ap[jj + q] : 0.0;
            for (int q = 0; q < 8; ++q)
                for (int i = 0; i < ni; ++i) acc[q][i] += ai[i] * b[q];
        }
        for (int q = 0; q < nq; ++q) {
            double* col = A + (int64_t)(jj + q) * n;
            const int lo = std::max(0, jj + q - i0);  // lower part only
            for (int i = lo; i < ni; ++i) col[i0 + i] -= acc[q][i];
        }
    }
}

// The partial Cholesky of the first c columns of the n x n column-major
// front A (lower part): 64-column panels, each factored column by column
// (left-looking inside the panel: POTRF + TRSM of factorize_kernel_m.f90:
// 4-16 on the panel's full height), then the trailing columns updated by the
// panel (the SYRK / GEMM of factorize_kernel_m.f90:18-33 and of the work
// tiles, border_subroutines_m.f90:41-73) in groups of 8 columns and 64 rows,
// the groups dealt round-robin to `threads` threads when the update is large.
// Returns -1, or the front column of the first pivot that is not positive.
int factor_front(double* A, int n, int c, int threads)
{
    for (int k0 = 0; k0 < c; k0 += kPanel) {
        const int kb = std::min(kPanel, c - k0);
        for (int j = k0; j < k0 + kb; ++j) {
            double* aj = A + (int64_t)j * n;
            for (int p = k0; p < j; ++p) {
                const double* ap = A + (int64_t)p * n;
                const double ljp = ap[j];
                for (int i = j; i < n; ++i) aj[i] -= ap[i] * ljp;
            }
            const double d = aj[j];
            if (!(d > 0.0)) return j;
            const double ljj = std::sqrt(d), r = 1.0 / ljj;
            aj[j] = ljj;
            for (int i = j + 1; i < n; ++i) aj[i] *= r;
        }
        const int j0 = k0 + kb, groups = (n - j0 + 7) / 8;
        const double flops = (double)(n - j0) * (n - j0) * kb;
        const int nt = flops >= 2e7 ? std::min(threads, groups) : 1;
        auto part = [&](int t) {
            for (int g = t; g < groups; g += nt) trailing_group(A, n, k0, j0, j0 + 8 * g);
        };
        std::vector<std::thread> help;
        for (int t = 1; t < nt; ++t) help.emplace_back(part, t);
        part(0);
        for (auto& t : help) t.join();
    }
    return -1;
}

// inv[r + 64 q] = L^{-1}(q, r) for each 64 x 64 diagonal block of the first
// c columns (zero outside its kb x kb block and above its diagonal): the
// layout of the device's stored inverses, used by the solves.
void block_inverses(const double* A, int n, int c, double* inv)
{
    for (int b = 0; b * kPanel < c; ++b) {
        double* I = inv + (int64_t)b * kPanel * kPanel;
        std::memset(I, 0, sizeof(double) * kPanel * kPanel);
        const int k0 = b * kPanel, kb = std::min(kPanel, c - k0);
        double x[kPanel];
        for (int r = 0; r < kb; ++r) {
            // column r of L^{-1}: L x = e_r by forward substitution
            for (int q = r; q < kb; ++q) {
                double s = q == r ? 1.0 : 0.0;
                for (int p = r; p < q; ++p) s -= A[(k0 + q) + (int64_t)(k0 + p) * n] * x[p];
                x[q] = s / A[(k0 + q) + (int64_t)(k0 + q) * n];
                I[r + kPanel * q] = x[q];
            }
        }
    }
}

// Front f: zero and assemble it, add its host children's update blocks in
// child order (extend_add, extend_add_kernel_m.f90:6-34; a child's update
// block stays in its front after its own factorization), then factor it and
// form its block inverses.  Returns -1 or the 1-based global column of its
// failing pivot.
int factor_one(const HostPlan& p, const Structure& S, const double* a, double* stage, int f, int threads)
{
    const HostFront& h = p.fronts[f];
    double* F = stage + h.front;
    for (int j = 0; j < h.n; ++j) std::memset(F + (int64_t)j * h.n + j, 0, sizeof(double) * (h.n - j));
    for (int64_t e = h.e0; e < h.e1; ++e) stage[p.dst[e]] = a[p.src[e]];
    for (int k = p.child_ptr[f]; k < p.child_ptr[f + 1]; ++k) {
        const HostFront& hc = p.fronts[p.child_pos[k]];
        const double* C = stage + hc.front;
        const int* li = S.lidx.data() + hc.lidx;
        for (int j = 0; j < hc.w; ++j) {
            const double* u = C + (int64_t)(hc.c + j) * hc.n + hc.c;
            double* pc = F + (int64_t)(li[j] - 1) * h.n - 1;
            for (int i = j; i < hc.w; ++i) pc[li[i]] += u[i];
        }
    }
    const int bad = factor_front(F, h.n, h.c, threads);
    if (bad >= 0) return h.first + bad + 1;
    block_inverses(F, h.n, h.c, stage + h.inv);
    return -1;
}

}  // namespace

int build_host_plan(const Structure& S, const std::vector<NodeDev>& nodes, const std::vector<int>& roots,
                    HostPlan& p)
{
    const int m = S.num_node;
    if ((int)nodes.size() != m) return MFD_ERR_STATE;
    p = HostPlan{};
    p.is_host.assign(m, 0);
    p.sub_ptr.assign(1, 0);
    std::vector<int> pos(m, -1);
    std::vector<double> work;
    for (int r : roots) {
        if (r < 0 || r >= m || p.is_host[r]) return MFD_ERR_ARG;
        // the subtree's nodes: ascending ids are a postorder (parent > child)
        std::vector<int> sub, stack(1, r);
        while (!stack.empty()) {
            const int s = stack.back();
            stack.pop_back();
            if (p.is_host[s]) return MFD_ERR_ARG;  // nested roots
            p.is_host[s] = 1;
            sub.push_back(s);
            for (int ch : S.children[s]) stack.push_back(ch - 1);
        }
        std::sort(sub.begin(), sub.end());
        // runs of consecutive node ids: their fronts, then their inverses,
        // are consecutive in the pool, so each run is two staging blocks
        double flops = 0;
        for (size_t q0 = 0; q0 < sub.size();) {
            size_t q1 = q0 + 1;
            while (q1 < sub.size() && sub[q1] == sub[q1 - 1] + 1) ++q1;
            const NodeDev& d0 = nodes[sub[q0]];
            const NodeDev& d1 = nodes[sub[q1 - 1]];
            const HostRun fr{p.stage_elems, d0.front, d1.front + slot_of(d1.n) - d0.front};
            const HostRun ir{fr.stage + fr.count, d0.inv,
                             d1.inv + (int64_t)((d1.c + kPanel - 1) / kPanel) * kPanel * kPanel - d0.inv};
            p.runs.push_back(fr);
            p.runs.push_back(ir);
            p.stage_elems = ir.stage + ir.count;
            for (size_t q = q0; q < q1; ++q) {
                const int s = sub[q];
                const NodeDev& d = nodes[s];
                HostFront h{};
                h.node = s;
                h.n = d.n;
                h.c = d.c;
                h.w = d.w;
                h.first = S.first[s] - 1;
                h.lidx = S.li_ptr[s];
                h.front = fr.stage + (d.front - fr.pool);
                h.inv = ir.stage + (d.inv - ir.pool);
                pos[s] = (int)p.fronts.size();
                p.fronts.push_back(h);
                for (int64_t t = d.n; t > d.w; --t) flops += (double)t * t;
            }
            q0 = q1;
        }
        p.sub_ptr.push_back((int)p.fronts.size());
        work.push_back(flops);
        p.roots.push_back(r);
    }
    for (auto& h : p.fronts) {
        const int par = S.parent[h.node] - 1;
        h.parent_pos = (par >= 0 && p.is_host[par]) ? pos[par] : -1;
    }
    // each front's host children in child order (all of a host front's
    // children are host fronts)
    p.child_ptr.assign(1, 0);
    for (auto& h : p.fronts) {
        for (int ch : S.children[h.node]) p.child_pos.push_back(pos[ch - 1]);
        p.child_ptr.push_back((int)p.child_pos.size());
    }
    // assembly entries of the host fronts (coefficient_setter_m.f90:26-41)
    for (auto& h : p.fronts) {
        const int s = h.node;
        h.e0 = (int64_t)p.dst.size();
        for (int j = S.first[s]; j < S.first[s + 1]; ++j) {
            const int64_t q = j - S.first[s];
            for (int64_t e = S.a_colptr[j - 1]; e < S.a_colptr[j]; ++e) {
                p.dst.push_back(h.front + (S.snidx[e] - 1) + q * h.n);
                p.src.push_back(S.ccs_perm[e] - 1);
            }
        }
        h.e1 = (int64_t)p.dst.size();
    }
    p.entries = (int64_t)p.dst.size();
    p.sub_order.resize(work.size());
    for (size_t k = 0; k < work.size(); ++k) p.sub_order[k] = (int)k;
    std::stable_sort(p.sub_order.begin(), p.sub_order.end(), [&](int x, int y) { return work[x] > work[y]; });
    return MFD_OK;
}

int host_factor(const HostPlan& p, const Structure& S, const double* a, double* stage, int threads)
{
    // A front is ready once its host children are factored (the reference's
    // StarPU graph orders the same accesses, starpu_factorize_m.f90:13-31);
    // workers take ready fronts, heaviest subtree's first, and each front
    // gathers its children's updates itself, so no two workers write one
    // front and the sums do not depend on the schedule.
    const int nf = (int)p.fronts.size();
    if (nf == 0) return 0;
    std::vector<std::atomic<int>> pending(nf);
    std::vector<int> ready;
    for (int k : p.sub_order)
        for (int f = p.sub_ptr[k]; f < p.sub_ptr[k + 1]; ++f) {
            pending[f].store(p.child_ptr[f + 1] - p.child_ptr[f]);
            if (p.child_ptr[f + 1] == p.child_ptr[f]) ready.push_back(f);
        }
    std::reverse(ready.begin(), ready.end());  // taken from the back
    std::mutex mu;
    std::condition_variable cv;
    int done = 0, running = 0, info = INT_MAX;
    auto worker = [&]() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv.wait(lk, [&] { return !ready.empty() || done == nf; });
            if (done == nf) return;
            const int f = ready.back();
            ready.pop_back();
            // the idle workers' share of threads for a front taken when no
            // other front is ready (the subtrees' top fronts)
            const int busy = ++running;
            const int inner = ready.empty() ? std::max(1, threads - busy + 1) : 1;
            lk.unlock();
            const int bad = factor_one(p, S, a, stage, f, inner);
            const int par = p.fronts[f].parent_pos;
            const bool go = par >= 0 && pending[par].fetch_sub(1) == 1;
            lk.lock();
            --running;
            if (bad > 0) info = std::min(info, bad);
            ++done;
            if (go) ready.push_back(par);
            if (go || done == nf) cv.notify_all();
        }
    };
    const int nt = std::max(1, std::min(threads, nf));
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
    return info == INT_MAX ? 0 : info;
}

}  // namespace mfd
