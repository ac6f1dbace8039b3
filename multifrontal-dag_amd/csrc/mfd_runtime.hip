// Runtime of the numeric phase: the handle behind the C-ABI (include/mfdag.h),
// the level-by-level GPU schedule built once per analysis, device memory,
// optional hipGraph replay, and the FP32 + FP64-refinement precision form.
//
// The reference drives the numeric phase node by node in postorder
// (seq_factorize_m.f90:13-29, seq_forward_m.f90:14-32,
// seq_backward_m.f90:14-32) or as a StarPU task DAG
// (starpu_factorize_m.f90:13-31).  Here the elimination tree is cut into
// levels (a node's level is one more than its highest child's), every level
// is a short fixed sequence of batched launches over all of its fronts, and
// the whole sequence is built on the host once and replayed per
// factorization.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "mfd_analysis.h"
#include "mfd_host.h"
#include "mfd_kernels.h"
#include "mfd_layout.h"
#include "mfdag.h"

using namespace mfd;

namespace {

struct Launch {
    int kind;       // MFD_K_*
    int64_t first;  // first task in the kind's task array
    int count;
    int wide = 0;   // GEMM launches: 1 = 128 x 128 tiles (k_gemm_seg + k_gemm_fix)
    // wide launches: K chunks over `groups` workgroups, their segments at
    // seg_ptr[sfirst .. sfirst + groups], then nfix split tiles from fix[xfirst]
    int64_t sfirst = 0, xfirst = 0;
    int groups = 0, nfix = 0;
    // MFD_K_POTRF launches: fused = 1 runs the step's update tiles
    // [first, first + count) of the GEMM tasks beside the diagonal tasks
    // [dfirst, dfirst + dcount) (k_upd_diag); otherwise [first, first +
    // count) are diagonal tasks (k_diag).
    int fused = 0;
    int64_t dfirst = 0;
    int dcount = 0;
    // Large in-panel updates run on the side stream (side = 1, numbered by
    // side_id) beside the next step's launch; wait_id >= 0: this launch
    // first waits for side launch wait_id.  Serial when profiling.
    int side = 0, side_id = -1, wait_id = -1;
    // solve launches: the step they belong to (a step's kFS and kFX
    // launches, and its kBS and kBX launches, run as one launch)
    int step = -1;
};

// A step's update tiles join the step's launch (k_upd_diag, one workgroup
// per CU at its LDS) up to this many; more run as their own launch.
constexpr int kFillGroups = 32;  // most level groups the front zeroing is split into
constexpr int kWholeRows = 128;   // update rows of a one-block front solved by its block task
constexpr int kFuseTiles = 1024;

// Step launches with at least this many diagonal tasks run their GEMM
// tiles (TRSM and fused update tiles) as a second launch (k_gemm).
constexpr int kApartDiag = 512;

// Backward solve steps with at most this many row chunks split each chunk
// into four 16-column tasks (shorter tasks where the step is latency-bound).
constexpr int kSplitChunks = 256;

// The in-panel update of the large fronts runs every `period` steps with
// K = period panels (fewer passes over the trailing columns); the TRSM tiles
// and the diagonal task of a step apply the panels still pending for their
// columns themselves (at most period, 64 columns each).  The handle's period
// defaults to kUpdPeriod; mfd_set_update_period changes it (1..kMaxPeriod, or
// 0: chosen per level from the length of its longest chain).
constexpr int kUpdPeriod = 2;
constexpr int kMaxPeriod = 8;
constexpr int kAutoSteps = 48;  // 3072 columns

// Schur updates of fronts with at least wide_min_rows update rows use
// 128 x 128 tiles with their K chunks divided evenly over the workgroups
// (k_gemm_seg, k_gemm_fix), the others 64 x 64 tiles (k_gemm); set per
// handle by mfd_set_tile_bound, 0 (64 x 64 only) by default: with the
// extend-add in its epilogue the 128 x 128 tile (two workgroups per CU)
// hides its scattered parent accesses worse than four 64 x 64 workgroups
// per CU, and on the 64^3 workload every bound tried (1024, 2048, 3000
// update rows) was slower (factor 36.3, 35.5, 27.7 ms against 21.8 ms).
constexpr int kWideTile = 128;
constexpr int kWideMinRows = 0;

// Fronts with at least super_min pivot columns solve forward and backward
// over 256-column superblocks (k_fwd_super, k_bwd_super: a quarter of the
// dependent launches of the 64-column steps), with their superblock
// matrices formed at the end of the factorization; set per handle by
// mfd_set_solve_block (0: never).  256: on the 64^3 workload the bounds 128,
// 256, 512 and 1024 gave FP64 steps of 25.08, 25.13, 25.17 and 25.39 ms,
// and on the 512-system batch 128 was slower than 512 (FP64 step 21.26
// against 20.90 ms: its fronts of one partial superblock gain nothing).
constexpr int kSuperMin = 256;

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

#define HIP_TRY(expr)                                \
    do {                                             \
        hipError_t e_ = (expr);                      \
        if (e_ != hipSuccess) {                      \
            h->last_hip = e_;                        \
            return MFD_ERR_HIP;                      \
        }                                            \
    } while (0)

#define MFD_TRY(expr)         \
    do {                      \
        int r_ = (expr);      \
        if (r_ != MFD_OK) return r_; \
    } while (0)

}  // namespace

struct mfd_handle {
    int nb = 0, flags = 0;
    int upd_period = kUpdPeriod;
    int level_rule = 1;  // mfd_set_level_rule
    // Medium fronts (order > 64) factored whole by one workgroup each
    // (k_front_wg): those with at most fuse_cmax pivot columns, unless
    // front_kind (one entry per node, when set) says otherwise: 1 = one
    // workgroup, 0 = 64-column steps over the chip, -1 = by fuse_cmax.
    int fuse_cmax = 0;
    std::vector<int> front_kind;
    int wide_min_rows = kWideMinRows;  // Schur tiles of 128 x 128 from this many update rows (0: never)
    // 128 x 128 Schur launches with their K chunks divided evenly over the
    // workgroups: segments, per-launch workgroup ranges, split tiles, and
    // the most partial-product slots one launch needs
    std::vector<SegTask> segs;
    std::vector<int> seg_ptr;
    std::vector<FixTask> fixes;
    int64_t seg_slots = 0;
    SegTask* d_segs = nullptr;
    int* d_seg_ptr = nullptr;
    FixTask* d_fixes = nullptr;
    double* d_segpart = nullptr;  // seg_slots x 16384 entries (FP32 launches use it as floats)
    int super_min = kSuperMin;         // forward superblocks from this many pivot columns (0: never)
    bool device_ok = false;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // large in-panel updates beside the diagonal chain
    hipEvent_t fork = nullptr, sev[2] = {};
    // Front-pool zeroing and assembly in groups of levels beside the
    // factorization: group g's fronts (fill tasks [fill_gptr[g],
    // fill_gptr[g+1]), entries [scat_gptr[g], scat_gptr[g+1])) are first
    // written by a launch of level group_level[g] or later; the main stream
    // waits for fill_ev[g] before the first launch of that level
    // (level_launch).
    hipStream_t fillst = nullptr;
    std::vector<hipEvent_t> fill_ev;
    std::vector<int> fill_gptr, group_level, level_launch, front_group;
    std::vector<int64_t> scat_gptr;
    hipError_t last_hip = hipSuccess;
    int64_t mem_limit = 0, device_bytes = 0, front_bytes = 0;
    // Device buffers in use (pointer, bytes) and released ones kept for
    // reuse.  A buffer that a captured hipGraph's kernels used is not given
    // back to the device by hipFree after the graph is destroyed (measured:
    // every schedule rebuild lost the front pool), so buffers are freed only
    // by mfd_destroy and otherwise reused by later allocations of the handle.
    // device_bytes counts both.
    std::vector<std::pair<void*, int64_t>> allocs, spare;

    bool analyzed = false, factored = false, from_parts = false;
    int precision = MFD_FP64;
    // 16-bit operand mode of the current FP32-pool factorization: 0 (FP32
    // fronts), 1 (MFD_FP16_IR), 2 (MFD_BF16_IR); the 16-bit forms factor
    // the diagonally scaled matrix S A S (scale vector d_dscale).
    int lowp = 0;
    Structure S;
    RefLayout* layout = nullptr;

    // host schedule
    std::vector<NodeDev> nodes;
    std::vector<ExtTask> ext;
    std::vector<int> small;
    std::vector<int> fused;             // medium fronts factored whole by one workgroup
    std::vector<int> fused_pos;         // node -> position in fused, or -1
    std::vector<DiagTask> diag;         // diagonal-block factorizations of large fronts
    std::vector<int2> fill;             // (node, first column) blocks of the pool's lower triangles
    std::vector<GemmTask> gemm;
    std::vector<Launch> fseq;
    std::vector<int> level_ptr, level_nodes, child_ptr, child_list;
    // solve schedule: kinds kFI/kFA/kFS (forward) and kBG/kBS (backward)
    std::vector<int2> fi, fa, bg;  // (front, first row) tasks of the gathers
    int2 *d_fi = nullptr, *d_fa = nullptr, *d_bg = nullptr;
    std::vector<SolveTask> fstep, bstep;
    std::vector<Launch> fwd_seq, bwd_seq;
    int64_t part_slots = 0;
    SolveTask *d_fstep = nullptr, *d_bstep = nullptr;
    // superblock forward steps (kind kFX), backward top tasks (kind kBX)
    // and the blocks of the superblock matrices: X by block distance i - j
    // ([xinv_ptr[d], xinv_ptr[d + 1]), d = 0..3), then W and V
    // ([xinv_ptr[4], xinv_ptr[5]))
    std::vector<SuperTask> xstep, xbstep;
    std::vector<SuperInvTask> xinv;
    int xinv_ptr[6] = {};
    int64_t sinv_elems = 0;
    SuperTask *d_xstep = nullptr, *d_xbstep = nullptr;
    SuperInvTask* d_xinv = nullptr;
    double* d_sinv64 = nullptr;
    float* d_sinv32 = nullptr;
    double* d_part = nullptr;
    int64_t pool_elems = 0, z_per_rhs = 0;
    int num_levels = 0;

    // device copies
    NodeDev* d_nodes = nullptr;
    int* d_lidx = nullptr;
    ExtTask* d_ext = nullptr;
    int* d_small = nullptr;
    int* d_fused = nullptr;
    long long* d_clocks = nullptr;  // start/end clock of each medium front's workgroup, in fused's order
    DiagTask* d_diag = nullptr;
    int2* d_fill = nullptr;
    GemmTask* d_gemm = nullptr;
    int *d_level_nodes = nullptr, *d_child_ptr = nullptr, *d_child_list = nullptr;
    int64_t *d_dst = nullptr, *d_src = nullptr;
    int2* d_scat_rc = nullptr;       // global (row, col) of each scattered entry, in d_dst's order
    int64_t* d_diag_src = nullptr;   // caller index of A(j,j), j postordered
    double* d_dscale = nullptr;      // diagonal scaling of the 16-bit forms
    int* d_perm = nullptr;  // new -> old, 0-based
    double* d_pool64 = nullptr;
    float* d_pool32 = nullptr;
    double* d_a = nullptr;  // caller-order values (nnz)
    int* d_info = nullptr;
    // solve buffers
    int nrhs_cap = 0;
    double *d_z = nullptr, *d_bin = nullptr, *d_bp = nullptr, *d_xp = nullptr;
    // refinement
    int *d_csr_ptr = nullptr, *d_csr_col = nullptr;
    int64_t* d_csr_src = nullptr;  // index into the caller-order values of each CSR entry
    double* d_csr_val = nullptr;   // the values in CSR order (gathered with ||A||_inf)
    int64_t csr_nnz = 0;
    double *d_r = nullptr, *d_d = nullptr;
    unsigned long long* d_norms = nullptr;
    double anorm = -1;  // ||A||_inf of the factored values (< 0: not formed yet)
    // mfd_solve_device_noalloc: device address of the fixed scratch array on
    // this handle's GPU (nullptr if it could not be resolved), and, during
    // such a call, no_alloc = true (no allocation, no graph capture; the
    // solve buffers above then point into the scratch array).
    double* scratch = nullptr;
    bool no_alloc = false;

    hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};  // FP64, FP32, FP16, BF16 factorizations
    struct SolveGraph {
        int prec, nrhs;
        const double* bp;
        double* xp;
        hipGraphExec_t exec;
    };
    std::vector<SolveGraph> sgraphs;
    hipEvent_t ev[4] = {};
    std::vector<hipEvent_t> prof_ev;

    // errors of the last solve, per right-hand side (backward_error)
    int solved_nrhs = 0;  // nrhs of the last solve; 0: no solve since the last factorization
    std::vector<double> err_eta, err_res2, err_ferr;

    double times[8] = {0};
    mfd_stats st{};
    double ktime[MFD_K_NUM] = {0};
    int kcount[MFD_K_NUM] = {0};

    // CPU+GPU hybrid: the subtrees rooted at host_roots (0-based) are
    // assembled and factored by host threads (mfd_host.cpp) into hstage,
    // which is copied into the pool before fseq[host_join] -- the first
    // launch of the lowest level with a parent of a host subtree root.
    std::vector<int> host_roots;
    int host_threads = 1;
    HostPlan hplan;
    int host_rc = 0;                   // build_host_plan's result for the current schedule
    int host_join = -1;
    double* hstage = nullptr;          // hplan.stage_elems doubles (pinned with a device)
    std::vector<double> hstage_vec;    // the staging array without a device
    const double* host_vals = nullptr; // caller-order values for the host part
    std::vector<double> host_vals_buf; // ... copied from the device for mfd_factorize_device
    std::thread host_thr;
    int host_info = 0;
    bool host_factored = false;        // hstage holds the host fronts of the current values
    double host_ms = 0;
};

namespace {

template <typename T> int dalloc(mfd_handle* h, T** p, int64_t n)
{
    if (h->no_alloc) return MFD_ERR_SCRATCH;
    size_t bytes = (size_t)std::max<int64_t>(n, 1) * sizeof(T);
    // the smallest released buffer that holds it and is at most twice its size
    int best = -1;
    for (size_t i = 0; i < h->spare.size(); ++i) {
        const int64_t b = h->spare[i].second;
        if (b >= (int64_t)bytes && b <= 2 * (int64_t)bytes && (best < 0 || b < h->spare[best].second)) best = (int)i;
    }
    if (best >= 0) {
        *p = static_cast<T*>(h->spare[best].first);
        h->allocs.push_back(h->spare[best]);
        h->spare.erase(h->spare.begin() + best);
        return MFD_OK;
    }
    if (h->mem_limit > 0 && h->device_bytes + (int64_t)bytes > h->mem_limit) return MFD_ERR_NOMEM;
    if (hipMalloc((void**)p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return MFD_ERR_NOMEM;
    }
    h->device_bytes += (int64_t)bytes;
    h->allocs.push_back({(void*)*p, (int64_t)bytes});
    return MFD_OK;
}

template <typename T> int upload(mfd_handle* h, T** p, const std::vector<T>& v)
{
    MFD_TRY(dalloc(h, p, (int64_t)v.size()));
    if (!v.empty()) HIP_TRY(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return MFD_OK;
}

void drop_solve_graphs(mfd_handle* h)
{
    for (auto& g : h->sgraphs) (void)hipGraphExecDestroy(g.exec);
    h->sgraphs.clear();
}

void free_device(mfd_handle* h)
{
    if (h->host_thr.joinable()) h->host_thr.join();
    if (h->hstage) (void)hipHostFree(h->hstage);
    h->hstage = nullptr;
    h->host_factored = false;
    for (auto& g : h->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    drop_solve_graphs(h);
    // every call synchronizes before it returns, so no kernel uses them now
    h->spare.insert(h->spare.end(), h->allocs.begin(), h->allocs.end());
    h->allocs.clear();
    h->d_nodes = nullptr;
    h->d_lidx = nullptr;
    h->d_ext = nullptr;
    h->d_small = nullptr;
    h->d_fused = nullptr;
    h->d_clocks = nullptr;
    h->d_diag = nullptr;
    h->d_fill = nullptr;
    h->d_gemm = nullptr;
    h->d_level_nodes = h->d_child_ptr = h->d_child_list = nullptr;
    h->d_fstep = h->d_bstep = nullptr;
    h->d_xstep = h->d_xbstep = nullptr;
    h->d_xinv = nullptr;
    h->d_segs = nullptr;
    h->d_seg_ptr = nullptr;
    h->d_fixes = nullptr;
    h->d_segpart = nullptr;
    h->d_sinv64 = nullptr;
    h->d_sinv32 = nullptr;
    h->d_fi = h->d_fa = h->d_bg = nullptr;
    h->d_part = nullptr;
    h->d_dst = h->d_src = nullptr;
    h->d_scat_rc = nullptr;
    h->d_diag_src = nullptr;
    h->d_dscale = nullptr;
    h->d_perm = nullptr;
    h->d_pool64 = nullptr;
    h->d_pool32 = nullptr;
    h->d_a = nullptr;
    h->d_info = nullptr;
    h->nrhs_cap = 0;
    h->d_z = h->d_bin = h->d_bp = h->d_xp = nullptr;
    h->d_csr_ptr = h->d_csr_col = nullptr;
    h->d_csr_src = nullptr;
    h->d_csr_val = nullptr;
    h->d_r = h->d_d = nullptr;
    h->d_norms = nullptr;
}

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// ------------------------------------------------------------- schedule
enum SolveKind { kFI = 100, kFA, kFS, kBG, kBS, kFX, kBX, kFC, kBC };

// A solve step with at least this many chunk tasks launches them apart from
// its block tasks (k_fwd_chunk, k_bwd_chunk: more workgroups per CU than the
// block tasks' two).
constexpr int kChunkLaunch = 1024;

// Solve schedule.  Forward: one launch setting every front's Z from b,
// then per level: one launch per child rank adding the children's update
// parts, then launches t = 0 .. steps: the block task of block t (which
// first applies block t - 1 to block t) and the row chunks applying block
// t - 1 to the rows after block t.  Fronts with at least super_min pivot columns step over
// superblocks of 256 columns instead (kind kFX, launched beside step t's
// kFS launch): the superblock task of superblock t and the row chunks
// applying superblock t - 1 to the rows after it.
// Backward, per level: a gather launch, then launches t = steps .. 0: the
// block task of block t (for fronts that have it) and the row chunks forming
// the partial sums block t - 1 needs from the rows after block t (launch
// t = steps only has chunks).  Partials of block j live in bank j % 2, so a
// launch never overwrites the partials its block tasks read.  Superblock
// fronts step 256 columns at a time here too: the top tasks of superblock t
// (kind kBX, after step t's kBS launch) and, in the kBS launch, chunk tasks
// forming the partial sums of superblock t - 1's four 64-column blocks from
// the rows after superblock t; their pivot part of x lives in Z (NodeDev::xz).
void build_solve_plan(mfd_handle* h)
{
    h->fstep.clear();
    h->bstep.clear();
    h->fi.clear();
    h->fa.clear();
    h->bg.clear();
    h->fwd_seq.clear();
    h->bwd_seq.clear();
    h->part_slots = 0;
    h->xstep.clear();
    h->xbstep.clear();
    h->xinv.clear();
    h->sinv_elems = 0;
    auto push = [](std::vector<Launch>& seq, int kind, int64_t first, int64_t end, int step = -1) {
        if (end > first) {
            seq.push_back(Launch{kind, first, (int)(end - first)});
            seq.back().step = step;
        }
    };
    auto super = [&](const NodeDev& d) { return h->super_min > 0 && d.c >= h->super_min; };
    // steps of a level when superblock fronts step 256 columns at a time
    // (forward and backward alike)
    auto fsteps_of = [&](int L) {
        int steps = 0;
        for (int k = h->level_ptr[L]; k < h->level_ptr[L + 1]; ++k) {
            const NodeDev& d = h->nodes[h->level_nodes[k]];
            steps = std::max(steps, (d.c + (super(d) ? kSuper : 64) - 1) / (super(d) ? kSuper : 64));
        }
        return steps;
    };
    for (auto& d : h->nodes) d.xz = super(d) ? 1 : 0;
    // superblock matrices (X, W and V, 3 x 256^2 entries per superblock):
    // four launches of X blocks by block distance, then one of W and V
    // blocks
    const int64_t per_super = kSuperMats;
    std::vector<int64_t> xoff(h->nodes.size(), -1);
    for (size_t s = 0; s < h->nodes.size(); ++s)
        if (super(h->nodes[s])) {
            xoff[s] = h->sinv_elems;
            h->sinv_elems += (int64_t)((h->nodes[s].c + kSuper - 1) / kSuper) * per_super;
        }
    for (int dist = 0; dist <= 4; ++dist) {
        h->xinv_ptr[dist] = (int)h->xinv.size();
        for (size_t s = 0; s < h->nodes.size(); ++s) {
            if (xoff[s] < 0) continue;
            const int c = h->nodes[s].c;
            for (int u = 0; u * kSuper < c; ++u) {
                const int64_t xo = xoff[s] + u * per_super;
                if (dist == 4) {
                    if (u > 0)
                        for (int i = 0; i < 4; ++i)
                            for (int j = 0; j < 4; ++j) h->xinv.push_back(SuperInvTask{(int)s, u, i, 4 + j, xo});
                    if ((u + 1) * kSuper < c)
                        for (int i = 0; i < 4; ++i)
                            for (int j = 0; j < 4; ++j) h->xinv.push_back(SuperInvTask{(int)s, u, i, 8 + j, xo});
                    continue;
                }
                for (int i = dist; i < 4; ++i)
                    if (dist == 0 || u * kSuper + 64 * i < c) h->xinv.push_back(SuperInvTask{(int)s, u, i, i - dist, xo});
            }
        }
    }
    h->xinv_ptr[5] = (int)h->xinv.size();
    // every front's Z (pivot part from b, update part zero) in one launch:
    // the fronts' Z regions are disjoint
    for (int L = 0; L < h->num_levels; ++L)
        for (int k = h->level_ptr[L]; k < h->level_ptr[L + 1]; ++k) {
            const int s = h->level_nodes[k];
            for (int r0 = 0; r0 < h->nodes[s].n; r0 += 256) h->fi.push_back(make_int2(s, r0));
        }
    push(h->fwd_seq, kFI, 0, (int64_t)h->fi.size());
    for (int L = 0; L < h->num_levels; ++L) {
        const int b0 = h->level_ptr[L], b1 = h->level_ptr[L + 1];
        {
            int64_t f;
            // children's update parts, one launch per child rank
            size_t maxch = 0;
            for (int k = b0; k < b1; ++k) {
                const int s = h->level_nodes[k];
                maxch = std::max(maxch, (size_t)(h->child_ptr[s + 1] - h->child_ptr[s]));
            }
            for (size_t r = 0; r < maxch; ++r) {
                f = (int64_t)h->fa.size();
                for (int k = b0; k < b1; ++k) {
                    const int s = h->level_nodes[k];
                    if ((size_t)(h->child_ptr[s + 1] - h->child_ptr[s]) <= r) continue;
                    const int ch = h->child_list[h->child_ptr[s] + r];
                    for (int r0 = 0; r0 < h->nodes[ch].w; r0 += 256) h->fa.push_back(make_int2(ch, r0));
                }
                push(h->fwd_seq, kFA, f, (int64_t)h->fa.size());
            }
        }
        const int steps = fsteps_of(L);
        std::vector<SolveTask> chunks;
        for (int t = 0; t <= steps; ++t) {
            int64_t f0 = (int64_t)h->fstep.size(), x0 = (int64_t)h->xstep.size();
            chunks.clear();
            for (int k = b0; k < b1; ++k) {
                const int s = h->level_nodes[k];
                const NodeDev& d = h->nodes[s];
                if (super(d)) {
                    const int U = (d.c + kSuper - 1) / kSuper;
                    const int64_t xo = xoff[s] + t * per_super;
                    if (t <= U - 1) {
                        const int sb = std::min(kSuper, d.c - kSuper * t);
                        for (int r0 = 0; r0 < sb; r0 += 16)
                            h->xstep.push_back(SuperTask{s, t, r0, sb, t > 0 ? kSuper : 0, 1, 0, 0, xo});
                    }
                    if (t >= 1 && t - 1 <= U - 1) {
                        // superblock t - 1 applied to the rows after superblock t
                        const int pb = std::min(kSuper, d.c - kSuper * (t - 1));
                        for (int r0 = std::min(kSuper * (t + 1), d.c); r0 < d.n; r0 += 16)
                            h->xstep.push_back(SuperTask{s, t, r0, 0, pb, 0, 0, 0, xo});
                    }
                    continue;
                }
                const int T = (d.c + 63) / 64;
                // a front of one block with few update rows: one task solves
                // the block and applies it to those rows
                const bool whole = T == 1 && d.w <= kWholeRows;
                if (t <= T - 1) {
                    const int k0 = 64 * t, kb = std::min(64, d.c - k0);
                    h->fstep.push_back(SolveTask{s, k0, kb, -1, 0, 0, whole ? 1 : 0, t > 0 ? 64 : 0});
                }
                if (!whole && t >= 1 && t - 1 <= T - 1) {
                    // block t - 1 applied to the rows after block t
                    const int k0 = 64 * (t - 1), kb = std::min(64, d.c - k0);
                    for (int r0 = std::min(64 * (t + 1), d.c); r0 < d.n; r0 += 64)
                        chunks.push_back(SolveTask{s, k0, kb, r0, 0, 0, 0, 0});
                }
            }
            // the step's chunk tasks after its block tasks, in the same
            // launch, or in a launch of their own when there are many
            const bool apart = (int)chunks.size() >= kChunkLaunch;
            if (!apart) h->fstep.insert(h->fstep.end(), chunks.begin(), chunks.end());
            push(h->fwd_seq, kFS, f0, (int64_t)h->fstep.size(), t);
            push(h->fwd_seq, kFX, x0, (int64_t)h->xstep.size(), t);
            if (apart) {
                const int64_t c0 = (int64_t)h->fstep.size();
                h->fstep.insert(h->fstep.end(), chunks.begin(), chunks.end());
                push(h->fwd_seq, kFC, c0, (int64_t)h->fstep.size(), t);
            }
        }
    }
    // chunks of the partial sums of block j of front s: rows [min(64 (j + 2), c), n)
    auto chunk_rows = [&](const NodeDev& d, int j) {
        const int start = std::min(64 * (j + 2), d.c);
        return std::make_pair(start, (d.n - start + 255) / 256);
    };
    // superblock fronts: the partial sums of superblock w, rows
    // [min(256 (w + 2), c), n), for each of its (at most four) 64-column
    // blocks
    auto super_chunk_rows = [&](const NodeDev& d, int w) {
        const int start = std::min(kSuper * (w + 2), d.c);
        return std::make_pair(start, (d.n - start + 255) / 256);
    };
    auto super_blocks = [&](const NodeDev& d, int w) { return std::min(4, (d.c - kSuper * w + 63) / 64); };
    auto blocks = [&](const NodeDev& d) { return super(d) ? (d.c + kSuper - 1) / kSuper : (d.c + 63) / 64; };
    // slots written by the chunk tasks of backward step t of front d
    auto step_slots = [&](const NodeDev& d, int t) -> int64_t {
        if (t < 1 || t - 1 >= blocks(d)) return 0;
        if (super(d)) return (int64_t)super_blocks(d, t - 1) * super_chunk_rows(d, t - 1).second;
        return std::max(0, chunk_rows(d, t - 1).second);
    };
    int64_t bank = 1;  // slots per bank: the most chunks any launch has
    for (int L = 0; L < h->num_levels; ++L) {
        const int b0 = h->level_ptr[L], b1 = h->level_ptr[L + 1];
        for (int t = fsteps_of(L); t >= 1; --t) {
            int64_t cnt = 0;
            for (int k = b0; k < b1; ++k) cnt += step_slots(h->nodes[h->level_nodes[k]], t);
            bank = std::max(bank, cnt);
        }
    }
    std::vector<int> qfirst(h->nodes.size(), 0), qcount(h->nodes.size(), 0);
    std::vector<SolveTask> bchunks;  // a backward step's chunk tasks
    for (int L = h->num_levels - 1; L >= 0; --L) {
        const int b0 = h->level_ptr[L], b1 = h->level_ptr[L + 1];
        {
            int64_t f = (int64_t)h->bg.size();
            for (int k = b0; k < b1; ++k) {
                const int s = h->level_nodes[k];
                if (h->nodes[s].parent < 0) continue;
                for (int r0 = 0; r0 < h->nodes[s].w; r0 += 256) h->bg.push_back(make_int2(s, r0));
            }
            push(h->bwd_seq, kBG, f, (int64_t)h->bg.size());
        }
        for (int t = fsteps_of(L); t >= 0; --t) {
            int64_t f0 = (int64_t)h->bstep.size(), x0 = (int64_t)h->xbstep.size();
            bchunks.clear();
            int idx = 0;
            // few chunks (latency-bound steps): four 16-column tasks per
            // chunk; many: one task per chunk
            int64_t nch = 0;
            for (int k = b0; k < b1; ++k) nch += step_slots(h->nodes[h->level_nodes[k]], t);
            const bool split = nch <= kSplitChunks;
            for (int k = b0; k < b1; ++k) {
                const int s = h->level_nodes[k];
                const NodeDev& d = h->nodes[s];
                const int T = blocks(d);
                if (super(d)) {
                    const int64_t xo = xoff[s] + (int64_t)t * per_super;
                    if (t <= T - 1) {
                        // superblock t from the partial sums of the previous
                        // launch and x of superblock t + 1
                        const int sb = std::min(kSuper, d.c - kSuper * t);
                        const int pb = t + 1 <= T - 1 ? std::min(kSuper, d.c - kSuper * (t + 1)) : 0;
                        for (int r0 = 0; r0 < sb; r0 += 16)
                            h->xbstep.push_back(SuperTask{s, t, r0, sb, pb, 2, qfirst[s], qcount[s], xo});
                    }
                    if (t >= 1 && t - 1 <= T - 1) {
                        // partial sums of superblock t - 1 over the rows after superblock t
                        const int w = t - 1;
                        const auto rc = super_chunk_rows(d, w);
                        qfirst[s] = (int)((w % 2) * bank + idx);
                        qcount[s] = rc.second;
                        for (int j = 0; j < super_blocks(d, w); ++j) {
                            const int k0 = kSuper * w + 64 * j, kb = std::min(64, d.c - k0);
                            for (int q = 0; q < rc.second; ++q, ++idx)
                                for (int cq = 0; cq < (split ? 4 : 1); ++cq)
                                    bchunks.push_back(SolveTask{s, k0, kb, rc.first + 256 * q,
                                                                 (int)((w % 2) * bank + idx), 0, split ? cq : -1, 0});
                        }
                    }
                    continue;
                }
                if (t <= T - 1) {
                    const int k0 = 64 * t, kb = std::min(64, d.c - k0);
                    const int kn = t + 1 <= T - 1 ? std::min(64, d.c - k0 - 64) : 0;
                    h->bstep.push_back(SolveTask{s, k0, kb, -1, qfirst[s], qcount[s], 0, kn});
                }
                if (t >= 1 && t - 1 <= T - 1) {
                    const int j = t - 1, k0 = 64 * j, kb = std::min(64, d.c - k0);
                    const auto rc = chunk_rows(d, j);
                    qfirst[s] = (int)((j % 2) * bank + idx);
                    qcount[s] = rc.second;
                    for (int q = 0; q < rc.second; ++q, ++idx)
                        for (int cq = 0; cq < (split ? 4 : 1); ++cq)
                            bchunks.push_back(
                                SolveTask{s, k0, kb, rc.first + 256 * q, (int)((j % 2) * bank + idx), 0, split ? cq : -1, 0});
                }
            }
            const bool apart = (int)bchunks.size() >= kChunkLaunch;
            if (!apart) h->bstep.insert(h->bstep.end(), bchunks.begin(), bchunks.end());
            push(h->bwd_seq, kBS, f0, (int64_t)h->bstep.size(), t);
            push(h->bwd_seq, kBX, x0, (int64_t)h->xbstep.size(), t);
            if (apart) {
                const int64_t c0 = (int64_t)h->bstep.size();
                h->bstep.insert(h->bstep.end(), bchunks.begin(), bchunks.end());
                push(h->bwd_seq, kBC, c0, (int64_t)h->bstep.size(), t);
            }
        }
    }
    h->part_slots = 2 * bank;
}

void build_plan(mfd_handle* h)
{
    const Structure& S = h->S;
    const int m = S.num_node;
    const int period_set = h->upd_period;
    h->nodes.assign(m, NodeDev{});
    int64_t off = 0, zoff = 0, inv_blocks = 0;
    for (int s = 0; s < m; ++s) {
        NodeDev& d = h->nodes[s];
        d.n = S.c[s] + S.w[s];
        d.c = S.c[s];
        d.w = S.w[s];
        d.front = off;
        d.lidx = S.li_ptr[s];
        d.z = zoff;
        d.parent = S.parent[s] - 1;
        d.first = S.first[s] - 1;
        off += align_up((int64_t)d.n * d.n, 32);
        zoff += d.n + d.c;
        d.inv = inv_blocks;  // rebased below
        inv_blocks += (d.c + kPanel - 1) / kPanel;
    }
    // pool layout: fronts | diagonal-block inverses
    for (int s = 0; s < m; ++s) h->nodes[s].inv = off + h->nodes[s].inv * kPanel * kPanel;
    off += inv_blocks * kPanel * kPanel;
    h->pool_elems = off;
    h->z_per_rhs = zoff;
    // subtrees factored on the host (CPU+GPU hybrid): their fronts get no
    // zeroing, assembly or factorization launch; a host subtree root adds
    // into its parent by the extend-add launch, like a small child
    h->hplan = HostPlan{};
    h->host_rc = h->host_roots.empty() ? MFD_OK : build_host_plan(S, h->nodes, h->host_roots, h->hplan);
    if (h->host_rc != MFD_OK) h->hplan = HostPlan{};
    auto on_host = [&](int s) { return !h->hplan.is_host.empty() && h->hplan.is_host[s]; };

    // Levels.  Rule 1 (the default): a front sits one level below its
    // parent (roots on the top level), so siblings share a level and their
    // diagonal chains run in the same launches; rule 0: one level above its
    // highest child (leaves on level 0).  Either way every child is below
    // its parent.
    std::vector<int> lev(m, 0);
    if (h->level_rule == 1) {
        std::vector<int> depth(m, 0);
        for (int s = m - 1; s >= 0; --s)  // parents follow their children
            if (S.parent[s] > 0) depth[s] = depth[S.parent[s] - 1] + 1;
        const int top = m ? *std::max_element(depth.begin(), depth.end()) : 0;
        for (int s = 0; s < m; ++s) lev[s] = top - depth[s];
    } else {
        for (int s = 0; s < m; ++s)
            for (int ch : S.children[s]) lev[s] = std::max(lev[s], lev[ch - 1] + 1);
    }
    h->num_levels = m ? *std::max_element(lev.begin(), lev.end()) + 1 : 0;
    h->level_ptr.assign(h->num_levels + 1, 0);
    for (int s = 0; s < m; ++s) h->level_ptr[lev[s] + 1]++;
    for (int L = 0; L < h->num_levels; ++L) h->level_ptr[L + 1] += h->level_ptr[L];
    h->level_nodes.assign(m, 0);
    {
        std::vector<int> pos(h->level_ptr.begin(), h->level_ptr.end() - 1);
        for (int s = 0; s < m; ++s) h->level_nodes[pos[lev[s]]++] = s;
    }
    // fill groups: a front is first written by its own level's launches or,
    // earlier, by a child's extend-add
    std::vector<int> group(m, 0);
    {
        const int G = std::min(h->num_levels, kFillGroups);
        h->group_level.resize(G);
        for (int g = 0; g < G; ++g) h->group_level[g] = (int)((int64_t)g * h->num_levels / G);
        for (int s = 0; s < m; ++s) {
            int ready = lev[s];
            for (int ch : S.children[s]) ready = std::min(ready, lev[ch - 1]);
            group[s] = (int)(std::upper_bound(h->group_level.begin(), h->group_level.end(), ready) -
                             h->group_level.begin()) - 1;
        }
        h->fill.clear();
        h->fill_gptr.assign(1, 0);
        for (int g = 0; g < G; ++g) {
            for (int s2 = 0; s2 < m; ++s2)
                if (group[s2] == g && !on_host(s2))
                    for (int j0 = 0; j0 < h->nodes[s2].n; j0 += 64) h->fill.push_back(make_int2(s2, j0));
            h->fill_gptr.push_back((int)h->fill.size());
        }
    }
    h->front_group = group;
    // rank of each node among its parent's children (0 for roots)
    std::vector<int> rank(m, 0);
    for (int s = 0; s < m; ++s)
        for (size_t r = 0; r < S.children[s].size(); ++r) rank[S.children[s][r] - 1] = (int)r;
    h->child_ptr.assign(m + 1, 0);
    h->child_list.clear();
    for (int s = 0; s < m; ++s) {
        for (int ch : S.children[s]) h->child_list.push_back(ch - 1);
        h->child_ptr[s + 1] = (int)h->child_list.size();
    }

    h->ext.clear();
    h->small.clear();
    h->fused.clear();
    h->fused_pos.assign(m, -1);
    // front classes: small (n <= 64, whole in LDS), medium (one workgroup
    // per front), large (64-column steps over the chip)
    enum { kClsSmall = 0, kClsMedium = 1, kClsLarge = 2, kClsHost = 3 };
    std::vector<int> cls(m, kClsLarge);
    for (int s = 0; s < m; ++s) {
        const NodeDev& d = h->nodes[s];
        if (on_host(s)) {
            cls[s] = kClsHost;
            continue;
        }
        if (d.n <= kSmallFront) {
            cls[s] = kClsSmall;
            continue;
        }
        const int k = (int)h->front_kind.size() == m ? h->front_kind[s] : -1;
        const bool medium = k >= 0 ? k == 1 : d.c <= h->fuse_cmax;
        cls[s] = medium ? kClsMedium : kClsLarge;
    }
    h->diag.clear();
    h->gemm.clear();
    h->fseq.clear();
    h->segs.clear();
    h->seg_ptr.clear();
    h->fixes.clear();
    h->seg_slots = 0;
    int nside = 0;
    // a wide launch's K chunks (16 columns) divided evenly over at most
    // kSegGroups workgroups (two per CU), no fewer than kSegMin chunks each
    constexpr int kSegGroups = 512, kSegMin = 16;
    auto balance = [&](Launch& l) {
        std::vector<int64_t> units(l.count);
        int64_t total = 0;
        for (int i = 0; i < l.count; ++i) total += units[i] = (h->gemm[l.first + i].k + 15) / 16;
        const int P = (int)std::max<int64_t>(1, std::min<int64_t>(kSegGroups, total / kSegMin));
        l.sfirst = (int64_t)h->seg_ptr.size();
        l.groups = P;
        const size_t s0 = h->segs.size();
        int i = 0;
        int64_t pos = 0;
        for (int w = 0; w < P; ++w) {
            h->seg_ptr.push_back((int)h->segs.size());
            int64_t need = total * (w + 1) / P - total * w / P;
            while (need > 0 && i < l.count) {
                const int64_t take = std::min(need, units[i] - pos);
                const int k = h->gemm[l.first + i].k;
                h->segs.push_back(SegTask{(int)(l.first + i), (int)(16 * pos), (int)std::min<int64_t>(16 * (pos + take), k), 0});
                pos += take;
                need -= take;
                if (pos == units[i]) ++i, pos = 0;
            }
        }
        h->seg_ptr.push_back((int)h->segs.size());
        // tiles in more than one segment keep their partial products in
        // consecutive slots, in K order
        l.xfirst = (int64_t)h->fixes.size();
        int slot = 0;
        for (size_t a = s0; a < h->segs.size();) {
            size_t b = a + 1;
            while (b < h->segs.size() && h->segs[b].task == h->segs[a].task) ++b;
            if (b - a == 1) {
                h->segs[a].slot = -1;
            } else {
                h->fixes.push_back(FixTask{h->segs[a].task, slot, (int)(b - a), 0});
                for (size_t q = a; q < b; ++q) h->segs[q].slot = slot++;
            }
            a = b;
        }
        l.nfix = (int)(h->fixes.size() - l.xfirst);
        h->seg_slots = std::max<int64_t>(h->seg_slots, slot);
    };
    auto push = [&](int kind, int64_t first, int64_t end, int wide = 0) {
        if (end > first) h->fseq.push_back(Launch{kind, first, (int)(end - first), wide});
    };
    h->level_launch.assign(h->num_levels, 0);
    for (int L = 0; L < h->num_levels; ++L) {
        h->level_launch[L] = (int)h->fseq.size();
        const int* lv = h->level_nodes.data() + h->level_ptr[L];
        const int cnt = h->level_ptr[L + 1] - h->level_ptr[L];
        // extend-add of the small children (a large child adds its update
        // into the parent in its own Schur launch): the r-th child of every
        // parent of this level in one launch
        size_t maxch = 0;
        for (int k = 0; k < cnt; ++k) maxch = std::max(maxch, S.children[lv[k]].size());
        for (size_t r = 0; r < maxch; ++r) {
            int64_t first = (int64_t)h->ext.size();
            for (int k = 0; k < cnt; ++k) {
                const auto& chs = S.children[lv[k]];
                if (chs.size() <= r || cls[lv[k]] == kClsHost) continue;
                int ch = chs[r] - 1, nt = (S.w[ch] + kTile - 1) / kTile;
                if (cls[ch] == kClsLarge) continue;
                for (int ti = 0; ti < nt; ++ti)
                    for (int tj = 0; tj <= ti; ++tj) h->ext.push_back(ExtTask{ch, ti, tj, 0});
            }
            push(MFD_K_EXTEND_ADD, first, (int64_t)h->ext.size());
        }
        // small fronts: whole front in LDS
        int64_t first = (int64_t)h->small.size();
        std::vector<int> large;
        for (int k = 0; k < cnt; ++k) {
            if (cls[lv[k]] == kClsSmall) h->small.push_back(lv[k]);
            else if (cls[lv[k]] == kClsLarge) large.push_back(lv[k]);
        }
        push(MFD_K_SMALL_FRONT, first, (int64_t)h->small.size());
        // medium fronts: one workgroup each, the whole partial factorization
        first = (int64_t)h->fused.size();
        for (int k = 0; k < cnt; ++k)
            if (cls[lv[k]] == kClsMedium) {
                h->fused_pos[lv[k]] = (int)h->fused.size();
                h->fused.push_back(lv[k]);
            }
        push(MFD_K_FRONT, first, (int64_t)h->fused.size());
        // large fronts: the first diagonal block, then one launch per
        // 64-column step t (k_upd_diag) holding
        //  * the TRSM of panel t for the rows after block t+1, each tile
        //    first applying panel t-1 to its rows (C = (C - A B^T) L_t^{-T}),
        //  * panel t-1's update of the column blocks after t+1 (in its own
        //    launch when it has many tiles; it is independent of the rest),
        //  * block t+1: panels t-1 and t applied to it (its panel-t rows
        //    solved and stored on the way), Cholesky and inverse;
        // no task of a launch reads what another writes, so the chain of
        // diagonal blocks advances one block per launch.  One Schur update
        // with K = c closes the level.
        int steps = 0;
        for (int s : large) steps = std::max(steps, (h->nodes[s].c + kPanel - 1) / kPanel);
        // period 0 (automatic): 1 for levels whose longest chain is short,
        // where light step tasks pay most, 2 for the long ones
        const int period = period_set > 0 ? period_set : (steps >= kAutoSteps ? 2 : 1);
        {
            int64_t f0 = (int64_t)h->diag.size();
            for (int s : large) h->diag.push_back(DiagTask{s, 0, std::min(kPanel, h->nodes[s].c), 0, 0, 0, 0, 0});
            push(MFD_K_POTRF, f0, (int64_t)h->diag.size());
        }
        int prev_side = -1;
        for (int t = 0; t < steps; ++t) {
            const int k0 = t * kPanel;
            // panels not yet applied to block column t (and to block t+1):
            // the in-panel update runs every period steps, on the period
            // panels before it
            const int pend = t == 0 ? 0 : ((t % period) == 0 ? period : t % period) * kPanel;
            const bool upd = t >= period && t % period == 0;
            int64_t f1 = (int64_t)h->gemm.size();
            // the TRSM tile of the rows of block t+2 also applies the panels
            // pending for block t+2 (and panel t) to its diagonal block, which
            // the diagonal task of step t+1 then skips
            auto syrk_at = [&](const NodeDev& d) { return k0 + 2 * kPanel < d.c; };
            for (int s : large) {
                const NodeDev& d = h->nodes[s];
                if (d.c <= k0) continue;
                const int kb = std::min(kPanel, d.c - k0), kn = std::min(kPanel, d.c - k0 - kb);
                for (int r0 = k0 + kb + kn; r0 < d.n; r0 += kTile) {
                    GemmTask g{};
                    if (r0 == k0 + 2 * kPanel && syrk_at(d)) {
                        g.syrk = std::min(kPanel, d.c - r0);
                        g.cd = d.front + r0 + (int64_t)r0 * d.n;
                    }
                    g.c = d.front + r0 + (int64_t)k0 * d.n;
                    g.ldc = g.lda = d.n;
                    g.m = std::min(kTile, d.n - r0);
                    g.n = kb;
                    if (t == 0) {
                        g.a = g.c;
                        g.b = d.inv;  // L^{-T}, read transposed
                        g.ldb = kPanel;
                        g.k = kb;
                        g.mode = 1;
                    } else {
                        g.a = d.front + r0 + (int64_t)(k0 - pend) * d.n;
                        g.b = d.front + k0 + (int64_t)(k0 - pend) * d.n;
                        g.ldb = d.n;
                        g.k = pend;
                        g.d = d.inv + (int64_t)t * kPanel * kPanel;
                        g.kd = kb;
                        g.mode = 2;
                    }
                    h->gemm.push_back(g);
                }
            }
            int64_t f3 = (int64_t)h->gemm.size();
            if (upd) {
                const int kq = k0 - period * kPanel;  // panels t-period .. t-1
                for (int s : large) {
                    const NodeDev& d = h->nodes[s];
                    for (int c0 = k0 + kPanel; c0 < d.c; c0 += kTile) {
                        const int nc = std::min(kTile, d.c - c0);
                        for (int r0 = c0; r0 < d.n; r0 += kTile) {
                            // block t+1 itself belongs to its diagonal task,
                            // block t+2 to the TRSM tile of its rows
                            const bool own = c0 == k0 + kPanel || (c0 == k0 + 2 * kPanel && syrk_at(d));
                            const int rs = (own && r0 == c0) ? c0 + nc : r0;
                            const int m = std::min(r0 + kTile, d.n) - rs;
                            if (m <= 0) continue;
                            GemmTask g{};
                            g.c = d.front + rs + (int64_t)c0 * d.n;
                            g.a = d.front + rs + (int64_t)kq * d.n;
                            g.b = d.front + c0 + (int64_t)kq * d.n;
                            g.ldc = g.lda = g.ldb = d.n;
                            g.m = m;
                            g.n = nc;
                            g.k = period * kPanel;
                            g.tri = rs == c0;
                            h->gemm.push_back(g);
                        }
                    }
                }
            }
            const int64_t f4u = (int64_t)h->gemm.size();
            const bool fuse = f4u - f3 <= kFuseTiles;
            if (!fuse) {  // panel t-1's update, beside this step's launch
                push(MFD_K_UPDATE, f3, f4u);
                h->fseq.back().side = 1;
                h->fseq.back().side_id = nside;
            }
            int64_t f2 = (int64_t)h->diag.size();
            for (int s : large) {
                const NodeDev& d = h->nodes[s];
                if (d.c <= k0 + kPanel) continue;
                // from step 1 on, the previous step's TRSM tile of its rows
                // has applied the pending panels to the block
                h->diag.push_back(
                    DiagTask{s, k0 + kPanel, std::min(kPanel, d.c - k0 - kPanel), kPanel, pend, t >= 1 ? 1 : 0, 0, 0});
            }
            const int64_t ng = (fuse ? f4u : f3) - f1, nd = (int64_t)h->diag.size() - f2;
            // many diagonal tasks (many fronts at this level): the GEMM tiles
            // go to a launch of their own, where they are not held to the
            // diagonal tasks' one workgroup per CU
            const bool apart = nd >= kApartDiag && ng > 0;
            if (ng + nd > 0) {
                Launch l{MFD_K_POTRF, f1, apart ? 0 : (int)ng};
                l.fused = 1;
                l.dfirst = f2;
                l.dcount = (int)nd;
                l.wait_id = prev_side;  // panel t-2's update of block column t
                h->fseq.push_back(l);
                if (apart) {
                    Launch g{MFD_K_TRSM, f1, (int)ng};
                    g.wait_id = prev_side;
                    h->fseq.push_back(g);
                }
            }
            prev_side = fuse ? -1 : nside++;
        }
        // Schur update with the extend-add of its result into the parent,
        // one launch per child rank (the r-th children of their parents), so
        // no two tiles of a launch add into one parent entry
        int maxrank = 0;
        for (int s : large) maxrank = std::max(maxrank, rank[s]);
        for (int r = 0; r <= maxrank; ++r)
            for (int wide = 1; wide >= 0; --wide) {  // the 128 x 128 launch, then the 64 x 64 one
                const int tile = wide ? kWideTile : kTile;
                int64_t f4 = (int64_t)h->gemm.size();
                for (int s : large) {
                    if (rank[s] != r) continue;
                    const NodeDev& d = h->nodes[s];
                    if ((h->wide_min_rows > 0 && d.w >= h->wide_min_rows) != (wide == 1)) continue;
                    const NodeDev* pa = d.parent >= 0 ? &h->nodes[d.parent] : nullptr;
                    for (int c0 = d.c; c0 < d.n; c0 += tile)
                        for (int r0 = c0; r0 < d.n; r0 += tile) {
                            GemmTask g{};
                            g.c = d.front + r0 + (int64_t)c0 * d.n;
                            g.a = d.front + r0;
                            g.b = d.front + c0;
                            g.ldc = g.lda = g.ldb = d.n;
                            g.m = std::min(tile, d.n - r0);
                            g.n = std::min(tile, d.n - c0);
                            g.k = d.c;
                            g.tri = r0 == c0;
                            if (pa) {
                                g.mode = 3;
                                g.p = pa->front;
                                g.np = pa->n;
                                g.li = d.lidx + (r0 - d.c);
                                g.lj = d.lidx + (c0 - d.c);
                            }
                            h->gemm.push_back(g);
                        }
                }
                push(MFD_K_SCHUR, f4, (int64_t)h->gemm.size(), wide);
                if (wide && (int64_t)h->gemm.size() > f4) balance(h->fseq.back());
            }

    }
    // the host factors go into the pool before the first launch of the
    // lowest level holding a parent of a host subtree root (at the end when
    // no root has a parent)
    h->host_join = -1;
    if (!h->hplan.roots.empty()) {
        int jl = h->num_levels;
        for (int r : h->hplan.roots)
            if (S.parent[r] > 0) jl = std::min(jl, lev[S.parent[r] - 1]);
        h->host_join = jl < h->num_levels ? h->level_launch[jl] : (int)h->fseq.size();
    }
    build_solve_plan(h);
}

// Full symmetric CSR (both triangles) of the postordered matrix for the
// refinement residual; entry e takes its value from caller entry csr_src[e].
void build_csr(const Structure& S, std::vector<int>& ptr, std::vector<int>& col, std::vector<int64_t>& src)
{
    const int n = S.n;
    std::vector<int> cnt(n + 1, 0);
    // rows of a_structure: snidx is front-relative; recover the global row
    // from the node's first column and row list.
    std::vector<int> node_of(n, 0);
    for (int s = 0; s < S.num_node; ++s)
        for (int j = S.first[s]; j < S.first[s + 1]; ++j) node_of[j - 1] = s;
    auto global_row = [&](int j, int64_t p) {
        int s = node_of[j - 1];
        int r = S.snidx[p];
        int c = S.c[s];
        return r <= c ? S.first[s] + r - 1 : S.l_rows[S.l_ptr[s] + (r - c - 1)];
    };
    for (int j = 1; j <= n; ++j)
        for (int64_t p = S.a_colptr[j - 1]; p < S.a_colptr[j]; ++p) {
            int r = global_row(j, p);
            cnt[r]++;
            if (r != j) cnt[j]++;
        }
    ptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + cnt[i + 1];
    col.assign(ptr[n], 0);
    src.assign(ptr[n], 0);
    std::vector<int> pos(ptr.begin(), ptr.end() - 1);
    for (int j = 1; j <= n; ++j)
        for (int64_t p = S.a_colptr[j - 1]; p < S.a_colptr[j]; ++p) {
            int r = global_row(j, p);
            int64_t v = S.ccs_perm[p] - 1;
            col[pos[r - 1]] = j - 1;
            src[pos[r - 1]++] = v;
            if (r != j) {
                col[pos[j - 1]] = r - 1;
                src[pos[j - 1]++] = v;
            }
        }
}

int setup_device(mfd_handle* h)
{
    free_device(h);
    const Structure& S = h->S;
    build_plan(h);
    MFD_TRY(h->host_rc);
    if (h->hplan.stage_elems > 0)
        HIP_TRY(hipHostMalloc((void**)&h->hstage, h->hplan.stage_elems * sizeof(double), hipHostMallocDefault));
    MFD_TRY(upload(h, &h->d_nodes, h->nodes));
    MFD_TRY(upload(h, &h->d_lidx, S.lidx));
    MFD_TRY(upload(h, &h->d_ext, h->ext));
    MFD_TRY(upload(h, &h->d_small, h->small));
    MFD_TRY(upload(h, &h->d_fused, h->fused));
    MFD_TRY(dalloc(h, &h->d_clocks, 2 * (int64_t)h->fused.size()));
    MFD_TRY(upload(h, &h->d_diag, h->diag));
    MFD_TRY(upload(h, &h->d_fill, h->fill));
    while (h->fill_ev.size() < h->group_level.size()) {
        hipEvent_t ev;
        HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        h->fill_ev.push_back(ev);
    }
    MFD_TRY(upload(h, &h->d_gemm, h->gemm));
    MFD_TRY(upload(h, &h->d_segs, h->segs));
    MFD_TRY(upload(h, &h->d_seg_ptr, h->seg_ptr));
    MFD_TRY(upload(h, &h->d_fixes, h->fixes));
    if (h->seg_slots > 0) MFD_TRY(dalloc(h, &h->d_segpart, h->seg_slots * 16384));
    MFD_TRY(upload(h, &h->d_level_nodes, h->level_nodes));
    MFD_TRY(upload(h, &h->d_child_ptr, h->child_ptr));
    MFD_TRY(upload(h, &h->d_child_list, h->child_list));
    MFD_TRY(upload(h, &h->d_fstep, h->fstep));
    MFD_TRY(upload(h, &h->d_bstep, h->bstep));
    MFD_TRY(upload(h, &h->d_xstep, h->xstep));
    MFD_TRY(upload(h, &h->d_xbstep, h->xbstep));
    MFD_TRY(upload(h, &h->d_xinv, h->xinv));
    if (h->sinv_elems > 0) MFD_TRY(dalloc(h, &h->d_sinv64, h->sinv_elems));
    MFD_TRY(upload(h, &h->d_fi, h->fi));
    MFD_TRY(upload(h, &h->d_fa, h->fa));
    MFD_TRY(upload(h, &h->d_bg, h->bg));
    // assembly map: A entry -> front position (coefficient_setter_m.f90:26-41),
    // the entries in the fill groups' order
    const int64_t nnz = (int64_t)S.snidx.size();
    std::vector<int64_t> dst(nnz), src(nnz), diag_src(S.n, 0);
    std::vector<int2> rc(nnz);
    h->scat_gptr.assign(1, 0);
    int64_t e = 0;
    for (size_t g = 0; g < h->group_level.size(); ++g) {
        for (int s = 0; s < S.num_node; ++s) {
            if (h->front_group[s] != (int)g) continue;
            const NodeDev& d = h->nodes[s];
            const bool host = !h->hplan.is_host.empty() && h->hplan.is_host[s];
            for (int j = S.first[s]; j < S.first[s + 1]; ++j) {
                int64_t q = j - S.first[s];
                for (int64_t p = S.a_colptr[j - 1]; p < S.a_colptr[j]; ++p) {
                    // global row: a pivot row of this front or one of its update rows
                    const int r = S.snidx[p];
                    const int grow = r <= S.c[s] ? S.first[s] + r - 1 : S.l_rows[S.l_ptr[s] + (r - S.c[s] - 1)];
                    if (grow == j) diag_src[j - 1] = S.ccs_perm[p] - 1;
                    if (host) continue;  // assembled on the host
                    dst[e] = d.front + (S.snidx[p] - 1) + q * d.n;
                    src[e] = S.ccs_perm[p] - 1;
                    rc[e] = make_int2(grow - 1, j - 1);
                    ++e;
                }
            }
        }
        h->scat_gptr.push_back(e);
    }
    if (e != nnz - h->hplan.entries) return MFD_ERR_INTERNAL;
    dst.resize(e);
    src.resize(e);
    rc.resize(e);
    MFD_TRY(upload(h, &h->d_dst, dst));
    MFD_TRY(upload(h, &h->d_src, src));
    MFD_TRY(upload(h, &h->d_scat_rc, rc));
    MFD_TRY(upload(h, &h->d_diag_src, diag_src));
    MFD_TRY(dalloc(h, &h->d_dscale, S.n));
    std::vector<int> perm0(S.n);
    for (int i = 0; i < S.n; ++i) perm0[i] = S.perm[i] - 1;
    MFD_TRY(upload(h, &h->d_perm, perm0));
    MFD_TRY(dalloc(h, &h->d_a, nnz));
    MFD_TRY(dalloc(h, &h->d_info, 1));
    MFD_TRY(dalloc(h, &h->d_pool64, h->pool_elems));
    h->front_bytes = h->pool_elems * (int64_t)sizeof(double);
    std::vector<int> cptr, ccol;
    std::vector<int64_t> csrc;
    build_csr(S, cptr, ccol, csrc);
    h->csr_nnz = (int64_t)ccol.size();
    MFD_TRY(upload(h, &h->d_csr_ptr, cptr));
    MFD_TRY(upload(h, &h->d_csr_col, ccol));
    MFD_TRY(upload(h, &h->d_csr_src, csrc));
    MFD_TRY(dalloc(h, &h->d_csr_val, h->csr_nnz));
    return MFD_OK;
}

int ensure_rhs(mfd_handle* h, int nrhs)
{
    if (nrhs <= h->nrhs_cap) return MFD_OK;
    const int64_t n = h->S.n;
    MFD_TRY(dalloc(h, &h->d_z, h->z_per_rhs * nrhs));
    MFD_TRY(dalloc(h, &h->d_bin, n * nrhs));
    MFD_TRY(dalloc(h, &h->d_bp, n * nrhs));
    MFD_TRY(dalloc(h, &h->d_xp, n * nrhs));
    MFD_TRY(dalloc(h, &h->d_r, n * nrhs));
    MFD_TRY(dalloc(h, &h->d_d, n * nrhs));
    MFD_TRY(dalloc(h, &h->d_norms, 4 * nrhs + 1 + 2 * kSumParts * nrhs));
    MFD_TRY(dalloc(h, &h->d_part, h->part_slots * 64 * nrhs));
    drop_solve_graphs(h);  // captured solves refer to the old buffers
    h->nrhs_cap = nrhs;
    return MFD_OK;
}

// ------------------------------------------------------------ factorize
template <typename T> T* pool_of(mfd_handle* h);
template <> double* pool_of<double>(mfd_handle* h) { return h->d_pool64; }
template <> float* pool_of<float>(mfd_handle* h) { return h->d_pool32; }
template <typename T> T* sinv_of(mfd_handle* h);
template <> double* sinv_of<double>(mfd_handle* h) { return h->d_sinv64; }
template <> float* sinv_of<float>(mfd_handle* h) { return h->d_sinv32; }

template <typename T> void enqueue_launch(mfd_handle* h, const Launch& l, hipStream_t st)
{
    T* pool = pool_of<T>(h);
    switch (l.kind) {
    case MFD_K_EXTEND_ADD: launch_extend_add<T>(pool, h->d_lidx, h->d_nodes, h->d_ext + l.first, l.count, st); break;
    case MFD_K_SMALL_FRONT: launch_small_front<T>(pool, h->d_nodes, h->d_small + l.first, l.count, h->d_info, st); break;
    case MFD_K_FRONT:
        launch_front_wg<T>(pool, h->d_nodes, h->d_fused + l.first, l.count, h->d_info, h->d_clocks + 2 * l.first, st,
                           sizeof(T) == 4 ? h->lowp : 0);
        break;
    case MFD_K_POTRF:
        if (l.fused)
            launch_upd_diag<T>(pool, h->d_nodes, h->d_diag + l.dfirst, l.dcount, h->d_gemm + l.first, l.count, h->d_info,
                               st);
        else
            launch_diag<T>(pool, h->d_nodes, h->d_diag + l.first, l.count, h->d_info, st);
        break;
    case MFD_K_TRSM:
    case MFD_K_UPDATE:
    case MFD_K_SCHUR:
        if (l.wide)
            launch_gemm_seg<T>(pool, h->d_gemm, h->d_segs, h->d_seg_ptr + l.sfirst, l.groups, h->d_fixes + l.xfirst,
                               l.nfix, reinterpret_cast<T*>(h->d_segpart), h->d_lidx, st, sizeof(T) == 4 ? h->lowp : 0);
        else
            launch_gemm<T>(pool, h->d_gemm + l.first, l.count,
                           l.kind == MFD_K_TRSM ? 0 : l.kind == MFD_K_UPDATE ? 1 : 2, h->d_lidx, st,
                           sizeof(T) == 4 ? h->lowp : 0);
        break;
    default: break;
    }
}

double* stage_of(mfd_handle* h) { return h->hstage ? h->hstage : h->hstage_vec.data(); }

// CPU+GPU hybrid: waits for the host part and copies its fronts and their
// diagonal-block inverses into the pool (one copy per run of consecutive
// host fronts and one per run of their inverses).
hipError_t host_join(mfd_handle* h, hipStream_t st)
{
    if (h->host_thr.joinable()) h->host_thr.join();
    hipError_t err = hipSuccess;
    for (const HostRun& r : h->hplan.runs) {
        const hipError_t e = hipMemcpyAsync(h->d_pool64 + r.pool, stage_of(h) + r.stage, r.count * sizeof(double),
                                            hipMemcpyHostToDevice, st);
        if (e != hipSuccess && err == hipSuccess) err = e;
    }
    return err;
}

// The factorization schedule: reset the pivot flag, zero the front pool,
// scatter A, then every level's launches.  No hipMemsetAsync, so the
// sequence can be captured into a graph.
// Returns the first error of the event and stream-wait calls that order the
// streams (a failed one would leave two launches unordered).
template <typename T> hipError_t enqueue_factor(mfd_handle* h, bool profile)
{
    hipError_t err = hipSuccess;
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess && err == hipSuccess) err = e;
    };
    hipStream_t st = h->stream;
    T* pool = pool_of<T>(h);
    auto mark = [&](int k) {
        if (!profile) return;
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, st);
        h->prof_ev.push_back(e);
        (void)k;
    };
    launch_set_int(h->d_info, INT_MAX, st);
    mark(0);
    // Zeroing and assembly of the fronts: beside the factorization, group by
    // group, when the fill stream and its events exist; otherwise (and when
    // profiling) all of it first, on the main stream.
    const int G = (int)h->group_level.size();
    const bool split_fill = !profile && h->fillst && (int)h->fill_ev.size() >= G;
    hipStream_t fst = split_fill ? h->fillst : st;
    if (split_fill) {
        ok(hipEventRecord(h->fork, st));
        ok(hipStreamWaitEvent(fst, h->fork, 0));
    }
    const bool scaled = sizeof(T) == 4 && h->lowp != 0;
    if (scaled) launch_diag_scale(h->d_dscale, h->d_a, h->d_diag_src, h->S.n, fst);
    for (int g = 0; g < G; ++g) {
        launch_fill_lower<T>(pool, h->d_nodes, h->d_fill + h->fill_gptr[g], h->fill_gptr[g + 1] - h->fill_gptr[g], fst);
        const int64_t e0 = h->scat_gptr[g], ne = h->scat_gptr[g + 1] - h->scat_gptr[g];
        if (scaled)
            launch_scatter_a_scaled<T>(pool, h->d_a, h->d_dst + e0, h->d_src + e0, h->d_scat_rc + e0, h->d_dscale, ne,
                                       fst);
        else
            launch_scatter_a<T>(pool, h->d_a, h->d_dst + e0, h->d_src + e0, ne, fst);
        if (split_fill) ok(hipEventRecord(h->fill_ev[g], fst));
    }
    mark(0);
    // Side launches fork from the main stream's current position and run
    // beside it; a step's launches join the side launch they depend on, and
    // any other launch joins every side launch before it.
    const bool two = !profile && h->side;
    // the host subtrees' factors join before launch host_join (FP64 only)
    bool hybrid = sizeof(T) == 8 && !h->hplan.roots.empty();
    int last_side = -1, joined = -1, gnext = split_fill ? 0 : G;
    for (size_t i = 0; i < h->fseq.size(); ++i) {
        const Launch& l = h->fseq[i];
        if (hybrid && (int)i == h->host_join) {
            ok(host_join(h, st));
            hybrid = false;
        }
        for (; gnext < G && h->level_launch[h->group_level[gnext]] <= (int)i; ++gnext)
            ok(hipStreamWaitEvent(st, h->fill_ev[gnext], 0));
        if (two && l.side) {
            ok(hipEventRecord(h->fork, st));
            ok(hipStreamWaitEvent(h->side, h->fork, 0));
            enqueue_launch<T>(h, l, h->side);
            ok(hipEventRecord(h->sev[l.side_id & 1], h->side));
            last_side = l.side_id;
            continue;
        }
        if (two && l.kind != MFD_K_POTRF && l.kind != MFD_K_TRSM && last_side > joined) {
            ok(hipStreamWaitEvent(st, h->sev[last_side & 1], 0));
            joined = last_side;
        } else if (two && l.wait_id > joined) {
            ok(hipStreamWaitEvent(st, h->sev[l.wait_id & 1], 0));
            joined = l.wait_id;
        }
        enqueue_launch<T>(h, l, st);
        mark(l.kind);
    }
    if (two && last_side > joined) ok(hipStreamWaitEvent(st, h->sev[last_side & 1], 0));
    for (; gnext < G; ++gnext) ok(hipStreamWaitEvent(st, h->fill_ev[gnext], 0));
    if (hybrid) ok(host_join(h, st));
    // superblock inverses of the solves, after every front is final
    for (int dist = 0; dist <= 4; ++dist)
        launch_superinv<T>(pool, sinv_of<T>(h), h->d_nodes, h->d_xinv + h->xinv_ptr[dist],
                           h->xinv_ptr[dist + 1] - h->xinv_ptr[dist], dist == 4, st);
    return err;
}

void collect_profile(mfd_handle* h, const std::vector<int>& kinds)
{
    // prof_ev[0..1] bracket assembly; then one event after each launch
    if (h->prof_ev.size() < 2) return;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, h->prof_ev[0], h->prof_ev[1]);
    h->ktime[MFD_K_ASSEMBLE] += ms;
    h->kcount[MFD_K_ASSEMBLE] += 2;
    for (size_t i = 0; i + 2 < h->prof_ev.size() && i < kinds.size(); ++i) {
        (void)hipEventElapsedTime(&ms, h->prof_ev[i + 1], h->prof_ev[i + 2]);
        h->ktime[kinds[i]] += ms;
        h->kcount[kinds[i]] += 1;
    }
    for (auto e : h->prof_ev) (void)hipEventDestroy(e);
    h->prof_ev.clear();
}

template <typename T> int run_factor(mfd_handle* h)
{
    const bool profile = (h->flags & MFD_FLAG_PROFILE) != 0;
    // the hybrid's host work sits between two parts of the enqueue, so it
    // is never captured
    const bool hybrid = sizeof(T) == 8 && !h->hplan.roots.empty();
    const bool graph = (h->flags & MFD_FLAG_GRAPH) != 0 && !profile && !h->no_alloc && !hybrid;
    const int gi = sizeof(T) == 8 ? 0 : 1 + h->lowp;
    if (hybrid) {
        if (!h->host_vals) return MFD_ERR_STATE;
        if (h->host_thr.joinable()) h->host_thr.join();
        h->host_factored = false;
        h->host_info = 0;
        h->host_thr = std::thread([h]() {
            const double t0 = now_ms();
            h->host_info = host_factor(h->hplan, h->S, h->host_vals, stage_of(h), h->host_threads);
            h->host_ms = now_ms() - t0;
        });
    }
    HIP_TRY(hipEventRecord(h->ev[0], h->stream));
    if (graph) {
        if (!h->graph[gi]) {
            hipGraph_t g;
            HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            const hipError_t qe = enqueue_factor<T>(h, false);
            HIP_TRY(hipStreamEndCapture(h->stream, &g));
            if (qe != hipSuccess) {
                (void)hipGraphDestroy(g);
                HIP_TRY(qe);
            }
            HIP_TRY(hipGraphInstantiate(&h->graph[gi], g, nullptr, nullptr, 0));
            (void)hipGraphDestroy(g);
        }
        HIP_TRY(hipGraphLaunch(h->graph[gi], h->stream));
    } else {
        if (profile)
            for (int k = 0; k < MFD_K_NUM; ++k) h->ktime[k] = 0, h->kcount[k] = 0;
        HIP_TRY(enqueue_factor<T>(h, profile));
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(h->ev[1], h->stream));
    HIP_TRY(hipEventSynchronize(h->ev[1]));
    if (profile) {
        std::vector<int> kinds;
        for (const Launch& l : h->fseq) kinds.push_back(l.kind);
        collect_profile(h, kinds);
    }
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
    h->st.factor_ms = ms;
    int info = INT_MAX;
    HIP_TRY(hipMemcpy(&info, h->d_info, sizeof(int), hipMemcpyDeviceToHost));
    if (hybrid) {
        if (h->host_thr.joinable()) h->host_thr.join();
        h->host_factored = true;
        if (h->host_info > 0) info = std::min(info, h->host_info);
    }
    return info == INT_MAX ? MFD_OK : info;
}

int factorize_common(mfd_handle* h, int precision)
{
    if (!h->analyzed) return MFD_ERR_STATE;
    if (!h->device_ok) return MFD_ERR_NODEVICE;
    if (precision < MFD_FP64 || precision > MFD_BF16_IR) return MFD_ERR_UNSUPPORTED;
    if (!h->hplan.roots.empty() && precision != MFD_FP64) return MFD_ERR_UNSUPPORTED;
    h->factored = false;
    h->anorm = -1;  // the values may change
    h->solved_nrhs = 0;
    int rc = MFD_OK;
    if (precision == MFD_FP16_IR || precision == MFD_BF16_IR) {
        if (!h->d_pool32) MFD_TRY(dalloc(h, &h->d_pool32, h->pool_elems));
        if (!h->d_sinv32 && h->sinv_elems > 0) MFD_TRY(dalloc(h, &h->d_sinv32, h->sinv_elems));
        h->lowp = precision == MFD_FP16_IR ? 1 : 2;
        rc = run_factor<float>(h);
        h->lowp = 0;
        // a pivot not positive with 16-bit updates: FP32 fronts next
        if (rc > 0) precision = MFD_FP32_IR;
    }
    if (precision == MFD_FP32_IR) {
        if (!h->d_pool32) MFD_TRY(dalloc(h, &h->d_pool32, h->pool_elems));
        if (!h->d_sinv32 && h->sinv_elems > 0) MFD_TRY(dalloc(h, &h->d_sinv32, h->sinv_elems));
        rc = run_factor<float>(h);
        // a pivot not positive in FP32: the FP64 refactorization the
        // refinement falls back to when it stalls; only an FP64 failure
        // says that the matrix is not positive definite
        if (rc > 0) precision = MFD_FP64;
    }
    if (precision == MFD_FP64) rc = run_factor<double>(h);
    if (rc < 0) return rc;
    h->precision = precision;
    h->factored = rc == MFD_OK;
    const double cost = (double)mfd::cost(h->S);
    h->st.factor_gflops = h->st.factor_ms > 0 ? cost / (h->st.factor_ms * 1e6) : 0;
    h->st.precision = precision;
    return rc;
}

// --------------------------------------------------------------- solve
template <typename T> void enqueue_solve(mfd_handle* h, const double* bp, double* xp, int nrhs)
{
    hipStream_t st = h->stream;
    const T* pool = pool_of<T>(h);
    const int n = h->S.n;
    const bool profile = (h->flags & MFD_FLAG_PROFILE) != 0;
    hipEvent_t pe[3] = {};
    if (profile)
        for (auto& e : pe) (void)hipEventCreate(&e);
    if (profile) (void)hipEventRecord(pe[0], st);
    // a step's kFS and kFX launches (kBS and kBX) touch different fronts (the
    // kBX tasks read the partial-sum bank the kBS tasks do not write): one
    // launch of both
    auto paired = [](const std::vector<Launch>& seq, size_t i, int a, int b) {
        return seq[i].kind == a && i + 1 < seq.size() && seq[i + 1].kind == b && seq[i].step == seq[i + 1].step;
    };
    int nfwd = 0, nbwd = 0;  // launches issued
    for (size_t i = 0; i < h->fwd_seq.size(); ++i, ++nfwd) {
        const Launch& l = h->fwd_seq[i];
        if (paired(h->fwd_seq, i, kFS, kFX)) {
            const Launch& x = h->fwd_seq[++i];
            launch_fwd_pair<T>(pool, sinv_of<T>(h), h->d_z, nrhs, h->d_nodes, h->d_xstep + x.first, x.count,
                               h->d_fstep + l.first, l.count, st);
            continue;
        }
        switch (l.kind) {
        case kFI: launch_fwd_init(h->d_z, bp, n, nrhs, h->d_nodes, h->d_fi + l.first, l.count, st); break;
        case kFA: launch_fwd_add(h->d_z, nrhs, h->d_nodes, h->d_lidx, h->d_fa + l.first, l.count, st); break;
        case kFS: launch_fwd_step<T>(pool, h->d_z, nrhs, h->d_nodes, h->d_fstep + l.first, l.count, st); break;
        case kFC: launch_fwd_chunk<T>(pool, h->d_z, nrhs, h->d_nodes, h->d_fstep + l.first, l.count, st); break;
        case kFX:
            launch_fwd_super<T>(pool, sinv_of<T>(h), h->d_z, nrhs, h->d_nodes, h->d_xstep + l.first, l.count, st);
            break;
        default: break;
        }
    }
    if (profile) (void)hipEventRecord(pe[1], st);
    for (size_t i = 0; i < h->bwd_seq.size(); ++i, ++nbwd) {
        const Launch& l = h->bwd_seq[i];
        if (paired(h->bwd_seq, i, kBS, kBX)) {
            const Launch& x = h->bwd_seq[++i];
            launch_bwd_pair<T>(pool, sinv_of<T>(h), h->d_z, h->d_part, xp, n, nrhs, h->d_nodes, h->d_xbstep + x.first,
                               x.count, h->d_bstep + l.first, l.count, st);
            continue;
        }
        switch (l.kind) {
        case kBG: launch_bwd_gather(h->d_z, nrhs, h->d_nodes, h->d_lidx, h->d_bg + l.first, l.count, st); break;
        case kBS:
            launch_bwd_step<T>(pool, h->d_z, h->d_part, xp, n, nrhs, h->d_nodes, h->d_bstep + l.first, l.count, st);
            break;
        case kBC:
            launch_bwd_chunk<T>(pool, h->d_z, h->d_part, nrhs, h->d_nodes, h->d_bstep + l.first, l.count, st);
            break;
        case kBX:
            launch_bwd_super<T>(pool, sinv_of<T>(h), h->d_z, h->d_part, xp, n, nrhs, h->d_nodes, h->d_xbstep + l.first,
                                l.count, st);
            break;
        default: break;
        }
    }
    if (profile) {
        (void)hipEventRecord(pe[2], st);
        (void)hipEventSynchronize(pe[2]);
        float a = 0, b = 0;
        (void)hipEventElapsedTime(&a, pe[0], pe[1]);
        (void)hipEventElapsedTime(&b, pe[1], pe[2]);
        h->ktime[MFD_K_SOLVE_FWD] += a;
        h->kcount[MFD_K_SOLVE_FWD] += nfwd;
        h->ktime[MFD_K_SOLVE_BWD] += b;
        h->kcount[MFD_K_SOLVE_BWD] += nbwd;
        for (auto e : pe) (void)hipEventDestroy(e);
    }
}

// One forward + backward sweep, replayed from a captured graph when graphs
// are enabled (one graph per precision, nrhs and buffer pair).
template <typename T> int run_solve(mfd_handle* h, const double* bp, double* xp, int nrhs)
{
    const bool graph = (h->flags & MFD_FLAG_GRAPH) != 0 && (h->flags & MFD_FLAG_PROFILE) == 0 && !h->no_alloc;
    if (!graph) {
        enqueue_solve<T>(h, bp, xp, nrhs);
        return MFD_OK;
    }
    const int prec = sizeof(T) == 8 ? 0 : 1;
    for (const auto& g : h->sgraphs)
        if (g.prec == prec && g.nrhs == nrhs && g.bp == bp && g.xp == xp) {
            HIP_TRY(hipGraphLaunch(g.exec, h->stream));
            return MFD_OK;
        }
    hipGraph_t g;
    hipGraphExec_t exec;
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueue_solve<T>(h, bp, xp, nrhs);
    HIP_TRY(hipStreamEndCapture(h->stream, &g));
    HIP_TRY(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    h->sgraphs.push_back({prec, nrhs, bp, xp, exec});
    HIP_TRY(hipGraphLaunch(exec, h->stream));
    return MFD_OK;
}

// Normwise backward error per right-hand side:
// eta = ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf), and res2 =
// ||b - A x||_2 / ||x||_2 (the measure of example.f90:147-180; both norms are
// unchanged by the analysis permutation), each per right-hand side into
// h->err_eta / h->err_res2, and their maxima over the right-hand sides into
// *eta / *res2.  With d != nullptr, also ||d||_inf / ||x||_inf per
// right-hand side into h->err_ferr.  r receives b - A x.  The sums of
// squares come back as per-block partials added here in a fixed order.
int backward_error(mfd_handle* h, const double* bp, const double* xp, double* r, int nrhs, double* eta, double* res2,
                   const double* d = nullptr)
{
    const int n = h->S.n;
    hipStream_t st = h->stream;
    const int nm = 4 * nrhs + 1, parts = sumsq_parts(n);
    double* sq = reinterpret_cast<double*>(h->d_norms + nm);
    // the values in CSR order and ||A||_inf once per factorization (the
    // values do not change in between)
    const bool anorm_known = h->anorm >= 0;
    if (!anorm_known) launch_gather_vals(h->d_csr_val, h->d_a, h->d_csr_src, h->csr_nnz, st);
    launch_residual(r, bp, xp, n, nrhs, h->d_csr_ptr, h->d_csr_col, h->d_csr_val, st);
    launch_zero_u64(h->d_norms, nm, st);
    launch_err_norms(h->d_norms, sq, r, xp, bp, d, n, nrhs, st);
    if (!anorm_known) launch_norm_inf(h->d_norms + 4 * nrhs, n, h->d_csr_ptr, h->d_csr_val, st);
    std::vector<unsigned long long> v(nm + 2 * parts * nrhs);
    HIP_TRY(hipMemcpyAsync(v.data(), h->d_norms, v.size() * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    auto val = [&](int i) {
        double x;
        memcpy(&x, &v[i], 8);
        return x;
    };
    auto sum = [&](int first) {
        double s = 0;
        for (int p = 0; p < parts; ++p) s += val(first + p);
        return s;
    };
    auto ratio = [](double num, double den) { return den > 0 ? num / den : (num > 0 ? INFINITY : 0.0); };
    h->err_eta.assign(nrhs, 0.0);
    h->err_res2.assign(nrhs, 0.0);
    if (d) h->err_ferr.assign(nrhs, 0.0);
    if (!anorm_known) h->anorm = val(4 * nrhs);
    const double an = h->anorm;
    double worst = 0, worst2 = 0;
    for (int k = 0; k < nrhs; ++k) {
        const double e = ratio(val(k), an * val(nrhs + k) + val(2 * nrhs + k));
        if (!(e <= worst)) worst = e;  // NaN propagates
        const double rr = sum(nm + k * parts), xx = sum(nm + (nrhs + k) * parts);
        const double e2 = ratio(std::sqrt(rr), std::sqrt(xx));
        if (!(e2 <= worst2)) worst2 = e2;
        h->err_eta[k] = e;
        h->err_res2[k] = e2;
        if (d) h->err_ferr[k] = ratio(val(3 * nrhs + k), val(nrhs + k));
    }
    *eta = worst;
    *res2 = worst2;
    return MFD_OK;
}

constexpr int kMaxRefine = 10;
constexpr double kRefineTarget = 1e-14;  // the backward error the FP64 form is held to

bool scaled_form(int precision) { return precision == MFD_FP16_IR || precision == MFD_BF16_IR; }

// d = F^{-1} r for the current lower-precision factor F: the 16-bit forms
// factor S A S, so d = S (S A S)^{-1} S r, with r scaled in place.
int correction_solve(mfd_handle* h, double* r, double* d, int nrhs)
{
    const int n = h->S.n;
    if (scaled_form(h->precision)) launch_scale_rows(r, r, h->d_dscale, n, nrhs, h->stream);
    MFD_TRY(run_solve<float>(h, r, d, nrhs));
    if (scaled_form(h->precision)) launch_scale_rows(d, d, h->d_dscale, n, nrhs, h->stream);
    return MFD_OK;
}

// FP64 iterative refinement of x (h->d_xp) for b (h->d_bp) with the current
// FP32-pool factor: the first solve, then r = b - A x with the FP64 matrix,
// the correction from the factor and x += d, until the backward error
// reaches the target or stops halving.  Returns the iterations taken.
int refine(mfd_handle* h, int nrhs, double* eta, double* res2, int* iters)
{
    const int64_t n = h->S.n;
    hipStream_t st = h->stream;
    if (scaled_form(h->precision)) {
        // x = S (S A S)^{-1} S b, with S b formed in the residual buffer
        launch_scale_rows(h->d_r, h->d_bp, h->d_dscale, (int)n, nrhs, st);
        MFD_TRY(run_solve<float>(h, h->d_r, h->d_xp, nrhs));
        launch_scale_rows(h->d_xp, h->d_xp, h->d_dscale, (int)n, nrhs, st);
    } else {
        MFD_TRY(run_solve<float>(h, h->d_bp, h->d_xp, nrhs));
    }
    double prev = INFINITY;
    int it = 0;
    for (;;) {
        MFD_TRY(backward_error(h, h->d_bp, h->d_xp, h->d_r, nrhs, eta, res2));
        if (!(*eta > kRefineTarget) || it >= kMaxRefine || !(*eta < 0.5 * prev)) break;
        prev = *eta;
        MFD_TRY(correction_solve(h, h->d_r, h->d_d, nrhs));
        launch_axpy(h->d_xp, h->d_d, n * nrhs, st);  // both n x nrhs with ld n
        ++it;
    }
    *iters = it;
    return MFD_OK;
}

// x for b, both postordered (ld n) in device memory.  The lower-precision
// forms are refined in FP64 (refine); when refinement stalls above the
// target the matrix is refactored in the next form -- FP16 / BF16 updates,
// then FP32 fronts, then FP64 -- and solved again.
int solve_permuted_device(mfd_handle* h, int nrhs)
{
    h->st.refine_iters = 0;
    h->st.backward_error = -1;
    h->st.refine_residual = -1;
    if (h->precision == MFD_FP64) {
        MFD_TRY(run_solve<double>(h, h->d_bp, h->d_xp, nrhs));
        return MFD_OK;
    }
    const int first = h->precision;
    double eta = 0, res2 = 0;
    int total = 0;
    for (;;) {
        int it = 0;
        MFD_TRY(refine(h, nrhs, &eta, &res2, &it));
        total += it;
        if (eta <= kRefineTarget) break;
        // refinement stalled: refactor in the next form
        const int next = h->precision == MFD_FP32_IR ? MFD_FP64 : MFD_FP32_IR;
        int rc = factorize_common(h, next);
        if (rc != MFD_OK) return rc;
        if (h->precision == MFD_FP64) {
            MFD_TRY(run_solve<double>(h, h->d_bp, h->d_xp, nrhs));
            MFD_TRY(backward_error(h, h->d_bp, h->d_xp, h->d_r, nrhs, &eta, &res2));
            break;
        }
    }
    h->st.backward_error = eta;
    h->st.refine_residual = res2;
    // negative: the form changed (-1 - iterations taken before the last change)
    h->st.refine_iters = h->precision == first ? total : -1 - total;
    return MFD_OK;
}

// Copy b (ldb x nrhs) to the device, apply the analysis permutation, solve,
// undo the permutation and copy back (hppllt.f90:152-173).
int solve_host(mfd_handle* h, double* b, int nrhs, int ldb, bool permuted, bool on_device)
{
    if (!h->factored) return MFD_ERR_STATE;
    if (!b || nrhs < 1 || ldb < h->S.n) return MFD_ERR_ARG;
    MFD_TRY(ensure_rhs(h, nrhs));
    const int n = h->S.n;
    hipStream_t st = h->stream;
    double t0 = now_ms();
    const hipMemcpyKind kin = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind kout = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    HIP_TRY(hipMemcpy2DAsync(h->d_bin, (size_t)n * 8, b, (size_t)ldb * 8, (size_t)n * 8, nrhs, kin, st));
    HIP_TRY(hipEventRecord(h->ev[2], st));
    if (permuted)
        HIP_TRY(hipMemcpyAsync(h->d_bp, h->d_bin, (size_t)n * nrhs * 8, hipMemcpyDeviceToDevice, st));
    else
        launch_gather_rows(h->d_bp, n, h->d_bin, n, h->d_perm, n, nrhs, st);
    MFD_TRY(solve_permuted_device(h, nrhs));
    if (permuted)
        HIP_TRY(hipMemcpyAsync(h->d_bin, h->d_xp, (size_t)n * nrhs * 8, hipMemcpyDeviceToDevice, st));
    else
        launch_scatter_rows(h->d_bin, n, h->d_xp, n, h->d_perm, n, nrhs, st);
    HIP_TRY(hipEventRecord(h->ev[3], st));
    HIP_TRY(hipMemcpy2DAsync(b, (size_t)ldb * 8, h->d_bin, (size_t)n * 8, (size_t)n * 8, nrhs, kout, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipGetLastError());
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev[2], h->ev[3]));
    h->st.solve_ms = ms;
    h->times[MFD_TIME_SOLVE] = (now_ms() - t0) / 1e3;
    h->solved_nrhs = nrhs;
    return MFD_OK;
}

// ----------------------------------------------------- fixed solve scratch
// mfd_solve_device_noalloc takes every solve buffer from this array instead
// of the handle's allocations.  Each GPU holds its own copy, at its own
// address, from the library's load until the process ends.  In host code
// the array's name is not a device address: the address comes from
// hipGetSymbolAddress, once per device (scratch_on_current_device).
constexpr int64_t kScratchBytes = int64_t(128) << 20;
static __device__ __align__(256) double scratch[kScratchBytes / 8];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// The scratch array's device address on the current device (resolved on the
// first call for that device), or nullptr if HIP cannot resolve it.
double* scratch_on_current_device()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        g_scratch[dev] = static_cast<double*>(p);
    }
    return g_scratch[dev];
}

// Bytes of scratch a solve of nrhs right-hand sides takes (each buffer
// 256-byte aligned): Z/Y slices, partial sums, permuted b and x and, for
// the refined forms, residual, correction and norm slots.
int64_t noalloc_bytes(const mfd_handle* h, int nrhs)
{
    auto al = [](int64_t b) { return (b + 255) & ~int64_t(255); };
    const int64_t n = h->S.n;
    int64_t b = al(h->z_per_rhs * nrhs * 8) + al(h->part_slots * 64 * nrhs * 8) + 2 * al(n * nrhs * 8);
    if (h->precision != MFD_FP64) b += 2 * al(n * nrhs * 8) + al((4 * nrhs + 1 + 2 * kSumParts * nrhs) * 8);
    return b;
}

// For the duration of a no-allocation solve: the handle's solve buffers
// point into the scratch array and no_alloc is set; both restored after.
struct NoAllocScope {
    mfd_handle* h;
    double *z, *part, *bp, *xp, *r, *d;
    unsigned long long* norms;
    NoAllocScope(mfd_handle* h_, int nrhs) : h(h_)
    {
        z = h->d_z, part = h->d_part, bp = h->d_bp, xp = h->d_xp, r = h->d_r, d = h->d_d, norms = h->d_norms;
        auto al = [](int64_t b) { return (b + 255) & ~int64_t(255); };
        const int64_t n = h->S.n;
        char* p = reinterpret_cast<char*>(h->scratch);
        auto take = [&](int64_t bytes) {
            char* q = p;
            p += al(bytes);
            return q;
        };
        h->d_z = reinterpret_cast<double*>(take(h->z_per_rhs * nrhs * 8));
        h->d_part = reinterpret_cast<double*>(take(h->part_slots * 64 * nrhs * 8));
        h->d_bp = reinterpret_cast<double*>(take(n * nrhs * 8));
        h->d_xp = reinterpret_cast<double*>(take(n * nrhs * 8));
        if (h->precision != MFD_FP64) {
            h->d_r = reinterpret_cast<double*>(take(n * nrhs * 8));
            h->d_d = reinterpret_cast<double*>(take(n * nrhs * 8));
            h->d_norms = reinterpret_cast<unsigned long long*>(take((4 * nrhs + 1 + 2 * kSumParts * nrhs) * 8));
        }
        h->no_alloc = true;
    }
    ~NoAllocScope()
    {
        h->d_z = z, h->d_part = part, h->d_bp = bp, h->d_xp = xp, h->d_r = r, h->d_d = d, h->d_norms = norms;
        h->no_alloc = false;
    }
};

int finish_analysis(mfd_handle* h)
{
    h->front_kind.clear();  // per-node choices belong to the previous structure
    delete h->layout;
    h->layout = new RefLayout(h->S.c, h->S.w, h->nb);
    h->analyzed = true;
    h->factored = false;
    if (!h->device_ok) return MFD_OK;
    int rc = setup_device(h);
    if (rc != MFD_OK) {
        free_device(h);
        h->analyzed = false;
    }
    return rc;
}

}  // namespace

// ------------------------------------------------ static schedule check
// The factorization schedule as enqueue_factor issues it on the graph
// path -- main stream, fill stream (zeroing and assembly by level groups),
// side stream (large in-panel updates) and their event edges -- turned into
// a DAG of launches, and every task's accesses: rectangles of fronts, the
// inverses of their 64 x 64 diagonal blocks, and the parent entries an
// extend-add adds into (the bounding rectangle of the local_index images).
// The check: no two tasks of one launch touch one entry where either
// writes, and every such pair in two launches is ordered by the DAG.
// Variant 1 models a schedule that is not built: the Schur tiles of the
// fronts whose chain ends before their level's last step issued on a fourth
// stream forked after the launch with the front's last diagonal task and
// joined at the level's end.
namespace {

struct Access {
    int region;  // front s, or num_node + s for its diagonal-block inverses (rows = block index)
    int r0, r1, c0, c1;  // [r0, r1) x [c0, c1)
    int node;            // DAG node of the launch
    int64_t task;        // global task id (unique across launches)
    bool write;
    int kind;            // MFD_K_* of the launch, -1 zeroing/assembly, -2 a Schur tile of variant 1's fourth stream
};

struct ScheduleCheck {
    const mfd_handle* h;
    int m = 0;
    std::vector<std::vector<int>> pred;  // DAG: predecessors of each node
    std::vector<Access> acc;
    std::vector<std::pair<int64_t, int>> starts;  // pool offset -> region (fronts, then inverses)
    int64_t next_task = 0;
    int kind = 0;  // kind recorded with the accesses being added

    int add_node() {
        pred.emplace_back();
        return (int)pred.size() - 1;
    }
    void edge(int a, int b) {
        if (a >= 0 && b >= 0 && a != b) pred[b].push_back(a);
    }
    void rect(int region, int r0, int r1, int c0, int c1, int node, int64_t task, bool write) {
        if (r1 > r0 && c1 > c0) acc.push_back(Access{region, r0, r1, c0, c1, node, task, write, kind});
    }
    // region and (row, col) of a pool offset
    void locate(int64_t off, int& region, int& row, int& col) const {
        auto it = std::upper_bound(starts.begin(), starts.end(), std::make_pair(off, INT_MAX));
        --it;
        region = it->second;
        const int64_t rel = off - it->first;
        if (region < m) {
            const int n = h->nodes[region].n;
            row = (int)(rel % n);
            col = (int)(rel / n);
        } else {
            row = (int)(rel / (kPanel * kPanel));
            col = 0;
        }
    }
    void front_all(int s, int node, int64_t task) {
        const NodeDev& d = h->nodes[s];
        rect(s, 0, d.n, 0, d.n, node, task, true);
        rect(m + s, 0, (d.c + kPanel - 1) / kPanel, 0, 1, node, task, true);
    }
    // parent entries of child ch's update rows [i0, i1) x columns [j0, j1)
    void parent_rect(int ch, int i0, int i1, int j0, int j1, int node, int64_t task) {
        const NodeDev& d = h->nodes[ch];
        if (d.parent < 0 || i1 <= i0 || j1 <= j0) return;
        const int* li = h->S.lidx.data() + d.lidx;
        rect(d.parent, li[i0] - 1, li[i1 - 1], li[j0] - 1, li[j1 - 1], node, task, true);
    }
    void gemm(const GemmTask& g, int node, int64_t task) {
        int rc, r, c;
        locate(g.c, rc, r, c);
        rect(rc, r, r + g.m, c, c + g.n, node, task, true);
        int ra, ar, ac;
        locate(g.a, ra, ar, ac);
        rect(ra, ar, ar + g.m, ac, ac + g.k, node, task, false);
        int rb, br, bc;
        locate(g.b, rb, br, bc);
        if (rb >= m)
            rect(rb, br, br + 1, 0, 1, node, task, false);  // a stored inverse block
        else
            rect(rb, br, br + g.n, bc, bc + g.k, node, task, false);
        if (g.mode == 2) {
            int rd, dr, dc;
            locate(g.d, rd, dr, dc);
            rect(rd, dr, dr + 1, 0, 1, node, task, false);
        }
        if (g.syrk > 0) {  // the diagonal block of the tile's rows
            int rs, sr, sc;
            locate(g.cd, rs, sr, sc);
            rect(rs, sr, sr + g.syrk, sc, sc + g.syrk, node, task, true);
        }
        if (g.mode == 3) {
            const NodeDev& d = h->nodes[rc];
            parent_rect(rc, r - d.c, r - d.c + g.m, c - d.c, c - d.c + g.n, node, task);
        }
    }
    void diag(const DiagTask& t, int node, int64_t task) {
        const int s = t.node, k0 = t.k0, kb = t.kb, kp = t.kprev, kpp = t.kpp;
        rect(s, k0, k0 + kb, k0, k0 + kb, node, task, true);
        rect(m + s, k0 / kPanel, k0 / kPanel + 1, 0, 1, node, task, true);
        if (kp > 0) {
            rect(s, k0, k0 + kb, k0 - kp, k0, node, task, true);
            rect(m + s, (k0 - kp) / kPanel, (k0 - kp) / kPanel + 1, 0, 1, node, task, false);
        }
        if (kpp > 0) {
            rect(s, k0, k0 + kb, k0 - kp - kpp, k0 - kp, node, task, false);
            rect(s, k0 - kp, k0, k0 - kp - kpp, k0 - kp, node, task, false);
        }
    }
    void launch(const Launch& l, int node) {
        kind = l.kind;
        switch (l.kind) {
        case MFD_K_EXTEND_ADD:
            for (int64_t i = l.first; i < l.first + l.count; ++i) {
                const ExtTask& t = h->ext[i];
                const NodeDev& d = h->nodes[t.child];
                const int64_t task = next_task++;
                const int i0 = t.ti * kTile, i1 = std::min(d.w, i0 + kTile);
                const int j0 = t.tj * kTile, j1 = std::min(d.w, j0 + kTile);
                rect(t.child, d.c + i0, d.c + i1, d.c + j0, d.c + j1, node, task, false);
                parent_rect(t.child, i0, i1, j0, j1, node, task);
            }
            break;
        case MFD_K_SMALL_FRONT:
            for (int64_t i = l.first; i < l.first + l.count; ++i) front_all(h->small[i], node, next_task++);
            break;
        case MFD_K_FRONT:
            for (int64_t i = l.first; i < l.first + l.count; ++i) front_all(h->fused[i], node, next_task++);
            break;
        case MFD_K_POTRF: {
            const int64_t d0 = l.fused ? l.dfirst : l.first, dn = l.fused ? l.dcount : l.count;
            for (int64_t i = d0; i < d0 + dn; ++i) diag(h->diag[i], node, next_task++);
            if (l.fused)
                for (int64_t i = l.first; i < l.first + l.count; ++i) gemm(h->gemm[i], node, next_task++);
            break;
        }
        default:
            for (int64_t i = l.first; i < l.first + l.count; ++i) {
                const GemmTask& g = h->gemm[i];
                gemm(g, node, next_task++);
            }
            break;
        }
    }
};

// Reachability over the DAG (nodes in enqueue order: predecessors first).
struct Reach {
    int n, words;
    std::vector<uint64_t> bits;
    explicit Reach(const std::vector<std::vector<int>>& pred) : n((int)pred.size()), words(((int)pred.size() + 63) / 64) {
        bits.assign((size_t)n * words, 0);
        for (int v = 0; v < n; ++v)
            for (int u : pred[v]) {
                uint64_t* bv = &bits[(size_t)v * words];
                const uint64_t* bu = &bits[(size_t)u * words];
                for (int w = 0; w < words; ++w) bv[w] |= bu[w];
                bv[u / 64] |= uint64_t(1) << (u % 64);
            }
    }
    bool before(int a, int b) const { return (bits[(size_t)b * words + a / 64] >> (a % 64)) & 1; }
};

int64_t check_schedule(mfd_handle* h, int variant, std::string& report)
{
    ScheduleCheck C;
    C.h = h;
    C.m = (int)h->nodes.size();
    const int m = C.m;
    for (int s = 0; s < m; ++s) C.starts.push_back({h->nodes[s].front, s});
    for (int s = 0; s < m; ++s)
        if (h->nodes[s].c > 0) C.starts.push_back({h->nodes[s].inv, m + s});
    std::sort(C.starts.begin(), C.starts.end());
    // streams as enqueue_factor issues them (graph path: fill and side streams present)
    const int init = C.add_node();
    int main_last = init;
    const int G = (int)h->group_level.size();
    std::vector<int> fill_node(G);
    int fill_last = init;  // the fill stream forks after the info reset
    for (int g = 0; g < G; ++g) {
        const int v = C.add_node();
        C.edge(fill_last, v);
        fill_last = v;
        fill_node[g] = v;
        C.kind = -1;
        for (int s = 0; s < m; ++s)
            if (h->front_group[s] == g) C.rect(s, 0, h->nodes[s].n, 0, h->nodes[s].n, v, C.next_task++, true);
    }
    // variants 1 and 2: per large front, the launch the fourth stream forks
    // after -- the one with its last diagonal task (1), or the last one with
    // any of its step tasks, TRSM and in-panel update tiles included (2)
    std::vector<int> last_diag_launch(m, -1);
    if (variant >= 1)
        for (size_t i = 0; i < h->fseq.size(); ++i) {
            const Launch& l = h->fseq[i];
            if (l.kind == MFD_K_POTRF) {
                const int64_t d0 = l.fused ? l.dfirst : l.first, dn = l.fused ? l.dcount : l.count;
                for (int64_t j = d0; j < d0 + dn; ++j) last_diag_launch[h->diag[j].node] = (int)i;
            }
            if (variant == 2 && (l.kind == MFD_K_TRSM || l.kind == MFD_K_UPDATE || (l.kind == MFD_K_POTRF && l.fused)))
                for (int64_t j = l.first; j < l.first + l.count; ++j) {
                    int rg, r, c;
                    C.locate(h->gemm[j].c, rg, r, c);
                    if (rg < m) last_diag_launch[rg] = (int)i;
                }
        }
    std::vector<int> side_rec(2, -1);  // side launch last recorded on sev[0/1]
    int side_last = -1, last_side = -1, joined = -1, gnext = 0;
    std::vector<int> launch_node(h->fseq.size(), -1), early_nodes;
    int cur_level = -1;
    for (size_t i = 0; i < h->fseq.size(); ++i) {
        const Launch& l = h->fseq[i];
        int lev = -1;
        for (int L = 0; L < h->num_levels; ++L)
            if (h->level_launch[L] <= (int)i) lev = L;
        if (lev != cur_level) {
            // variant 1: the fourth stream joins before the next level starts
            if (!early_nodes.empty()) {
                const int j = C.add_node();
                C.edge(main_last, j);
                for (int e : early_nodes) C.edge(e, j);
                main_last = j;
                early_nodes.clear();
            }
            cur_level = lev;
        }
        std::vector<int> waits;
        for (; gnext < G && h->level_launch[h->group_level[gnext]] <= (int)i; ++gnext) waits.push_back(fill_node[gnext]);
        if (l.side) {
            const int v = C.add_node();
            C.edge(main_last, v);  // fork from the main stream's position
            for (int w : waits) C.edge(w, v);
            C.edge(side_last, v);
            side_last = v;
            side_rec[l.side_id & 1] = v;
            last_side = l.side_id;
            launch_node[i] = v;
            C.launch(l, v);
            // the waits were recorded on the main stream
            if (!waits.empty()) {
                const int j = C.add_node();
                C.edge(main_last, j);
                for (int w : waits) C.edge(w, j);
                main_last = j;
            }
            continue;
        }
        if (l.kind != MFD_K_POTRF && l.kind != MFD_K_TRSM && last_side > joined) {
            waits.push_back(side_rec[last_side & 1]);
            joined = last_side;
        } else if (l.wait_id > joined) {
            waits.push_back(side_rec[l.wait_id & 1]);
            joined = l.wait_id;
        }
        const int v = C.add_node();
        C.edge(main_last, v);
        for (int w : waits) C.edge(w, v);
        main_last = v;
        launch_node[i] = v;
        if (variant >= 1 && l.kind == MFD_K_SCHUR) {
            // the Schur tiles of fronts whose chain ended early go to the fourth stream
            int steps_end = -1;
            for (size_t k = 0; k < i; ++k)
                if (h->fseq[k].kind == MFD_K_POTRF && (int)k >= h->level_launch[lev]) steps_end = (int)k;
            std::vector<int> early;
            for (int64_t t = l.first; t < l.first + l.count; ++t) {
                int rg, r, c;
                C.locate(h->gemm[t].c, rg, r, c);
                if (rg < m && last_diag_launch[rg] >= 0 && last_diag_launch[rg] < steps_end &&
                    std::find(early.begin(), early.end(), rg) == early.end())
                    early.push_back(rg);
            }
            C.kind = -2;
            for (int s : early) {
                const int e = C.add_node();
                C.edge(launch_node[last_diag_launch[s]], e);
                if (!early_nodes.empty()) C.edge(early_nodes.back(), e);  // one stream: in order
                early_nodes.push_back(e);
                for (int64_t t = l.first; t < l.first + l.count; ++t) {
                    int rg, r, c;
                    C.locate(h->gemm[t].c, rg, r, c);
                    if (rg == s) C.gemm(h->gemm[t], e, C.next_task++);
                }
            }
            // the rest stays on the main stream
            C.kind = MFD_K_SCHUR;
            for (int64_t t = l.first; t < l.first + l.count; ++t) {
                int rg, r, c;
                C.locate(h->gemm[t].c, rg, r, c);
                if (std::find(early.begin(), early.end(), rg) == early.end()) C.gemm(h->gemm[t], v, C.next_task++);
            }
            continue;
        }
        C.launch(l, v);
    }
    Reach R(C.pred);
    // conflicts: accesses bucketed by region and 64 x 64 cell
    std::unordered_map<int64_t, std::vector<int>> cells;
    for (size_t a = 0; a < C.acc.size(); ++a) {
        const Access& x = C.acc[a];
        for (int I = x.r0 / 64; I <= (x.r1 - 1) / 64; ++I)
            for (int J = x.c0 / 64; J <= (x.c1 - 1) / 64; ++J)
                cells[((int64_t)x.region << 40) | ((int64_t)I << 20) | J].push_back((int)a);
    }
    int64_t bad = 0;
    std::set<std::pair<int64_t, int64_t>> seen;
    char line[256];
    for (const auto& kv : cells) {
        const auto& v = kv.second;
        for (size_t p = 0; p < v.size(); ++p)
            for (size_t q = p + 1; q < v.size(); ++q) {
                const Access &x = C.acc[v[p]], &y = C.acc[v[q]];
                if (x.task == y.task || !(x.write || y.write)) continue;
                if (x.r1 <= y.r0 || y.r1 <= x.r0 || x.c1 <= y.c0 || y.c1 <= x.c0) continue;
                const bool ordered = x.node != y.node && (R.before(x.node, y.node) || R.before(y.node, x.node));
                if (ordered) continue;
                const auto key = std::make_pair(std::min(x.task, y.task), std::max(x.task, y.task));
                if (!seen.insert(key).second) continue;
                ++bad;
                if (bad <= 8) {
                    snprintf(line, sizeof line,
                             "%s: region %d, kinds %d (%s) / %d (%s), rows [%d,%d) / [%d,%d), cols [%d,%d) / [%d,%d)\n",
                             x.node == y.node ? "same launch" : "unordered", x.region, x.kind, x.write ? "w" : "r",
                             y.kind, y.write ? "w" : "r", x.r0, x.r1, y.r0, y.r1, x.c0, x.c1, y.c0, y.c1);
                    report += line;
                }
            }
    }
    return bad;
}

}  // namespace

// ================================================================== C-ABI
extern "C" {

int mfd_create(mfd_handle** out, int nb, int flags)
{
    if (!out || nb < 1) return MFD_ERR_ARG;
    *out = nullptr;
    mfd_handle* h = new mfd_handle();
    h->nb = nb;
    h->flags = flags;
    if (const char* lim = getenv("MFD_DEVICE_MEM_LIMIT")) h->mem_limit = atoll(lim);
    int count = 0;
    if (hipGetDeviceCount(&count) == hipSuccess && count > 0) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) == hipSuccess && strncmp(prop.gcnArchName, "gfx950", 6) == 0 &&
            hipSetDevice(0) == hipSuccess && hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) == hipSuccess) {
            h->device_ok = true;
            h->scratch = scratch_on_current_device();  // set-up of mfd_solve_device_noalloc
            for (auto& e : h->ev) (void)hipEventCreate(&e);
            if (hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess) h->side = nullptr;
            if (hipStreamCreateWithFlags(&h->fillst, hipStreamNonBlocking) != hipSuccess) h->fillst = nullptr;
            (void)hipEventCreateWithFlags(&h->fork, hipEventDisableTiming);
            for (auto& e : h->sev) (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
        }
    }
    (void)hipGetLastError();
    *out = h;
    return h->device_ok ? MFD_OK : MFD_ERR_NODEVICE;
}

void mfd_destroy(mfd_handle* h)
{
    if (!h) return;
    if (h->device_ok) {
        (void)hipStreamSynchronize(h->stream);
        free_device(h);
        for (auto& b : h->spare) (void)hipFree(b.first);
        h->spare.clear();
        h->device_bytes = 0;
        for (auto e : h->ev)
            if (e) (void)hipEventDestroy(e);
        if (h->fork) (void)hipEventDestroy(h->fork);
        for (auto e : h->sev)
            if (e) (void)hipEventDestroy(e);
        if (h->side) (void)hipStreamDestroy(h->side);
        for (auto e : h->fill_ev) (void)hipEventDestroy(e);
        if (h->fillst) (void)hipStreamDestroy(h->fillst);
        (void)hipStreamDestroy(h->stream);
    }
    delete h->layout;
    delete h;
}

int mfd_analyze(mfd_handle* h, int n, const int* ccs_col, const int* ccs_row, int max_zero, const int* perm)
{
    if (!h) return MFD_ERR_ARG;
    double t0 = now_ms();
    h->analyzed = h->factored = false;
    h->from_parts = false;
    std::vector<int> nd;
    if (!perm && (h->flags & MFD_FLAG_ORDER_ND)) {  // use_metis = 1 (hppllt.f90:47-51)
        if (n < 1) return MFD_ERR_ARG;
        nd.resize(n);
        const double t1 = now_ms();
        int rc = mfd_order_nd(n, ccs_col, ccs_row, nd.data());
        if (rc != MFD_OK) return rc;
        h->times[MFD_TIME_ORDERING] = (now_ms() - t1) / 1e3;
        perm = nd.data();
    }
    int rc = analyze(n, ccs_col, ccs_row, max_zero, perm, h->S);
    if (rc != MFD_OK) return rc;
    rc = finish_analysis(h);
    h->times[MFD_TIME_ANALYZE] = (now_ms() - t0) / 1e3;
    h->st.analyze_ms = now_ms() - t0;
    return rc;
}

int mfd_analyze_batch(mfd_handle* h, int nsys, const int* n, const int* const* ccs_col, const int* const* ccs_row,
                      int max_zero, const int* const* perm)
{
    if (!h || nsys < 1 || !n || !ccs_col || !ccs_row) return MFD_ERR_ARG;
    double t0 = now_ms();
    h->analyzed = h->factored = false;
    h->from_parts = false;
    std::vector<std::vector<int>> nd;
    std::vector<const int*> perms;
    if (h->flags & MFD_FLAG_ORDER_ND) {  // systems without an ordering get the library's
        const double t1 = now_ms();
        nd.resize(nsys);
        perms.resize(nsys);
        for (int k = 0; k < nsys; ++k) {
            perms[k] = perm ? perm[k] : nullptr;
            if (perms[k]) continue;
            if (n[k] < 1 || !ccs_col[k] || !ccs_row[k]) return MFD_ERR_ARG;
            nd[k].resize(n[k]);
            int rc = mfd_order_nd(n[k], ccs_col[k], ccs_row[k], nd[k].data());
            if (rc != MFD_OK) return rc;
            perms[k] = nd[k].data();
        }
        h->times[MFD_TIME_ORDERING] = (now_ms() - t1) / 1e3;
        perm = perms.data();
    }
    int rc = analyze_batch(nsys, n, ccs_col, ccs_row, max_zero, perm, h->S);
    if (rc != MFD_OK) return rc;
    rc = finish_analysis(h);
    h->times[MFD_TIME_ANALYZE] = (now_ms() - t0) / 1e3;
    h->st.analyze_ms = now_ms() - t0;
    return rc;
}

int mfd_set_structure(mfd_handle* h, int num_node, const int* c, const int* w, const int* parent, const int* lidx,
                      const int64_t* a_colptr, const int* snidx)
{
    if (!h || !a_colptr || !snidx) return MFD_ERR_ARG;
    h->analyzed = h->factored = false;
    int rc = structure_from_parts(num_node, c, w, parent, lidx, a_colptr, snidx, h->S);
    if (rc != MFD_OK) return rc;
    h->from_parts = true;
    return finish_analysis(h);
}

namespace {
// A schedule setting changed: rebuild the schedule of the current analysis
// (its factor is gone).
int replan(mfd_handle* h)
{
    if (!h->analyzed) return MFD_OK;
    h->factored = false;
    if (!h->device_ok) return MFD_OK;
    int rc = setup_device(h);
    if (rc != MFD_OK) {
        free_device(h);
        h->analyzed = false;
    }
    return rc;
}
}  // namespace

int mfd_set_update_period(mfd_handle* h, int period)
{
    if (!h || period < 0 || period > kMaxPeriod) return MFD_ERR_ARG;
    if (period == h->upd_period) return MFD_OK;
    h->upd_period = period;
    return replan(h);
}

int mfd_get_update_period(mfd_handle* h) { return h ? h->upd_period : MFD_ERR_ARG; }


int mfd_check_schedule(mfd_handle* h, int variant, int64_t* conflicts, char* report, int report_len)
{
    if (!h || !conflicts || variant < 0 || variant > 2) return MFD_ERR_ARG;
    if (!h->analyzed) return MFD_ERR_STATE;
    if (!h->device_ok) build_plan(h);  // a handle without a GPU keeps no schedule
    std::string rep;
    *conflicts = check_schedule(h, variant, rep);
    if (report && report_len > 0) {
        const size_t n = std::min(rep.size(), (size_t)report_len - 1);
        memcpy(report, rep.data(), n);
        report[n] = 0;
    }
    return MFD_OK;
}

int mfd_set_level_rule(mfd_handle* h, int rule)
{
    if (!h || rule < 0 || rule > 1) return MFD_ERR_ARG;
    if (rule == h->level_rule) return MFD_OK;
    h->level_rule = rule;
    return replan(h);
}
int mfd_get_level_rule(mfd_handle* h) { return h ? h->level_rule : MFD_ERR_ARG; }

int mfd_set_tile_bound(mfd_handle* h, int min_rows)
{
    if (!h || min_rows < 0) return MFD_ERR_ARG;
    if (min_rows == h->wide_min_rows) return MFD_OK;
    h->wide_min_rows = min_rows;
    return replan(h);
}
int mfd_get_tile_bound(mfd_handle* h) { return h ? h->wide_min_rows : MFD_ERR_ARG; }

int mfd_set_solve_block(mfd_handle* h, int min_c)
{
    if (!h || min_c < 0) return MFD_ERR_ARG;
    if (min_c == h->super_min) return MFD_OK;
    h->super_min = min_c;
    return replan(h);
}
int mfd_get_solve_block(mfd_handle* h) { return h ? h->super_min : MFD_ERR_ARG; }

int mfd_set_host_subtrees(mfd_handle* h, const int* roots, int count, int threads)
{
    if (!h || count < 0 || (count > 0 && !roots) || threads < 1) return MFD_ERR_ARG;
    std::vector<int> r(roots, roots + count);
    if (h->analyzed) {
        // every root in range, none twice and none inside another's subtree
        const int m = h->S.num_node;
        std::vector<char> mark(m, 0);
        for (int x : r) {
            if (x < 0 || x >= m || mark[x]) return MFD_ERR_ARG;
            mark[x] = 1;
        }
        for (int x : r)
            for (int a = h->S.parent[x] - 1; a >= 0; a = h->S.parent[a] - 1)
                if (mark[a]) return MFD_ERR_ARG;
    }
    h->host_roots = r;
    h->host_threads = threads;
    h->host_factored = false;
    if (!h->analyzed) return MFD_OK;
    if (!h->device_ok) {
        build_plan(h);
        return h->host_rc;
    }
    return replan(h);
}

int mfd_factorize_host_subtrees(mfd_handle* h, const double* a)
{
    if (!h || !a) return MFD_ERR_ARG;
    if (!h->analyzed || h->host_roots.empty()) return MFD_ERR_STATE;
    if (!h->device_ok) {
        build_plan(h);
        MFD_TRY(h->host_rc);
        h->hstage_vec.assign(h->hplan.stage_elems, 0.0);
    }
    if (h->host_thr.joinable()) h->host_thr.join();
    const double t0 = now_ms();
    const int info = host_factor(h->hplan, h->S, a, stage_of(h), h->host_threads);
    h->host_ms = now_ms() - t0;
    h->host_factored = true;
    return info;
}

int mfd_get_host_part(mfd_handle* h, double* ms, int* fronts, double* flops)
{
    if (!h) return MFD_ERR_ARG;
    if (ms) *ms = h->host_ms;
    if (fronts) *fronts = (int)h->hplan.fronts.size();
    if (flops) {
        double f = 0;
        for (const HostFront& x : h->hplan.fronts)
            for (int64_t t = x.n; t > x.w; --t) f += (double)t * t;
        *flops = f;
    }
    return MFD_OK;
}

int mfd_set_front_split(mfd_handle* h, int cmax)
{
    if (!h || cmax < 0) return MFD_ERR_ARG;
    if (cmax == h->fuse_cmax) return MFD_OK;
    h->fuse_cmax = cmax;
    return replan(h);
}

int mfd_set_front_kinds(mfd_handle* h, const int* kinds)
{
    if (!h) return MFD_ERR_ARG;
    if (!kinds) {
        if (h->front_kind.empty()) return MFD_OK;
        h->front_kind.clear();
        return replan(h);
    }
    if (!h->analyzed) return MFD_ERR_STATE;
    const int m = h->S.num_node;
    for (int s = 0; s < m; ++s)
        if (kinds[s] < -1 || kinds[s] > 1) return MFD_ERR_ARG;
    std::vector<int> k(kinds, kinds + m);
    if (k == h->front_kind) return MFD_OK;
    h->front_kind = std::move(k);
    return replan(h);
}

int mfd_get_front_times(mfd_handle* h, double* us)
{
    if (!h || !us) return MFD_ERR_ARG;
    if (!h->analyzed) return MFD_ERR_STATE;
    if (!h->device_ok) return MFD_ERR_NODEVICE;
    const int m = h->S.num_node;
    for (int s = 0; s < m; ++s) us[s] = -1.0;
    if (h->fused.empty()) return MFD_OK;
    int khz = 0;
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0));
    if (khz <= 0) return MFD_ERR_UNSUPPORTED;
    std::vector<long long> c(2 * h->fused.size());
    HIP_TRY(hipMemcpy(c.data(), h->d_clocks, c.size() * sizeof(long long), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < h->fused.size(); ++i) us[h->fused[i]] = (double)(c[2 * i + 1] - c[2 * i]) * 1e3 / khz;
    return MFD_OK;
}

int mfd_factorize(mfd_handle* h, const double* a, int precision)
{
    if (!h || !a) return MFD_ERR_ARG;
    if (!h->analyzed) return MFD_ERR_STATE;
    if (!h->device_ok) return MFD_ERR_NODEVICE;
    double t0 = now_ms();
    const size_t nnz = h->S.snidx.size();
    HIP_TRY(hipMemcpyAsync(h->d_a, a, nnz * 8, hipMemcpyHostToDevice, h->stream));
    h->host_vals = a;
    int rc = factorize_common(h, precision);
    h->host_vals = nullptr;
    h->times[MFD_TIME_FACTORIZE] = (now_ms() - t0) / 1e3;
    h->times[MFD_TIME_REAL_FACTORIZE] = h->st.factor_ms / 1e3;
    return rc;
}

int mfd_factorize_device(mfd_handle* h, const double* d_a, int precision)
{
    if (!h || !d_a) return MFD_ERR_ARG;
    if (!h->analyzed) return MFD_ERR_STATE;
    if (!h->device_ok) return MFD_ERR_NODEVICE;
    double t0 = now_ms();
    if (d_a != h->d_a)
        HIP_TRY(hipMemcpyAsync(h->d_a, d_a, h->S.snidx.size() * 8, hipMemcpyDeviceToDevice, h->stream));
    if (!h->hplan.roots.empty()) {  // the host part reads the values on the host
        h->host_vals_buf.resize(h->S.snidx.size());
        HIP_TRY(hipMemcpyAsync(h->host_vals_buf.data(), h->d_a, h->S.snidx.size() * 8, hipMemcpyDeviceToHost,
                               h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        h->host_vals = h->host_vals_buf.data();
    }
    int rc = factorize_common(h, precision);
    h->host_vals = nullptr;
    h->times[MFD_TIME_FACTORIZE] = (now_ms() - t0) / 1e3;
    h->times[MFD_TIME_REAL_FACTORIZE] = h->st.factor_ms / 1e3;
    return rc;
}

int mfd_solve(mfd_handle* h, double* b, int nrhs, int ldb)
{
    if (!h) return MFD_ERR_ARG;
    return solve_host(h, b, nrhs, ldb, false, false);
}

int mfd_solve_permuted(mfd_handle* h, double* b, int nrhs, int ldb)
{
    if (!h) return MFD_ERR_ARG;
    return solve_host(h, b, nrhs, ldb, true, false);
}

int mfd_solve_device(mfd_handle* h, double* d_b, int nrhs, int ldb)
{
    if (!h) return MFD_ERR_ARG;
    return solve_host(h, d_b, nrhs, ldb, false, true);
}

int mfd_solve_device_noalloc(mfd_handle* h, double* d_b, int nrhs, int ldb)
{
    if (!h) return MFD_ERR_ARG;
    if (!h->device_ok) return MFD_ERR_NODEVICE;
    if (!h->factored) return MFD_ERR_STATE;
    if (!d_b || nrhs < 1 || ldb < h->S.n) return MFD_ERR_ARG;
    if (!h->scratch) return MFD_ERR_HIP;
    if (noalloc_bytes(h, nrhs) > kScratchBytes) return MFD_ERR_SCRATCH;
    const int n = h->S.n;
    hipStream_t st = h->stream;
    double t0 = now_ms();
    int rc;
    {
        NoAllocScope scope(h, nrhs);
        HIP_TRY(hipEventRecord(h->ev[2], st));
        launch_gather_rows(h->d_bp, n, d_b, ldb, h->d_perm, n, nrhs, st);
        rc = solve_permuted_device(h, nrhs);
        if (rc == MFD_OK) launch_scatter_rows(d_b, ldb, h->d_xp, n, h->d_perm, n, nrhs, st);
        HIP_TRY(hipEventRecord(h->ev[3], st));
        HIP_TRY(hipStreamSynchronize(st));
        HIP_TRY(hipGetLastError());
    }
    if (rc != MFD_OK) return rc;
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev[2], h->ev[3]));
    h->st.solve_ms = ms;
    h->times[MFD_TIME_SOLVE] = (now_ms() - t0) / 1e3;
    h->solved_nrhs = 0;  // mfd_get_solve_errors reads the handle's own buffers, not this call's
    return MFD_OK;
}

int mfd_get_solve_errors(mfd_handle* h, int nrhs, double* eta, double* res2, double* ferr)
{
    if (!h) return MFD_ERR_ARG;
    if (!h->device_ok) return MFD_ERR_NODEVICE;
    if (!h->factored || h->solved_nrhs == 0) return MFD_ERR_STATE;
    if (nrhs != h->solved_nrhs) return MFD_ERR_ARG;
    double e = 0, e2 = 0;
    MFD_TRY(backward_error(h, h->d_bp, h->d_xp, h->d_r, nrhs, &e, &e2));
    if (ferr) {
        // d = A^{-1} (b - A x) by the current factor, then ||d|| / ||x||
        if (h->precision == MFD_FP64)
            MFD_TRY(run_solve<double>(h, h->d_r, h->d_d, nrhs));
        else
            MFD_TRY(correction_solve(h, h->d_r, h->d_d, nrhs));
        MFD_TRY(backward_error(h, h->d_bp, h->d_xp, h->d_r, nrhs, &e, &e2, h->d_d));
    }
    HIP_TRY(hipGetLastError());
    for (int k = 0; k < nrhs; ++k) {
        if (eta) eta[k] = h->err_eta[k];
        if (res2) res2[k] = h->err_res2[k];
        if (ferr) ferr[k] = h->err_ferr[k];
    }
    return MFD_OK;
}

int mfd_export_front(mfd_handle* h, int node, double* out)
{
    if (!h || !out) return MFD_ERR_ARG;
    if (node < 1 || node > h->S.num_node) return MFD_ERR_ARG;
    // a front of a host subtree, from the host staging array
    const bool host = h->host_factored && (int)h->hplan.is_host.size() == h->S.num_node && h->hplan.is_host[node - 1];
    if (!h->factored && !host) return MFD_ERR_STATE;
    const NodeDev& d = h->nodes[node - 1];
    const int64_t cnt = (int64_t)d.n * d.n;
    if (host) {
        for (const HostFront& f : h->hplan.fronts)
            if (f.node == node - 1) memcpy(out, stage_of(h) + f.front, cnt * sizeof(double));
    } else if (h->precision == MFD_FP64) {
        HIP_TRY(hipMemcpy(out, h->d_pool64 + d.front, cnt * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(cnt);
        HIP_TRY(hipMemcpy(tmp.data(), h->d_pool32 + d.front, cnt * 4, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < cnt; ++i) out[i] = tmp[i];
    }
    // only the lower part is kept on the device; the rest reads as zero
    for (int q = 1; q < d.n; ++q)
        for (int r = 0; r < q; ++r) out[(int64_t)q * d.n + r] = 0.0;
    return MFD_OK;
}

int mfd_export_tiles(mfd_handle* h, double* supernode, double* work)
{
    if (!h) return MFD_ERR_ARG;
    if (!h->factored) return MFD_ERR_STATE;
    std::vector<double> front;
    for (int s = 0; s < h->S.num_node; ++s) {
        const NodeDev& d = h->nodes[s];
        front.assign((size_t)d.n * d.n, 0.0);
        int rc = mfd_export_front(h, s + 1, front.data());
        if (rc != MFD_OK) return rc;
        h->layout->put_front(s, front.data(), supernode, work);
    }
    return MFD_OK;
}

int64_t mfd_get_cost(mfd_handle* h) { return h && h->analyzed ? mfd::cost(h->S) : -1; }
int64_t mfd_get_nonzero(mfd_handle* h) { return h && h->analyzed ? mfd::nonzero(h->S) : -1; }
int64_t mfd_get_array_size(mfd_handle* h, int which) { return h && h->layout ? h->layout->size(which) : -1; }
double mfd_get_time_info(mfd_handle* h, int id)
{
    if (!h || id < 1 || id > 7) return -1;
    return h->times[id];
}

int mfd_get_dims(mfd_handle* h, int* n, int* num_node, int64_t* nnz, int64_t* nlidx)
{
    if (!h) return MFD_ERR_ARG;
    if (!h->analyzed) return MFD_ERR_STATE;
    if (n) *n = h->S.n;
    if (num_node) *num_node = h->S.num_node;
    if (nnz) *nnz = (int64_t)h->S.snidx.size();
    if (nlidx) *nlidx = (int64_t)h->S.lidx.size();
    return MFD_OK;
}

int mfd_get_structure(mfd_handle* h, int* c, int* w, int* parent, int* perm, int* lidx, int* snidx, int64_t* ccs_perm,
                      int64_t* a_colptr)
{
    if (!h) return MFD_ERR_ARG;
    if (!h->analyzed) return MFD_ERR_STATE;
    const Structure& S = h->S;
    auto cp = [](auto* dst, const auto& v) {
        if (dst) std::copy(v.begin(), v.end(), dst);
    };
    cp(c, S.c);
    cp(w, S.w);
    cp(parent, S.parent);
    cp(perm, S.perm);
    cp(lidx, S.lidx);
    cp(snidx, S.snidx);
    cp(ccs_perm, S.ccs_perm);
    cp(a_colptr, S.a_colptr);
    return MFD_OK;
}

int mfd_get_stats(mfd_handle* h, mfd_stats* st)
{
    if (!h || !st) return MFD_ERR_ARG;
    h->st.front_bytes = h->front_bytes;
    h->st.device_bytes = h->device_bytes;
    h->st.num_levels = h->num_levels;
    h->st.num_launches_factor = (int)h->fseq.size() + 3;
    h->st.num_launches_solve = (int)(h->fwd_seq.size() + h->bwd_seq.size()) + 2;
    *st = h->st;
    return MFD_OK;
}

int mfd_get_kernel_time(mfd_handle* h, int k, double* total_ms, int* launches)
{
    if (!h || k < 0 || k >= MFD_K_NUM) return MFD_ERR_ARG;
    if (total_ms) *total_ms = h->ktime[k];
    if (launches) *launches = h->kcount[k];
    return MFD_OK;
}

int mfd_get_kernel_work(mfd_handle* h, int k, int s, double* flops, double* bytes)
{
    if (!h || k < 0 || k >= MFD_K_NUM || s < 1) return MFD_ERR_ARG;
    if (!h->analyzed) return MFD_ERR_STATE;
    double f = 0, b = 0;
    auto tri = [](double x) { return x * (x + 1) / 2; };
    // operands read by more than one tile of a launch (the row blocks of a
    // panel, a diagonal block's inverse) count once per launch
    std::set<int64_t> seen;
    auto once = [&](int64_t at, double elems) {
        if (seen.insert(at).second) b += s * elems;
    };
    // a GEMM tile of a step launch
    auto step_tile = [&](const GemmTask& g) {
        if (g.syrk > 0) {  // the diagonal block of its rows: R R^T, K = the columns of A and C
            const double kk = g.mode == 2 ? g.k + g.kd : g.k;
            f += (double)g.syrk * (g.syrk + 1) * kk;
            b += s * 2.0 * tri(g.syrk);  // R: the tile's own rows, counted below
        }
        if (g.mode == 1) {  // C = C L^{-T}
            f += (double)g.m * g.n * g.n;
            b += s * 2.0 * g.m * g.n;
            once(g.b, (double)g.n * g.n);
        } else if (g.mode == 2) {  // C = (C - A B^T) L^{-T}
            f += 2.0 * g.m * g.n * g.k + (double)g.m * g.n * g.n;
            b += s * 2.0 * g.m * g.n;
            once(g.a, (double)g.m * g.k);
            once(g.b, (double)g.n * g.k);
            once(g.d, (double)g.n * g.n);
        } else {
            double e = g.tri ? (double)g.n * g.m - tri(g.n - 1) : (double)g.m * g.n;
            f += 2.0 * g.k * e;
            b += s * 2.0 * e;
            once(g.a, (double)g.m * g.k);
            once(g.b, (double)g.n * g.k);
        }
    };
    switch (k) {
    case MFD_K_ASSEMBLE: {
        double nnz = (double)h->S.snidx.size();
        b = s * (double)h->pool_elems + nnz * (8 + 8 + 8 + s);
        break;
    }
    case MFD_K_EXTEND_ADD:  // small and medium children (large ones add in their Schur launch)
        for (size_t q = 0; q < h->nodes.size(); ++q) {
            const NodeDev& d = h->nodes[q];
            if (d.parent >= 0 && (d.n <= kSmallFront || h->fused_pos[q] >= 0)) {
                double e = tri(d.w);
                f += e;
                b += e * 3 * s + 4.0 * d.w;
            }
        }
        break;
    case MFD_K_FRONT:  // the cost formula's flops of each medium front; its lower triangle read and written
        for (int s0 : h->fused) {
            const NodeDev& d = h->nodes[s0];
            for (int64_t m = d.w + 1; m <= d.n; ++m) f += (double)m * m;
            b += 2.0 * s * tri(d.n);
        }
        break;
    case MFD_K_SMALL_FRONT:
        for (int s0 : h->small) {
            const NodeDev& d = h->nodes[s0];
            for (int64_t m = d.w + 1; m <= d.n; ++m) f += (double)m * m;
            b += 2.0 * s * tri(d.n);
        }
        break;
    case MFD_K_POTRF:  // the per-step launches: diagonal blocks (with their
                       // panel updates and TRSM), panel TRSM tiles, small updates
        for (const Launch& l : h->fseq) {
            if (l.kind != k) continue;
            const int64_t d0 = l.fused ? l.dfirst : l.first, dn = l.fused ? l.dcount : l.count;
            for (int64_t i = d0; i < d0 + dn; ++i) {
                const DiagTask& t = h->diag[i];
                const double kb = t.kb, kp = t.kprev, kpp = t.kpp;
                f += 2.0 * kb * kb * kb / 3.0 + kb * (kb + 1) * (kp + (t.applied ? 0 : kpp)) + 2.0 * kb * kp * kpp +
                     kb * kp * kp;
                b += s * (2.0 * tri(t.kb) + kb * kb + 2.0 * kb * kp + kp * kp + kpp * (kb + kp));
            }
            if (!l.fused) continue;
            seen.clear();
            for (int64_t i = l.first; i < l.first + l.count; ++i) step_tile(h->gemm[i]);
        }
        break;
    case MFD_K_TRSM:  // a step's GEMM tiles launched apart from its diagonal tasks
        for (const Launch& l : h->fseq) {
            if (l.kind != k) continue;
            seen.clear();
            for (int64_t i = l.first; i < l.first + l.count; ++i) step_tile(h->gemm[i]);
        }
        break;
    case MFD_K_UPDATE:
    case MFD_K_SCHUR:
        for (const Launch& l : h->fseq) {
            if (l.kind != k) continue;
            seen.clear();
            for (int64_t i = l.first; i < l.first + l.count; ++i) {
                const GemmTask& g = h->gemm[i];
                double e = g.tri ? (double)g.n * g.m - tri(g.n - 1) : (double)g.m * g.n;
                f += 2.0 * g.k * e;
                b += s * 2.0 * e;  // C read and written
                once(g.a, (double)g.m * g.k);  // the panel's row blocks, once each
                once(g.b, (double)g.n * g.k);
                if (g.mode == 3) {  // the extend-add: e adds into the parent, plus the index maps
                    f += e;
                    b += 2.0 * s * e + 4.0 * (g.m + g.n);
                }
            }
        }
        break;
    case MFD_K_SOLVE_FWD:
    case MFD_K_SOLVE_BWD:
        for (const auto& d : h->nodes) {
            double lent = tri(d.c) + (double)d.w * d.c;
            f += 2.0 * lent;
            b += s * lent + 8.0 * 3 * d.n;
        }
        break;
    default: break;
    }
    if (flops) *flops = f;
    if (bytes) *bytes = b;
    return MFD_OK;
}

const char* mfd_error_string(int code)
{
    switch (code) {
    case MFD_OK: return "ok";
    case MFD_ERR_ARG: return "invalid argument";
    case MFD_ERR_STATE: return "call out of order";
    case MFD_ERR_HIP: return "HIP runtime error";
    case MFD_ERR_NOMEM: return "device memory budget exhausted";
    case MFD_ERR_REDUCIBLE: return "matrix is reducible (elimination forest)";
    case MFD_ERR_INTERNAL: return "internal structure inconsistency";
    case MFD_ERR_NODEVICE: return "no gfx950 device";
    case MFD_ERR_UNSUPPORTED: return "unsupported option";
    case MFD_ERR_NUMERIC: return "non-finite input";
    default: return code > 0 ? "matrix is not positive definite" : "unknown error";
    }
}

}  // extern "C"
