// Batched execution of many independent NMF problems with grouped launches.
//
// The reference runs independent problems one after another (the block "groups"
// of FacesNMFDlg.cpp:582-656, each a separate NMF object and a separate launch
// sequence).  Here every registered problem advances with the same six grouped
// launches per iteration (H pass per rule, H update, W pass per rule, W update);
// the per-problem arguments live in a device table that the host refreshes when
// the set of problems changes.  Because the launches read everything from that
// table, a hipGraph captured once stays valid across changes of the set.
//
// Autotuner (gpumlib_nmf_batch_autotune): the tile shape of the pass work items -- along K
// (the engine: whole contraction per item, or K slices whose depth follows the per-problem
// workgroup target), across the columns (H pass) / rows (W pass) of V (128 or 64 per item)
// and across the ranks (r <= 32 on the one-rank-block kernels, or not), each chosen per
// shape bucket of problems -- and the persistent grid sizes of the grouped launches are
// chosen by timing trial iterations of the current mix on the device.  W and H are saved before and
// restored after every trial, so tuning changes no problem's state; the winner is
// cached per mix signature.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <new>
#include <utility>
#include <vector>

#include "../../include/gpumlib_nmf.h"
#include "nmf_common.hpp"
#include "nmf_internal.hpp"
#include "nmf_plan.hpp"

using namespace gpumlib_nmf;

namespace {

// Launch tables.  Split-K engine: the passes (per rule) and the updates.  Fused engine:
// the grams of W and H (Euclidean problems) and the fused passes (per rule).  Both: the
// per-problem objective reduction.
// Every pass table comes in four variants, one per pass kernel shape: bit 0 set, the kernels
// built for one rank block of 32 (problems with r <= 32 when their bucket splits ranks: half
// the accumulator registers and operand LDS), else two (r <= 64); bit 1 set, 64-column (row)
// tiles, else 128.  A problem sits in the variant of its shape.
enum { T_UH = 0, T_UW, T_RED, T_GW, T_GH, T_PASS0 };
enum { P_PH_EUC = 0, P_PH_KL, P_PW_EUC, P_PW_KL, P_FH_EUC, P_FH_KL, P_FW_EUC, P_FW_KL, P_COUNT };
constexpr int kVariants = 4;
constexpr int T_COUNT = T_PASS0 + P_COUNT * kVariants;
constexpr int pass_table(int pk, int v) { return T_PASS0 + pk * kVariants + v; }
constexpr int variant_mb(int v) { return (v & 1) ? 1 : 2; }
constexpr int variant_tile(int v) { return (v & 2) ? 64 : TILE; }
enum { ENGINE_SPLIT = 0, ENGINE_FUSED = 1 };
constexpr int kPassGrid = 1024;  // persistent workgroups of the grouped passes (before tuning)
constexpr int kUpdGrid = 512;
constexpr int kRedGrid = 256;
constexpr int kBatchTargetWorkgroups = 32;  // per problem and pass, before tuning
constexpr int kBatchSizingTarget = 128;     // workspaces are sized for targets up to this

// Candidates of the autotuner.  Targets above the batch's sizing target are skipped
// (a problem's workspace would be too small for that many K slices).
constexpr int kTuneTargets[] = {4, 8, 16, 32, 64, 128};
// Shape buckets of the K-split target: problems of fewer than 2^18 elements of V, fewer than
// 2^21, and the rest (the serving mix spans 4 K to 8 M elements).  Each bucket's problems are
// planned with their own target, so small and large problems can take different K-tile depths.
constexpr int kBuckets = 3;
int shape_bucket(int n, int m) {
    const long e = (long)n * m;
    return e < (1L << 18) ? 0 : (e < (1L << 21) ? 1 : 2);
}
constexpr int kTunePassGrids[] = {512, 1024, 2048};
constexpr int kTuneUpdGrids[] = {256, 512, 1024};

// Tile shape of one shape bucket's problems.
struct BucketCfg {
    int target = kBatchTargetWorkgroups;  // K-split target (split-K engine)
    int tile = TILE;                      // columns (H pass) / rows (W pass) per work item
    int split = 0;                        // 1: problems with r <= 32 run on the one-rank-block kernels
    bool operator==(const BucketCfg& o) const { return target == o.target && tile == o.tile && split == o.split; }
};
typedef BucketCfg BucketSet[kBuckets];

struct Slot {
    bool active = false;
    int rule = 0, n = 0, m = 0, r = 0, bucket = 0;
    int tile = TILE;  // tile width the problem is planned with
    const float* V = nullptr;
    float *W = nullptr, *H = nullptr;
    void* ws = nullptr;
    size_t ws_bytes = 0;
    float eps = 0.f;
    JobArgs args;
};

struct TuneConfig {
    int engine, pass_grid, upd_grid;
    BucketSet bk;
    double ms_per_iteration, ms_default;
    int trials;
};

// Fused engine: one slice per pass, the whole contraction per work item.  The grams of W
// (over the n rows) and of H (over the m columns) come from the gram launches in K slices of
// at most kGramSlice elements, one work item each (so a long contraction no longer holds the
// launch back on one workgroup); the fused passes sum a gram's slices in slice order as they
// stage it.  Workspace use: the W gram's slices (r x rpad floats each) from offset 0, then the
// H gram's, then one objective partial per column tile of the H pass (inside the split-K plan's
// workspace when that is larger: the fused engine never reads the split-K partials and the
// split-K passes rewrite them).
constexpr int kGramSlice = 1024;
struct FusedLayout {
    int gsl_w, gks_w, gsl_h, gks_h;  // gram slices and their length (W gram over n, H gram over m)
    size_t off_gh, off_objp, bytes;
};
FusedLayout fused_layout(int n, int m, int r, int tile) {
    FusedLayout L;
    L.gsl_w = (n + kGramSlice - 1) / kGramSlice;
    L.gks_w = round_up((n + L.gsl_w - 1) / L.gsl_w, 128);
    L.gsl_h = (m + kGramSlice - 1) / kGramSlice;
    L.gks_h = round_up((m + L.gsl_h - 1) / L.gsl_h, 128);
    const size_t g = (size_t)r * round_up(r, 4) * sizeof(float);
    L.off_gh = L.gsl_w * g;
    L.off_objp = (L.off_gh + L.gsl_h * g + 255) / 256 * 256;
    L.bytes = L.off_objp + ((size_t)((m + tile - 1) / tile) * sizeof(double) + 255) / 256 * 256;
    return L;
}
size_t fused_bytes(int n, int m, int r, int tile) { return fused_layout(n, m, r, tile).bytes; }

}  // namespace

struct gpumlib_nmf_batch {
    int maxj = 0;
    int engine = ENGINE_FUSED;             // how the problems are planned and launched
    BucketSet bk;                          // tile shape per shape bucket (new problems take their bucket's)
    int size_target = kBatchSizingTarget;  // target their workspaces are sized for
    int pass_grid = kPassGrid, upd_grid = kUpdGrid;
    std::map<unsigned long long, TuneConfig> tuned;  // winners by mix signature
    std::vector<Slot> slots;
    JobArgs* d_jobs = nullptr;
    int* d_tables = nullptr;
    double* d_obj = nullptr;
    JobArgs* h_jobs = nullptr;  // pinned staging
    int* h_tables = nullptr;
    hipEvent_t staged = nullptr;
    bool staging_busy = false;
    bool dirty = true;
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;
};

static int table_stride() { return group_table_ints(kGroupMaxJobs); }

static void wait_staging(gpumlib_nmf_batch* b) {
    if (b->staging_busy) {
        (void)hipEventSynchronize(b->staged);
        b->staging_busy = false;
    }
}

// Work items of slot sl in a table of base kind kb (0: not in it), and the key the fused tables
// order their problems by (longest first: the grid strides over them in order).
static int table_items(const gpumlib_nmf_batch* b, const Slot& sl, int kind, long* cost) {
    const bool fused = b->engine == ENGINE_FUSED;
    const bool euc = sl.rule == GPUMLIB_NMF_EUCLIDEAN;
    *cost = 0;
    switch (kind) {
        case T_UH: return !fused ? (sl.m + UPD_COLS - 1) / UPD_COLS : 0;
        case T_UW: return !fused ? (sl.n + UPD_ROWS - 1) / UPD_ROWS : 0;
        case T_RED: return 1;
        case T_GW: *cost = (long)sl.args.h.gks * sl.r; return fused && euc ? sl.args.h.gsl : 0;
        case T_GH: *cost = (long)sl.args.w.gks * sl.r; return fused && euc ? sl.args.w.gsl : 0;
        default: break;
    }
    const int pk = (kind - T_PASS0) / kVariants, v = (kind - T_PASS0) % kVariants;
    const bool small = b->bk[sl.bucket].split && sl.r <= 32;
    if (small != ((v & 1) != 0) || sl.tile != variant_tile(v)) return 0;
    switch (pk) {
        case P_PH_EUC: return !fused && euc ? sl.args.h.tiles * sl.args.h.S : 0;
        case P_PH_KL: return !fused && !euc ? sl.args.h.tiles * sl.args.h.S : 0;
        case P_PW_EUC: return !fused && euc ? sl.args.w.tiles * sl.args.w.S : 0;
        case P_PW_KL: return !fused && !euc ? sl.args.w.tiles * sl.args.w.S : 0;
        case P_FH_EUC: *cost = sl.n; return fused && euc ? sl.args.h.tiles : 0;
        case P_FH_KL: *cost = sl.n; return fused && !euc ? sl.args.h.tiles : 0;
        case P_FW_EUC: *cost = sl.m; return fused && euc ? sl.args.w.tiles : 0;
        case P_FW_KL: *cost = sl.m; return fused && !euc ? sl.args.w.tiles : 0;
    }
    return 0;
}

// Rebuilds the launch tables from the active slots and stages them to the device.
static hipError_t upload(gpumlib_nmf_batch* b, hipStream_t st) {
    if (!b->dirty) return hipSuccess;
    wait_staging(b);
    const int stride = table_stride();
    memset(b->h_tables, 0, sizeof(int) * (size_t)stride * T_COUNT);
    std::vector<std::pair<long, int>> order;  // (-cost, slot)
    std::vector<int> count(b->maxj);
    for (int k = 0; k < T_COUNT; k++) {
        int* t = b->h_tables + (size_t)k * stride;
        int J = 0, T = 0;
        int* first = t + 2;
        int* idx = t + 3 + kGroupMaxJobs;
        order.clear();
        for (int s = 0; s < b->maxj; s++) {
            const Slot& sl = b->slots[s];
            if (!sl.active) continue;
            long cost = 0;
            count[s] = table_items(b, sl, k, &cost);
            if (count[s] > 0) order.emplace_back(-cost, s);
        }
        std::stable_sort(order.begin(), order.end(),
                         [](const std::pair<long, int>& x, const std::pair<long, int>& y) { return x.first < y.first; });
        for (const auto& o : order) {
            first[J] = T;
            idx[J] = o.second;
            T += count[o.second];
            J++;
        }
        first[J] = T;
        t[0] = J;
        t[1] = T;
    }
    for (int s = 0; s < b->maxj; s++) b->h_jobs[s] = b->slots[s].args;
    hipError_t e = hipMemcpyAsync(b->d_jobs, b->h_jobs, sizeof(JobArgs) * (size_t)b->maxj, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_tables, b->h_tables, sizeof(int) * (size_t)stride * T_COUNT, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(b->staged, st);
    if (e != hipSuccess) return e;
    b->staging_busy = true;
    b->dirty = false;
    return hipSuccess;
}

// Pass-table variants launched: those some bucket's configuration can put a problem in (whether
// or not one is there now, so that a captured graph stays valid while problems come and go).
static unsigned variants_launched(const gpumlib_nmf_batch* b) {
    unsigned set = 0;
    for (int k = 0; k < kBuckets; k++) {
        const int v = b->bk[k].tile == 64 ? 2 : 0;
        set |= 1u << v;
        if (b->bk[k].split) set |= 1u << (v | 1);
    }
    return set;
}

static hipError_t enqueue(gpumlib_nmf_batch* b, int iterations, int objectives, hipStream_t st) {
    const int stride = table_stride();
    const int* T = b->d_tables;
    const int pg = b->pass_grid, ug = b->upd_grid;
    const unsigned vs = variants_launched(b);
    auto tab = [&](int pk, int v) { return T + (size_t)pass_table(pk, v) * stride; };
    // every launched variant of one pass kind
#define EACH_VARIANT(call)                          \
    for (int v = 0; v < kVariants; v++) {           \
        if (!((vs >> v) & 1)) continue;             \
        const int mb = variant_mb(v), tw = variant_tile(v); \
        if ((e = (call)) != hipSuccess) return e;   \
    }
    if (b->engine == ENGINE_FUSED) {
        // per iteration: gram of W; fused H passes (objective of the entry state fused in);
        // gram of the new H; fused W passes
        for (int it = 0; it < iterations; it++) {
            const bool obj = it < objectives;
            const int oe = obj ? OBJ_FROB : OBJ_NONE, ok = obj ? OBJ_KL : OBJ_NONE;
            hipError_t e;
            if ((e = launch_group_gram(b->d_jobs, T + T_GW * stride, 0, ug, st))) return e;
            EACH_VARIANT(launch_group_fused_h(b->d_jobs, tab(P_FH_EUC, v), 0, oe, mb, tw, pg, st))
            EACH_VARIANT(launch_group_fused_h(b->d_jobs, tab(P_FH_KL, v), 1, ok, mb, tw, pg, st))
            if (obj && (e = launch_group_reduce_obj(b->d_jobs, T + T_RED * stride, it, kRedGrid, st))) return e;
            if ((e = launch_group_gram(b->d_jobs, T + T_GH * stride, 1, ug, st))) return e;
            EACH_VARIANT(launch_group_fused_w(b->d_jobs, tab(P_FW_EUC, v), 0, mb, tw, pg, st))
            EACH_VARIANT(launch_group_fused_w(b->d_jobs, tab(P_FW_KL, v), 1, mb, tw, pg, st))
        }
        return hipSuccess;
    }
    for (int it = 0; it < iterations; it++) {
        const bool obj = it < objectives;
        const int oe = obj ? OBJ_FROB : OBJ_NONE, ok = obj ? OBJ_KL : OBJ_NONE;
        hipError_t e;
        EACH_VARIANT(launch_group_pass_h(b->d_jobs, tab(P_PH_EUC, v), 0, oe, mb, tw, pg, st))
        EACH_VARIANT(launch_group_pass_h(b->d_jobs, tab(P_PH_KL, v), 1, ok, mb, tw, pg, st))
        if (obj && (e = launch_group_reduce_obj(b->d_jobs, T + T_RED * stride, it, kRedGrid, st))) return e;
        if ((e = launch_group_update_h(b->d_jobs, T + T_UH * stride, ug, st))) return e;
        EACH_VARIANT(launch_group_pass_w(b->d_jobs, tab(P_PW_EUC, v), 0, mb, tw, pg, st))
        EACH_VARIANT(launch_group_pass_w(b->d_jobs, tab(P_PW_KL, v), 1, mb, tw, pg, st))
        if ((e = launch_group_update_w(b->d_jobs, T + T_UW * stride, ug, st))) return e;
    }
#undef EACH_VARIANT
    return hipSuccess;
}

// Plans slot sl with its bucket's tile shape c in its own workspace and zeroes the arrival
// counters at the plan's offset (stream-ordered, outside graph capture).
static hipError_t plan_slot(gpumlib_nmf_batch* b, Slot& sl, const BucketCfg& c, hipStream_t st) {
    Plan pl;
    if (!make_plan(sl.rule, GPUMLIB_NMF_FP32, sl.n, sl.m, sl.r, &pl, c.target, c.tile) || pl.bytes > sl.ws_bytes)
        return hipErrorInvalidValue;
    JobArgs a;
    memset(&a, 0, sizeof(a));
    pass_args(pl, sl.V, sl.W, sl.H, sl.ws, sl.eps, 1, &a.h);
    pass_args(pl, sl.V, sl.W, sl.H, sl.ws, sl.eps, 3, &a.w);
    a.W = sl.W;
    a.H = sl.H;
    a.obj = sl.args.obj;
    a.rule = sl.rule;
    a.nobj = pl.tiles1 * pl.S1;
    sl.tile = c.tile;
    if (b->engine == ENGINE_FUSED) {
        const FusedLayout L = fused_layout(sl.n, sl.m, sl.r, c.tile);
        if (L.bytes > sl.ws_bytes) return hipErrorInvalidValue;
        char* base = (char*)sl.ws;
        for (PassArgs* p : {&a.h, &a.w}) {
            p->S = 1;
            p->P = nullptr;
            p->Psum = nullptr;
            p->gcount = nullptr;
        }
        a.h.tiles = (sl.m + c.tile - 1) / c.tile;
        a.h.kslice = round_up(sl.n, KC);
        a.h.G = (float*)base;  // W^T W, in gsl slices over n
        a.h.gsl = L.gsl_w;
        a.h.gks = L.gks_w;
        a.h.objp = (double*)(base + L.off_objp);
        a.w.tiles = (sl.n + c.tile - 1) / c.tile;
        a.w.kslice = round_up(sl.m, KC);
        a.w.G = (float*)(base + L.off_gh);  // H H^T, in gsl slices over m
        a.w.gsl = L.gsl_h;
        a.w.gks = L.gks_h;
        a.w.objp = nullptr;
        a.nobj = a.h.tiles;
        sl.args = a;
        b->dirty = true;
        return hipSuccess;
    }
    sl.args = a;
    b->dirty = true;
    return hipMemsetAsync(a.h.gcount, 0, 2 * sizeof(int), st);
}

static hipError_t replan_all(gpumlib_nmf_batch* b, int engine, const BucketSet& bk, hipStream_t st) {
    b->engine = engine;
    for (int k = 0; k < kBuckets; k++) b->bk[k] = bk[k];
    b->dirty = true;
    for (Slot& sl : b->slots) {
        if (!sl.active) continue;
        const hipError_t e = plan_slot(b, sl, bk[sl.bucket], st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

static void drop_graphs(gpumlib_nmf_batch* b) {
    for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.second);
    b->graphs.clear();
}

// Order-independent signature of the active mix (FNV-1a over the sorted shapes).
static unsigned long long mix_signature(const gpumlib_nmf_batch* b) {
    std::vector<unsigned long long> keys;
    for (const Slot& sl : b->slots)
        if (sl.active)
            keys.push_back(((unsigned long long)sl.rule << 62) ^ ((unsigned long long)sl.n << 40) ^
                           ((unsigned long long)sl.m << 16) ^ (unsigned long long)sl.r);
    std::sort(keys.begin(), keys.end());
    unsigned long long h = 1469598103934665603ull;
    for (unsigned long long k : keys)
        for (int i = 0; i < 8; i++) {
            h ^= (k >> (8 * i)) & 0xff;
            h *= 1099511628211ull;
        }
    return h ^ ((unsigned long long)b->size_target << 1);
}

// A launch sequence captured for one engine stays valid for that engine only.

// Milliseconds per iteration of the current configuration: one untimed iteration,
// then `iterations` timed with events on the stream.
static hipError_t time_config(gpumlib_nmf_batch* b, int iterations, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
                              double* ms) {
    hipError_t e = upload(b, st);
    if (e == hipSuccess) e = enqueue(b, 1, 0, st);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    if (e == hipSuccess) e = enqueue(b, iterations, 0, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    *ms = (double)t / iterations;
    return e;
}

extern "C" {

hipError_t gpumlib_nmf_batch_create(int max_jobs, gpumlib_nmf_batch_handle* out) {
    if (!out || max_jobs < 1 || max_jobs > kGroupMaxJobs) return hipErrorInvalidValue;
    *out = nullptr;
    gpumlib_nmf_batch* b = new (std::nothrow) gpumlib_nmf_batch;
    if (!b) return hipErrorOutOfMemory;
    b->maxj = max_jobs;
    b->slots.resize(max_jobs);
    hipError_t e = hipMalloc((void**)&b->d_jobs, sizeof(JobArgs) * (size_t)max_jobs);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_tables, sizeof(int) * (size_t)table_stride() * T_COUNT);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_obj, sizeof(double) * 2 * (size_t)max_jobs);
    if (e == hipSuccess) e = hipMemset(b->d_obj, 0, sizeof(double) * 2 * (size_t)max_jobs);  // unused slots read 0
    if (e == hipSuccess) e = hipHostMalloc((void**)&b->h_jobs, sizeof(JobArgs) * (size_t)max_jobs, hipHostMallocDefault);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&b->h_tables, sizeof(int) * (size_t)table_stride() * T_COUNT, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->staged, hipEventDisableTiming);
    if (e != hipSuccess) {
        gpumlib_nmf_batch_destroy(b);
        return e;
    }
    memset(b->h_jobs, 0, sizeof(JobArgs) * (size_t)max_jobs);
    *out = b;
    return hipSuccess;
}

void gpumlib_nmf_batch_destroy(gpumlib_nmf_batch_handle b) {
    if (!b) return;
    wait_staging(b);
    drop_graphs(b);
    if (b->staged) (void)hipEventDestroy(b->staged);
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->d_tables) (void)hipFree(b->d_tables);
    if (b->d_obj) (void)hipFree(b->d_obj);
    if (b->h_jobs) (void)hipHostFree(b->h_jobs);
    if (b->h_tables) (void)hipHostFree(b->h_tables);
    delete b;
}

hipError_t gpumlib_nmf_batch_set_target(gpumlib_nmf_batch_handle b, int workgroups_per_pass) {
    if (!b || workgroups_per_pass < 1) return hipErrorInvalidValue;
    b->size_target = workgroups_per_pass;
    for (int k = 0; k < kBuckets; k++) b->bk[k].target = workgroups_per_pass;
    return hipSuccess;
}

hipError_t gpumlib_nmf_batch_set_engine(gpumlib_nmf_batch_handle b, int engine) {
    if (!b || (engine != GPUMLIB_NMF_BATCH_SPLIT_K && engine != GPUMLIB_NMF_BATCH_FUSED)) return hipErrorInvalidValue;
    if (engine == b->engine) return hipSuccess;
    wait_staging(b);
    drop_graphs(b);
    hipError_t e = replan_all(b, engine, b->bk, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    return e;
}

int gpumlib_nmf_batch_engine(gpumlib_nmf_batch_handle b) { return b ? b->engine : -1; }

hipError_t gpumlib_nmf_batch_set_rank_split(gpumlib_nmf_batch_handle b, int on) {
    if (!b || (on != 0 && on != 1)) return hipErrorInvalidValue;
    bool same = true;
    for (int k = 0; k < kBuckets; k++) same = same && b->bk[k].split == on;
    if (same) return hipSuccess;
    wait_staging(b);
    drop_graphs(b);
    for (int k = 0; k < kBuckets; k++) b->bk[k].split = on;
    b->dirty = true;
    return hipSuccess;
}

int gpumlib_nmf_batch_rank_split(gpumlib_nmf_batch_handle b) {
    if (!b) return -1;
    int any = 0;
    for (int k = 0; k < kBuckets; k++) any |= b->bk[k].split;
    return any;
}

hipError_t gpumlib_nmf_batch_set_tile_cols(gpumlib_nmf_batch_handle b, int cols) {
    if (!b || (cols != 64 && cols != TILE)) return hipErrorInvalidValue;
    bool same = true;
    for (int k = 0; k < kBuckets; k++) same = same && b->bk[k].tile == cols;
    if (same) return hipSuccess;
    wait_staging(b);
    drop_graphs(b);
    BucketSet bk;
    for (int k = 0; k < kBuckets; k++) {
        bk[k] = b->bk[k];
        bk[k].tile = cols;
    }
    hipError_t e = replan_all(b, b->engine, bk, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    return e;
}

size_t gpumlib_nmf_batch_workspace_bytes(gpumlib_nmf_batch_handle b, int rule, int n, int m, int r) {
    if (!b || r > 64 || (rule != GPUMLIB_NMF_EUCLIDEAN && rule != GPUMLIB_NMF_KL)) return 0;
    // the largest plan of the targets up to the sizing target, at either tile width
    size_t most = 0;
    for (int tile : {TILE, 64}) {
        Plan pl;
        if (!make_plan(rule, GPUMLIB_NMF_FP32, n, m, r, &pl, std::max(kBatchTargetWorkgroups, b->size_target), tile))
            return 0;
        most = std::max(most, std::max(pl.bytes, fused_bytes(n, m, r, tile)));
    }
    return most;
}

hipError_t gpumlib_nmf_batch_add(gpumlib_nmf_batch_handle b, int rule, int n, int m, int r, const float* V, float* W,
                                 float* H, float eps, void* workspace, size_t workspace_bytes, int* slot) {
    Plan pl;
    if (!b || !slot || !V || !W || !H || !workspace || r > 64) return hipErrorInvalidValue;
    if (rule != GPUMLIB_NMF_EUCLIDEAN && rule != GPUMLIB_NMF_KL) return hipErrorInvalidValue;  // multiplicative only
    const int bucket = shape_bucket(n, m);
    if (!make_plan(rule, GPUMLIB_NMF_FP32, n, m, r, &pl, b->bk[bucket].target, b->bk[bucket].tile) ||
        workspace_bytes < pl.bytes)
        return hipErrorInvalidValue;
    int s = 0;
    while (s < b->maxj && b->slots[s].active) s++;
    if (s == b->maxj) return hipErrorOutOfMemory;
    wait_staging(b);
    Slot& sl = b->slots[s];
    sl.rule = rule;
    sl.n = n;
    sl.m = m;
    sl.r = r;
    sl.bucket = bucket;
    sl.V = V;
    sl.W = W;
    sl.H = H;
    sl.ws = workspace;
    sl.ws_bytes = workspace_bytes;
    sl.eps = eps;
    sl.args.obj = b->d_obj + 2 * (size_t)s;
    // split-K: the gram arrival counters start at zero (null stream, outside any captured graph)
    hipError_t e = plan_slot(b, sl, b->bk[bucket], nullptr);
    if (e == hipSuccess && b->engine == ENGINE_SPLIT) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return e;
    sl.active = true;
    *slot = s;
    return hipSuccess;
}

hipError_t gpumlib_nmf_batch_remove(gpumlib_nmf_batch_handle b, int slot) {
    if (!b || slot < 0 || slot >= b->maxj || !b->slots[slot].active) return hipErrorInvalidValue;
    b->slots[slot].active = false;
    b->dirty = true;
    return hipSuccess;
}

int gpumlib_nmf_batch_active(gpumlib_nmf_batch_handle b) {
    if (!b) return 0;
    int k = 0;
    for (const Slot& s : b->slots) k += s.active ? 1 : 0;
    return k;
}

hipError_t gpumlib_nmf_batch_iterate(gpumlib_nmf_batch_handle b, int iterations, int objectives, hipStream_t stream) {
    if (!b || iterations < 0 || objectives < 0 || objectives > 2 || objectives > iterations)
        return hipErrorInvalidValue;
    hipError_t e = upload(b, stream);
    if (e != hipSuccess) return e;
    return enqueue(b, iterations, objectives, stream);
}

hipError_t gpumlib_nmf_batch_graph_launch(gpumlib_nmf_batch_handle b, int iterations, int objectives,
                                          hipStream_t stream) {
    if (!b || !stream || iterations < 1 || objectives < 0 || objectives > 2 || objectives > iterations)
        return hipErrorInvalidValue;
    hipError_t e = upload(b, stream);
    if (e != hipSuccess) return e;
    const auto key = std::make_pair(iterations, objectives);
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        hipGraph_t g = nullptr;
        if ((e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal)) != hipSuccess) return e;
        const hipError_t ee = enqueue(b, iterations, objectives, stream);
        e = hipStreamEndCapture(stream, &g);
        if (ee != hipSuccess) e = ee;
        hipGraphExec_t exec = nullptr;
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) return e;
        it = b->graphs.emplace(key, exec).first;
    }
    return hipGraphLaunch(it->second, stream);
}

hipError_t gpumlib_nmf_batch_objectives(gpumlib_nmf_batch_handle b, double* host_out, hipStream_t stream) {
    if (!b || !host_out) return hipErrorInvalidValue;
    hipError_t e =
        hipMemcpyAsync(host_out, b->d_obj, sizeof(double) * 2 * (size_t)b->maxj, hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(stream);
}

const double* gpumlib_nmf_batch_objective_device_ptr(gpumlib_nmf_batch_handle b) { return b ? b->d_obj : nullptr; }

}  // extern "C"

extern "C" hipError_t gpumlib_nmf_batch_autotune(gpumlib_nmf_batch_handle b, int trial_iterations, hipStream_t stream,
                                                 gpumlib_nmf_batch_tuning* out) {
    if (!b || !stream || trial_iterations < 1) return hipErrorInvalidValue;
    if (gpumlib_nmf_batch_active(b) == 0) return hipErrorInvalidValue;
    // buckets with active problems (and with problems of r <= 32, the ones rank blocking moves);
    // the single target reported is that of the largest bucket present
    bool present[kBuckets] = {false, false, false}, narrow[kBuckets] = {false, false, false};
    for (const Slot& sl : b->slots)
        if (sl.active) {
            present[sl.bucket] = true;
            narrow[sl.bucket] = narrow[sl.bucket] || sl.r <= 32;
        }
    int top = 0;
    for (int k = 0; k < kBuckets; k++)
        if (present[k]) top = k;
    auto report = [&](const TuneConfig& c, int cached) {
        if (!out) return;
        out->engine = c.engine;
        out->target_workgroups = c.bk[top].target;
        out->pass_grid = c.pass_grid;
        out->update_grid = c.upd_grid;
        out->ms_per_iteration = c.ms_per_iteration;
        out->ms_per_iteration_before = c.ms_default;
        out->trials = cached ? 0 : c.trials;
        out->cached = cached;
        out->split_ranks = 0;
        for (int k = 0; k < kBuckets; k++) {
            out->bucket_targets[k] = present[k] ? c.bk[k].target : 0;
            out->bucket_tile_cols[k] = present[k] ? c.bk[k].tile : 0;
            out->bucket_split_ranks[k] = present[k] ? c.bk[k].split : 0;
            out->split_ranks |= out->bucket_split_ranks[k];
        }
    };
    BucketSet bk0;
    for (int k = 0; k < kBuckets; k++) bk0[k] = b->bk[k];
    const int en0 = b->engine, pg0 = b->pass_grid, ug0 = b->upd_grid;
    const unsigned vs0 = variants_launched(b);
    auto changed = [&]() {
        return b->engine != en0 || b->pass_grid != pg0 || b->upd_grid != ug0 || variants_launched(b) != vs0;
    };
    const unsigned long long sig = mix_signature(b);
    auto hit = b->tuned.find(sig);
    if (hit != b->tuned.end()) {  // this mix was tuned before: apply its winner
        const TuneConfig& c = hit->second;
        hipError_t e = replan_all(b, c.engine, c.bk, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return e;
        b->pass_grid = c.pass_grid;
        b->upd_grid = c.upd_grid;
        if (changed()) drop_graphs(b);
        report(c, 1);
        return hipSuccess;
    }
    // every problem of a bucket (all buckets: bucket < 0) must hold the plan of a candidate shape
    auto fits = [&](int target, int tile, int bucket) {
        for (const Slot& sl : b->slots) {
            if (!sl.active || (bucket >= 0 && sl.bucket != bucket)) continue;
            Plan pl;
            if (!make_plan(sl.rule, GPUMLIB_NMF_FP32, sl.n, sl.m, sl.r, &pl, target, tile) || pl.bytes > sl.ws_bytes ||
                fused_bytes(sl.n, sl.m, sl.r, tile) > sl.ws_bytes)
                return false;
        }
        return true;
    };
    // W and H of every active problem are saved, and restored after each trial
    size_t floats = 0;
    for (const Slot& sl : b->slots)
        if (sl.active) floats += (size_t)sl.n * sl.r + (size_t)sl.r * sl.m;
    float* save = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc((void**)&save, floats * sizeof(float));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto copy_wh = [&](bool to_save) {
        size_t off = 0;
        hipError_t ce = hipSuccess;
        for (const Slot& sl : b->slots) {
            if (!sl.active || ce != hipSuccess) continue;
            const size_t nw = (size_t)sl.n * sl.r, nh = (size_t)sl.r * sl.m;
            float* sw = save + off;
            float* sh = save + off + nw;
            if (to_save) {
                ce = hipMemcpyAsync(sw, sl.W, nw * sizeof(float), hipMemcpyDeviceToDevice, stream);
                if (ce == hipSuccess) ce = hipMemcpyAsync(sh, sl.H, nh * sizeof(float), hipMemcpyDeviceToDevice, stream);
            } else {
                ce = hipMemcpyAsync(sl.W, sw, nw * sizeof(float), hipMemcpyDeviceToDevice, stream);
                if (ce == hipSuccess) ce = hipMemcpyAsync(sl.H, sh, nh * sizeof(float), hipMemcpyDeviceToDevice, stream);
            }
            off += nw + nh;
        }
        return ce;
    };
    TuneConfig best;
    best.engine = en0;
    best.pass_grid = pg0;
    best.upd_grid = ug0;
    for (int k = 0; k < kBuckets; k++) best.bk[k] = bk0[k];
    best.ms_per_iteration = 1e30;
    best.ms_default = 0.0;
    best.trials = 0;
    auto trial = [&](int engine, const BucketSet& bk, int pg, int ug) -> hipError_t {
        hipError_t te = replan_all(b, engine, bk, stream);
        if (te != hipSuccess) return te;
        b->pass_grid = pg;
        b->upd_grid = ug;
        double ms = 0.0;
        te = time_config(b, trial_iterations, stream, e0, e1, &ms);
        if (te == hipSuccess) te = copy_wh(false);
        if (te != hipSuccess) return te;
        if (++best.trials == 1) best.ms_default = ms;
        if (ms < best.ms_per_iteration) {
            best.engine = engine;
            for (int k = 0; k < kBuckets; k++) best.bk[k] = bk[k];
            best.pass_grid = pg;
            best.upd_grid = ug;
            best.ms_per_iteration = ms;
        }
        return hipSuccess;
    };
    // a copy of the best bucket shapes with bucket k's replaced
    auto with = [&](int k, const BucketCfg& c, BucketSet& bk) {
        for (int j = 0; j < kBuckets; j++) bk[j] = best.bk[j];
        bk[k] = c;
    };
    // The configuration in effect first.  Then the tile depth along K: the fused engine (one
    // slice, the whole contraction per tile) with every bucket on 128- and on 64-wide tiles, and
    // the split-K targets, one for every bucket, on 128-wide tiles; then, split-K, each present
    // bucket's target one step deeper and shallower than the best (the other buckets kept).  Then,
    // per present bucket, the other tile width; then the pass grids and the update (fused: gram)
    // grids around the best so far; then, per bucket with problems of r <= 32, the other rank
    // blocking.
    if (e == hipSuccess) e = copy_wh(true);
    // the shapes in effect with every bucket on one tile width (and the split-K target of each)
    auto uniform = [&](int tile, int target, BucketSet& bk) {
        bool same = true;
        for (int k = 0; k < kBuckets; k++) {
            bk[k] = bk0[k];
            bk[k].tile = tile;
            if (target > 0) bk[k].target = target;
            same = same && bk[k] == bk0[k];
        }
        return same;
    };
    if (e == hipSuccess) e = trial(en0, bk0, pg0, ug0);
    for (int tile : {TILE, 64}) {  // the fused engine at either uniform width
        BucketSet bk;
        const bool same = uniform(tile, 0, bk) && en0 == ENGINE_FUSED;
        if (e == hipSuccess && !same && fits(kBatchTargetWorkgroups, tile, -1)) e = trial(ENGINE_FUSED, bk, pg0, ug0);
    }
    for (int t : kTuneTargets) {  // split-K targets on 128-wide tiles
        BucketSet bk;
        const bool same = uniform(TILE, t, bk) && en0 == ENGINE_SPLIT;
        if (e == hipSuccess && !same && fits(t, TILE, -1)) e = trial(ENGINE_SPLIT, bk, pg0, ug0);
    }
    const int nt = (int)(sizeof(kTuneTargets) / sizeof(kTuneTargets[0]));
    if (best.engine == ENGINE_SPLIT) {
        for (int k = kBuckets - 1; k >= 0; k--) {
            if (!present[k]) continue;
            int cur = 0;
            while (cur < nt && kTuneTargets[cur] != best.bk[k].target) cur++;
            for (int d : {-1, 1}) {
                const int c = cur + d;
                if (e != hipSuccess || cur == nt || c < 0 || c >= nt || !fits(kTuneTargets[c], best.bk[k].tile, k))
                    continue;
                BucketCfg nc = best.bk[k];
                nc.target = kTuneTargets[c];
                BucketSet bk;
                with(k, nc, bk);
                e = trial(ENGINE_SPLIT, bk, pg0, ug0);
            }
        }
    }
    for (int k = kBuckets - 1; k >= 0; k--) {
        BucketCfg nc = best.bk[k];
        nc.tile = nc.tile == TILE ? 64 : TILE;
        if (e != hipSuccess || !present[k] || !fits(nc.target, nc.tile, k)) continue;
        BucketSet bk;
        with(k, nc, bk);
        e = trial(best.engine, bk, pg0, ug0);
    }
    {
        const int enb = best.engine;
        BucketSet bkb;
        for (int k = 0; k < kBuckets; k++) bkb[k] = best.bk[k];
        for (int pg : kTunePassGrids)
            if (e == hipSuccess && pg != pg0) e = trial(enb, bkb, pg, ug0);
        const int pgb = best.pass_grid;
        for (int ug : kTuneUpdGrids)
            if (e == hipSuccess && ug != ug0) e = trial(enb, bkb, pgb, ug);
    }
    for (int k = kBuckets - 1; k >= 0; k--) {
        BucketCfg nc = best.bk[k];
        nc.split = !nc.split;
        if (e != hipSuccess || !narrow[k]) continue;
        BucketSet bk;
        with(k, nc, bk);
        e = trial(best.engine, bk, best.pass_grid, best.upd_grid);
    }
    // apply the winner, or the configuration in effect before if a trial failed
    const bool ok = (e == hipSuccess);
    hipError_t ae = ok ? replan_all(b, best.engine, best.bk, stream) : replan_all(b, en0, bk0, stream);
    b->pass_grid = ok ? best.pass_grid : pg0;
    b->upd_grid = ok ? best.upd_grid : ug0;
    if (changed()) drop_graphs(b);
    if (ae == hipSuccess) ae = hipStreamSynchronize(stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (save) (void)hipFree(save);
    if (!ok) return e;
    if (ae != hipSuccess) return ae;
    b->tuned[sig] = best;
    report(best, 0);
    return hipSuccess;
}
