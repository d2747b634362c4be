// Batched execution of many independent NMF problems with grouped launches.
//
// The reference runs independent problems one after another (the block "groups"
// of FacesNMFDlg.cpp:582-656, each a separate NMF object and a separate launch
// sequence).  Here every registered problem advances with the same six grouped
// launches per iteration (H pass per rule, H update, W pass per rule, W update);
// the per-problem arguments live in a device table that the host refreshes when
// the set of problems changes.  Because the launches read everything from that
// table, a hipGraph captured once stays valid across changes of the set.
#include <hip/hip_runtime.h>
#include <string.h>

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

enum { T_PH_EUC = 0, T_PH_KL, T_PW_EUC, T_PW_KL, T_UH, T_UW, T_RED, T_COUNT };
constexpr int kPassGrid = 1024;  // persistent workgroups of the grouped passes
constexpr int kUpdGrid = 512;
constexpr int kRedGrid = 256;
constexpr int kBatchTargetWorkgroups = 32;  // per problem and pass, before tuning

struct Slot {
    bool active = false;
    int rule = 0, n = 0, m = 0, r = 0;
    JobArgs args;
};

}  // namespace

struct gpumlib_nmf_batch {
    int maxj = 0;
    int target = kBatchTargetWorkgroups;
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

// Rebuilds the launch tables from the active slots and stages them to the device.
static hipError_t upload(gpumlib_nmf_batch* b, hipStream_t st) {
    if (!b->dirty) return hipSuccess;
    wait_staging(b);
    const int stride = table_stride();
    memset(b->h_tables, 0, sizeof(int) * (size_t)stride * T_COUNT);
    for (int k = 0; k < T_COUNT; k++) {
        int* t = b->h_tables + (size_t)k * stride;
        int J = 0, T = 0;
        int* first = t + 2;
        int* idx = t + 3 + kGroupMaxJobs;
        for (int s = 0; s < b->maxj; s++) {
            const Slot& sl = b->slots[s];
            if (!sl.active) continue;
            int items = 0;
            switch (k) {
                case T_PH_EUC: items = sl.rule == GPUMLIB_NMF_EUCLIDEAN ? sl.args.h.tiles * sl.args.h.S : 0; break;
                case T_PH_KL: items = sl.rule == GPUMLIB_NMF_KL ? sl.args.h.tiles * sl.args.h.S : 0; break;
                case T_PW_EUC: items = sl.rule == GPUMLIB_NMF_EUCLIDEAN ? sl.args.w.tiles * sl.args.w.S : 0; break;
                case T_PW_KL: items = sl.rule == GPUMLIB_NMF_KL ? sl.args.w.tiles * sl.args.w.S : 0; break;
                case T_UH: items = (sl.m + UPD_COLS - 1) / UPD_COLS; break;
                case T_UW: items = (sl.n + UPD_ROWS - 1) / UPD_ROWS; break;
                case T_RED: items = 1; break;
            }
            if (items == 0) continue;
            first[J] = T;
            idx[J] = s;
            T += items;
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

static hipError_t enqueue(gpumlib_nmf_batch* b, int iterations, int objectives, hipStream_t st) {
    const int stride = table_stride();
    const int* T = b->d_tables;
    for (int it = 0; it < iterations; it++) {
        const bool obj = it < objectives;
        hipError_t e;
        if ((e = launch_group_pass_h(b->d_jobs, T + T_PH_EUC * stride, 0, obj ? OBJ_FROB : OBJ_NONE, kPassGrid, st)))
            return e;
        if ((e = launch_group_pass_h(b->d_jobs, T + T_PH_KL * stride, 1, obj ? OBJ_KL : OBJ_NONE, kPassGrid, st)))
            return e;
        if (obj && (e = launch_group_reduce_obj(b->d_jobs, T + T_RED * stride, it, kRedGrid, st))) return e;
        if ((e = launch_group_update_h(b->d_jobs, T + T_UH * stride, kUpdGrid, st))) return e;
        if ((e = launch_group_pass_w(b->d_jobs, T + T_PW_EUC * stride, 0, kPassGrid, st))) return e;
        if ((e = launch_group_pass_w(b->d_jobs, T + T_PW_KL * stride, 1, kPassGrid, st))) return e;
        if ((e = launch_group_update_w(b->d_jobs, T + T_UW * stride, kUpdGrid, st))) return e;
    }
    return hipSuccess;
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
    for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.second);
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
    b->target = workgroups_per_pass;
    return hipSuccess;
}

size_t gpumlib_nmf_batch_workspace_bytes(gpumlib_nmf_batch_handle b, int rule, int n, int m, int r) {
    Plan pl;
    if (!b || r > 64 || (rule != GPUMLIB_NMF_EUCLIDEAN && rule != GPUMLIB_NMF_KL) ||
        !make_plan(rule, GPUMLIB_NMF_FP32, n, m, r, &pl, b->target))
        return 0;
    return pl.bytes;
}

hipError_t gpumlib_nmf_batch_add(gpumlib_nmf_batch_handle b, int rule, int n, int m, int r, const float* V, float* W,
                                 float* H, float eps, void* workspace, size_t workspace_bytes, int* slot) {
    Plan pl;
    if (!b || !slot || !V || !W || !H || !workspace || r > 64) return hipErrorInvalidValue;
    if (rule != GPUMLIB_NMF_EUCLIDEAN && rule != GPUMLIB_NMF_KL) return hipErrorInvalidValue;  // multiplicative only
    if (!make_plan(rule, GPUMLIB_NMF_FP32, n, m, r, &pl, b->target) || workspace_bytes < pl.bytes)
        return hipErrorInvalidValue;
    int s = 0;
    while (s < b->maxj && b->slots[s].active) s++;
    if (s == b->maxj) return hipErrorOutOfMemory;
    wait_staging(b);
    Slot& sl = b->slots[s];
    sl.active = true;
    sl.rule = rule;
    sl.n = n;
    sl.m = m;
    sl.r = r;
    memset(&sl.args, 0, sizeof(sl.args));
    pass_args(pl, V, W, H, workspace, eps, 1, &sl.args.h);
    pass_args(pl, V, W, H, workspace, eps, 3, &sl.args.w);
    // gram arrival counters start at zero (outside any captured graph)
    const hipError_t ze = hipMemset(sl.args.h.gcount, 0, 2 * sizeof(int));
    if (ze != hipSuccess) {
        sl.active = false;
        return ze;
    }
    sl.args.W = W;
    sl.args.H = H;
    sl.args.obj = b->d_obj + 2 * (size_t)s;
    sl.args.rule = rule;
    sl.args.nobj = pl.tiles1 * pl.S1;
    b->dirty = true;
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
