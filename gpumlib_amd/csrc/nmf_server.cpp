// Continuous-batching NMF server (no reference counterpart).
//
// Requests (one NMF problem each: rule, V, initial W and H, eps, tolerance,
// iteration cap) are queued by any host thread.  One engine thread owns the
// device work:
//   * admission: a queued problem is admitted when a batch slot is free and its
//     HBM footprint (V, W, H and its workspace) fits in the device pool; the
//     queue is scanned in arrival order, skipping problems that do not fit yet;
//   * rounds: every admitted problem advances `round_iterations` iterations with
//     the grouped launches of the batch engine (replayed from one hipGraph),
//     the FP64 objective being computed on entry to the first two iterations;
//   * retirement: a problem whose relative objective change over one iteration
//     is below its tolerance, or that reached its iteration cap, has W and H
//     copied back and its memory returned to the pool; its slot is reused in
//     the next round.
// Problems admitted while others are in flight join the next round, so work of
// different sizes and ages overlaps on the device.
//   * autotuning (config.autotune > 0): when the number of resident problems has
//     doubled or halved since the last tuning (and is at least 4), the batch
//     autotuner picks the K-split depth and grid sizes for the current mix before
//     the round (trials restore W and H, so no problem's state changes).
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <set>
#include <thread>
#include <vector>

#include "../../include/gpumlib_nmf.h"

namespace {

using Clock = std::chrono::steady_clock;

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// First-fit sub-allocator over one device allocation (offsets only).
class Pool {
  public:
    void init(size_t bytes) {
        total_ = bytes;
        free_.clear();
        free_[0] = bytes;
        used_ = high_ = 0;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = align256(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes, start = it->first + bytes;
            free_.erase(it);
            if (rest) free_[start] = rest;
            used_ += bytes;
            high_ = std::max(high_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align256(bytes);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t total() const { return total_; }
    size_t high_water() const { return high_; }

  private:
    std::map<size_t, size_t> free_;
    size_t total_ = 0, used_ = 0, high_ = 0;
};

struct Job {
    long long id = 0;
    int rule = 0, n = 0, m = 0, r = 0;
    float eps = 0;
    double tol = 0;
    int max_iters = 0;
    std::vector<float> V, W, H;  // inputs, then W and H results
    // device placement
    size_t off = 0, bytes = 0, offV = 0, offW = 0, offH = 0, offWs = 0, ws_bytes = 0;
    int slot = -1;
    int iters = 0;
    int converged = 0;
    double objective = 0;
    double t_submit = 0, t_admit = 0, t_finish = 0;
    hipError_t status = hipSuccess;
    bool done = false;
};

}  // namespace

struct gpumlib_nmf_server;

namespace {
// Servers still alive when the library is unloaded, i.e. the process exits without
// destroying them (an exception in the caller, say).  Their engine threads must not
// keep launching work while the HIP runtime is being torn down (in-flight kernels would
// then touch memory that is going away), so the library's exit path stops each engine
// after its round in flight and waits for its stream.  The registry is created after
// the HIP runtime has been initialised (on the first server), so it is destroyed
// before the runtime.
struct LiveServers {
    std::mutex mu;
    std::set<gpumlib_nmf_server*> set;
    ~LiveServers();
};
LiveServers& live_servers() {
    static LiveServers l;
    return l;
}
}  // namespace

struct gpumlib_nmf_server {
    gpumlib_nmf_server_config cfg;
    int device = 0;
    Clock::time_point t0;
    Pool pool;
    char* d_pool = nullptr;
    gpumlib_nmf_batch_handle batch = nullptr;
    hipStream_t stream = nullptr;
    double* h_obj = nullptr;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Job>> pending;
    std::map<long long, std::shared_ptr<Job>> jobs;  // submitted and not yet collected
    std::vector<std::shared_ptr<Job>> active;        // by slot
    long long next_id = 1;
    bool stop = false;
    bool abandon = false;  // process exit: leave after the round in flight, do not finish queued work
    hipError_t fatal = hipSuccess;
    gpumlib_nmf_server_stats stats;
    int tuned_at = 0;  // resident problems at the last autotuning (0: never tuned)
    std::thread engine;

    double* d_obj = nullptr;  // one-at-a-time mode: the round's two objectives
    // continuous batching: uploads of the next round's problems and copies of finished ones
    // run on their own stream while a round runs on `stream`
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_staged = nullptr, ev_retired = nullptr, ev_ready = nullptr;
    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    void run();
    void run_batched();
    hipError_t run_one(Job& j);
    bool admit_locked(std::vector<std::shared_ptr<Job>>& admitted, int reserved = 0);
    void fail_all_locked(hipError_t e, std::vector<std::shared_ptr<Job>>& extra);
};

// Device bytes of a problem: V, W, H and its workspace (the batch's plan, or in the
// one-at-a-time mode the single-problem iteration's).
static size_t job_footprint(gpumlib_nmf_batch_handle b, bool sequential, Job& j) {
    j.ws_bytes = sequential ? gpumlib_nmf_workspace_bytes(j.rule, GPUMLIB_NMF_FP32, j.n, j.m, j.r)
                            : gpumlib_nmf_batch_workspace_bytes(b, j.rule, j.n, j.m, j.r);
    if (j.ws_bytes == 0) return 0;
    j.offV = 0;
    j.offW = align256(sizeof(float) * (size_t)j.n * j.m);
    j.offH = j.offW + align256(sizeof(float) * (size_t)j.n * j.r);
    j.offWs = j.offH + align256(sizeof(float) * (size_t)j.r * j.m);
    return j.offWs + align256(j.ws_bytes);
}

// reserved: slots already promised to problems admitted but not yet in the batch.
bool gpumlib_nmf_server::admit_locked(std::vector<std::shared_ptr<Job>>& admitted, int reserved) {
    int free_slots =
        cfg.max_active - reserved - (int)std::count_if(active.begin(), active.end(), [](auto& p) { return !!p; });
    bool any = false;
    for (auto it = pending.begin(); it != pending.end() && free_slots > 0;) {
        Job& j = **it;
        size_t off = 0;
        if (pool.alloc(j.bytes, &off)) {
            j.off = off;
            admitted.push_back(*it);
            it = pending.erase(it);
            free_slots--;
            any = true;
        } else {
            ++it;
        }
    }
    return any;
}

// One-at-a-time mode: the admitted problem runs alone, round after round, through the
// single-problem fused iteration (gpumlib_nmf_iterate) on its pool memory, with the
// stopping rule of the batched rounds (objectives on entry to the round's first two
// iterations, read back once per round).
hipError_t gpumlib_nmf_server::run_one(Job& j) {
    char* base = d_pool + j.off;
    float* dV = (float*)(base + j.offV);
    float* dW = (float*)(base + j.offW);
    float* dH = (float*)(base + j.offH);
    void* ws = base + j.offWs;
    const int R = cfg.round_iterations;
    const int kind = j.rule == GPUMLIB_NMF_KL ? GPUMLIB_NMF_OBJ_KL : GPUMLIB_NMF_OBJ_FROBENIUS;
    hipError_t e = hipMemcpyAsync(dV, j.V.data(), sizeof(float) * j.V.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dW, j.W.data(), sizeof(float) * j.W.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dH, j.H.data(), sizeof(float) * j.H.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    j.t_admit = now();
    std::vector<float>().swap(j.V);
    while (e == hipSuccess) {
        e = gpumlib_nmf_iterate(j.rule, GPUMLIB_NMF_FP32, j.n, j.m, j.r, dV, dW, dH, j.eps, 1, 1, kind, d_obj, ws,
                                j.ws_bytes, stream);
        if (e == hipSuccess && R >= 2)
            e = gpumlib_nmf_iterate(j.rule, GPUMLIB_NMF_FP32, j.n, j.m, j.r, dV, dW, dH, j.eps, 1, 1, kind, d_obj + 1,
                                    ws, j.ws_bytes, stream);
        if (e == hipSuccess && R > 2)
            e = gpumlib_nmf_iterate(j.rule, GPUMLIB_NMF_FP32, j.n, j.m, j.r, dV, dW, dH, j.eps, R - 2, 1,
                                    GPUMLIB_NMF_OBJ_NONE, nullptr, ws, j.ws_bytes, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h_obj, d_obj, 2 * sizeof(double), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) break;
        j.iters += R;
        const double f0 = h_obj[0], f1 = R >= 2 ? h_obj[1] : f0;
        j.objective = f1;
        const double rel = (f0 != 0) ? std::fabs(f0 - f1) / std::fabs(f0) : 0.0;
        if (R >= 2 && rel < j.tol) j.converged = 1;
        {
            std::lock_guard<std::mutex> lk(mu);
            stats.rounds++;
            stats.iterations_batched += R;
            stats.peak_active = std::max(stats.peak_active, 1);
        }
        if (j.converged || j.iters >= j.max_iters) break;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(j.W.data(), dW, sizeof(float) * j.W.size(), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(j.H.data(), dH, sizeof(float) * j.H.size(), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    return e;
}

void gpumlib_nmf_server::run() {
    (void)hipSetDevice(device);
    const int rounds_iters = cfg.round_iterations;
    if (cfg.sequential) {
        for (;;) {
            std::shared_ptr<Job> jp;
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    if (abandon) return;
                    std::vector<std::shared_ptr<Job>> admitted;
                    admit_locked(admitted);  // max_active is 1 in this mode
                    if (!admitted.empty()) {
                        jp = admitted[0];
                        break;
                    }
                    if (stop) return;
                    cv_work.wait(lk);
                }
            }
            const hipError_t e = run_one(*jp);
            {
                std::lock_guard<std::mutex> lk(mu);
                pool.release(jp->off, jp->bytes);
                jp->status = e;
                jp->t_finish = now();
                jp->done = true;
                stats.completed++;
                stats.pool_high_water = pool.high_water();
                if (e != hipSuccess) {
                    fatal = e;
                    for (auto& q : pending) {
                        q->status = e;
                        q->done = true;
                    }
                    pending.clear();
                }
            }
            cv_done.notify_all();
            if (e != hipSuccess) return;
        }
    }
    run_batched();
}

// Marks every unfinished problem (queued, staged or active, plus `extra`) as failed with e.
void gpumlib_nmf_server::fail_all_locked(hipError_t e, std::vector<std::shared_ptr<Job>>& extra) {
    fatal = e;
    auto fail = [&](const std::shared_ptr<Job>& jp) {
        if (!jp || jp->done) return;
        jp->status = e;
        jp->done = true;
    };
    for (auto& jp : pending) fail(jp);
    pending.clear();
    for (auto& jp : active) fail(jp);
    for (auto& jp : extra) fail(jp);
}

// Continuous batching, pipelined: while round k runs on `stream`, the engine thread
// completes the retirement of the problems that finished after round k - 1 (their W and H
// were copied back on the copy stream), releases their memory, admits queued problems
// into it and uploads them on the copy stream; at the round boundary only the objectives
// are read, the finished problems leave the batch (their copies queued) and the uploaded
// ones join it (the round stream waits for their upload, long done).  Problems admitted at
// the boundary itself are uploaded then and join at the next boundary, unless the batch
// would otherwise have nothing to run.
void gpumlib_nmf_server::run_batched() {
    const int rounds_iters = cfg.round_iterations;
    std::vector<std::shared_ptr<Job>> staged;    // uploaded (or uploading) on copy_stream, not yet in the batch
    std::vector<std::shared_ptr<Job>> retiring;  // finished, W and H being copied back on copy_stream
    hipError_t e = hipSuccess;
    auto upload = [&](std::vector<std::shared_ptr<Job>>& admitted) {
        for (auto& jp : admitted) {
            Job& j = *jp;
            char* base = d_pool + j.off;
            if (e == hipSuccess)
                e = hipMemcpyAsync(base + j.offV, j.V.data(), sizeof(float) * j.V.size(), hipMemcpyHostToDevice,
                                   copy_stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(base + j.offW, j.W.data(), sizeof(float) * j.W.size(), hipMemcpyHostToDevice,
                                   copy_stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(base + j.offH, j.H.data(), sizeof(float) * j.H.size(), hipMemcpyHostToDevice,
                                   copy_stream);
            staged.push_back(jp);
        }
        if (e == hipSuccess && !admitted.empty()) e = hipEventRecord(ev_staged, copy_stream);
        admitted.clear();
    };
    auto finish_retiring = [&]() {
        if (retiring.empty()) return;
        const hipError_t ce = hipEventSynchronize(ev_retired);
        if (e == hipSuccess) e = ce;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto& jp : retiring) {
                Job& j = *jp;
                pool.release(j.off, j.bytes);
                if (j.status == hipSuccess) j.status = ce;
                j.t_finish = now();
                j.done = true;
                stats.completed++;
            }
            stats.pool_high_water = pool.high_water();
        }
        retiring.clear();
        cv_done.notify_all();
    };
    for (;;) {
        // admission: wait for work when nothing is resident, staged or retiring
        std::vector<std::shared_ptr<Job>> admitted;
        bool have_active = false;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                if (abandon) return;
                admit_locked(admitted, (int)staged.size());
                have_active = std::any_of(active.begin(), active.end(), [](auto& p) { return !!p; });
                if (!admitted.empty() || have_active || !staged.empty() || !retiring.empty()) break;
                if (stop) return;
                cv_work.wait(lk);
            }
        }
        // The problems uploaded while the last round ran join at this boundary; those admitted
        // just now are uploaded now and join at the next one, so that the round does not wait
        // for their copies (from pageable host memory) -- unless nothing would run otherwise.
        std::vector<std::shared_ptr<Job>> ready;
        ready.swap(staged);
        if (e == hipSuccess && !ready.empty()) e = hipEventRecord(ev_ready, copy_stream);
        upload(admitted);
        finish_retiring();
        if (!have_active && ready.empty()) {
            if (staged.empty()) continue;  // only retirements were pending
            ready.swap(staged);
            if (e == hipSuccess) e = hipEventRecord(ev_ready, copy_stream);
        }
        // round boundary: the ready problems join the batch
        std::vector<std::shared_ptr<Job>> joined;
        if (e == hipSuccess && !ready.empty()) e = hipStreamWaitEvent(stream, ev_ready, 0);
        for (auto& jp : ready) {
            Job& j = *jp;
            char* base = d_pool + j.off;
            int slot = -1;
            if (e == hipSuccess)
                e = gpumlib_nmf_batch_add(batch, j.rule, j.n, j.m, j.r, (float*)(base + j.offV),
                                          (float*)(base + j.offW), (float*)(base + j.offH), j.eps, base + j.offWs,
                                          j.ws_bytes, &slot);
            j.slot = slot;
            j.t_admit = now();
            if (e == hipSuccess) {
                active[slot] = jp;
                joined.push_back(jp);
            } else {
                std::lock_guard<std::mutex> lk(mu);
                pool.release(j.off, j.bytes);
                j.status = e;
                j.done = true;
            }
        }
        ready.clear();
        if (e == hipSuccess && cfg.autotune > 0) {
            const int resident = gpumlib_nmf_batch_active(batch);
            if (resident >= 4 && (tuned_at == 0 || resident >= 2 * tuned_at || 2 * resident <= tuned_at)) {
                gpumlib_nmf_batch_tuning t;
                e = gpumlib_nmf_batch_autotune(batch, cfg.autotune, stream, &t);
                if (e == hipSuccess) {
                    tuned_at = resident;
                    std::lock_guard<std::mutex> lk(mu);
                    stats.autotunes++;
                    stats.tuned_target = t.target_workgroups;
                    stats.tuned_pass_grid = t.pass_grid;
                    stats.tuned_update_grid = t.update_grid;
                    stats.tuned_engine = t.engine;
                    stats.tuned_split_ranks = t.split_ranks;
                    for (int k = 0; k < 3; k++) {
                        stats.tuned_bucket_targets[k] = t.bucket_targets[k];
                        stats.tuned_bucket_tile_cols[k] = t.bucket_tile_cols[k];
                        stats.tuned_bucket_split_ranks[k] = t.bucket_split_ranks[k];
                    }
                }
            }
        }
        // one round
        if (e == hipSuccess) {
            if (cfg.use_graph)
                e = gpumlib_nmf_batch_graph_launch(batch, rounds_iters, rounds_iters >= 2 ? 2 : 1, stream);
            else
                e = gpumlib_nmf_batch_iterate(batch, rounds_iters, rounds_iters >= 2 ? 2 : 1, stream);
        }
        // while it runs: the joined problems' host inputs are released once uploaded, and
        // queued problems are admitted and uploaded for the next round
        if (!joined.empty()) {
            const hipError_t ce = hipEventSynchronize(ev_ready);
            if (e == hipSuccess) e = ce;
            for (auto& jp : joined) std::vector<float>().swap(jp->V);
        }
        if (e == hipSuccess) {
            std::vector<std::shared_ptr<Job>> next;
            {
                std::lock_guard<std::mutex> lk(mu);
                admit_locked(next, (int)staged.size());  // (staged: admitted at the boundary, not yet in the batch)
            }
            upload(next);
        }
        if (e == hipSuccess) e = gpumlib_nmf_batch_objectives(batch, h_obj, stream);
        std::vector<std::shared_ptr<Job>> finished;
        int in_round = 0;
        for (auto& jp : active) {
            if (!jp) continue;
            in_round++;
            Job& j = *jp;
            j.iters += rounds_iters;
            const double f0 = h_obj[2 * j.slot], f1 = rounds_iters >= 2 ? h_obj[2 * j.slot + 1] : f0;
            j.objective = f1;
            const double rel = (f0 != 0) ? std::fabs(f0 - f1) / std::fabs(f0) : 0.0;
            if (rounds_iters >= 2 && rel < j.tol) j.converged = 1;
            if (e != hipSuccess || j.converged || j.iters >= j.max_iters) finished.push_back(jp);
        }
        // the finished problems leave the batch; their results come back on the copy stream
        // (the round is complete: the objectives were read after it)
        for (auto& jp : finished) {
            Job& j = *jp;
            char* base = d_pool + j.off;
            if (e == hipSuccess)
                e = hipMemcpyAsync(j.W.data(), base + j.offW, sizeof(float) * j.W.size(), hipMemcpyDeviceToHost,
                                   copy_stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(j.H.data(), base + j.offH, sizeof(float) * j.H.size(), hipMemcpyDeviceToHost,
                                   copy_stream);
            j.status = e;
            (void)gpumlib_nmf_batch_remove(batch, j.slot);
            active[j.slot].reset();
            retiring.push_back(jp);
        }
        if (e == hipSuccess && !finished.empty()) e = hipEventRecord(ev_retired, copy_stream);
        {
            std::lock_guard<std::mutex> lk(mu);
            stats.rounds++;
            stats.iterations_batched += (long long)in_round * rounds_iters;
            stats.peak_active = std::max(stats.peak_active, in_round);
        }
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(copy_stream);
            (void)hipStreamSynchronize(stream);
            std::lock_guard<std::mutex> lk(mu);
            for (auto& jp : retiring) {
                pool.release(jp->off, jp->bytes);
                jp->status = e;
                jp->t_finish = now();
                jp->done = true;
                stats.completed++;
            }
            retiring.clear();
            for (auto& jp : staged) pool.release(jp->off, jp->bytes);
            fail_all_locked(e, staged);
            staged.clear();
            cv_done.notify_all();
            return;
        }
    }
}

LiveServers::~LiveServers() {
    std::lock_guard<std::mutex> g(mu);
    for (gpumlib_nmf_server* s : set) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->abandon = true;
        }
        s->cv_work.notify_all();
        if (s->engine.joinable()) s->engine.join();
        (void)hipStreamSynchronize(s->stream);
        (void)hipStreamSynchronize(s->copy_stream);
    }
    set.clear();
}

extern "C" {

hipError_t gpumlib_nmf_server_create(const gpumlib_nmf_server_config* cfg, gpumlib_nmf_server_handle* out) {
    if (!cfg || !out || cfg->hbm_budget_bytes == 0 || cfg->max_active < 1 || cfg->max_active > 1024 ||
        cfg->round_iterations < 1 || cfg->split_k < 0 || cfg->split_k > 1 || cfg->sequential < 0 ||
        cfg->sequential > 1)
        return hipErrorInvalidValue;
    *out = nullptr;
    gpumlib_nmf_server* s = new (std::nothrow) gpumlib_nmf_server;
    if (!s) return hipErrorOutOfMemory;
    s->cfg = *cfg;
    if (s->cfg.sequential) s->cfg.max_active = 1;
    s->t0 = Clock::now();
    memset(&s->stats, 0, sizeof(s->stats));
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_pool, cfg->hbm_budget_bytes);
    if (e == hipSuccess) e = gpumlib_nmf_batch_create(cfg->max_active, &s->batch);
    if (e == hipSuccess && cfg->target_workgroups > 0) e = gpumlib_nmf_batch_set_target(s->batch, cfg->target_workgroups);
    if (e == hipSuccess && cfg->split_k) e = gpumlib_nmf_batch_set_engine(s->batch, GPUMLIB_NMF_BATCH_SPLIT_K);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_staged, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_retired, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_ready, hipEventDisableTiming);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_obj, sizeof(double) * 2 * cfg->max_active, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_obj, 2 * sizeof(double));
    if (e != hipSuccess) {
        if (s->d_obj) (void)hipFree(s->d_obj);
        if (s->h_obj) (void)hipHostFree(s->h_obj);
        if (s->stream) (void)hipStreamDestroy(s->stream);
        if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
        if (s->ev_staged) (void)hipEventDestroy(s->ev_staged);
        if (s->ev_retired) (void)hipEventDestroy(s->ev_retired);
        if (s->ev_ready) (void)hipEventDestroy(s->ev_ready);
        if (s->batch) gpumlib_nmf_batch_destroy(s->batch);
        if (s->d_pool) (void)hipFree(s->d_pool);
        delete s;
        return e;
    }
    s->pool.init(cfg->hbm_budget_bytes);
    s->active.resize(cfg->max_active);
    {
        LiveServers& l = live_servers();
        std::lock_guard<std::mutex> g(l.mu);
        l.set.insert(s);
    }
    s->engine = std::thread([s] { s->run(); });
    *out = s;
    return hipSuccess;
}

hipError_t gpumlib_nmf_server_submit(gpumlib_nmf_server_handle s, int rule, int n, int m, int r, const float* V,
                                     const float* W, const float* H, float eps, double tol, int max_iterations,
                                     long long* job_id) {
    if (!s || !V || !W || !H || !job_id || n < 1 || m < 1 || r < 1 || r > 64 || max_iterations < 1 ||
        (rule != GPUMLIB_NMF_EUCLIDEAN && rule != GPUMLIB_NMF_KL))
        return hipErrorInvalidValue;
    auto j = std::make_shared<Job>();
    j->rule = rule;
    j->n = n;
    j->m = m;
    j->r = r;
    j->eps = eps;
    j->tol = tol;
    j->max_iters = max_iterations;
    j->bytes = job_footprint(s->batch, s->cfg.sequential != 0, *j);
    if (j->bytes == 0) return hipErrorInvalidValue;
    if (j->bytes > s->pool.total()) return hipErrorOutOfMemory;  // can never be admitted
    j->V.assign(V, V + (size_t)n * m);
    j->W.assign(W, W + (size_t)n * r);
    j->H.assign(H, H + (size_t)r * m);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return hipErrorInvalidValue;
        if (s->fatal != hipSuccess) return s->fatal;
        j->id = s->next_id++;
        j->t_submit = s->now();
        s->jobs[j->id] = j;
        s->pending.push_back(j);
        s->stats.submitted++;
        *job_id = j->id;
    }
    s->cv_work.notify_one();
    return hipSuccess;
}

hipError_t gpumlib_nmf_server_wait(gpumlib_nmf_server_handle s, long long job_id, float* W_out, float* H_out,
                                   gpumlib_nmf_job_result* res) {
    if (!s) return hipErrorInvalidValue;
    std::shared_ptr<Job> j;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->jobs.find(job_id);
        if (it == s->jobs.end()) return hipErrorInvalidValue;
        j = it->second;
        s->cv_done.wait(lk, [&] { return j->done; });
        s->jobs.erase(it);
    }
    if (W_out) memcpy(W_out, j->W.data(), sizeof(float) * j->W.size());
    if (H_out) memcpy(H_out, j->H.data(), sizeof(float) * j->H.size());
    if (res) {
        res->iterations = j->iters;
        res->converged = j->converged;
        res->objective = j->objective;
        res->submit_s = j->t_submit;
        res->admit_s = j->t_admit;
        res->finish_s = j->t_finish;
    }
    return j->status;
}

hipError_t gpumlib_nmf_server_stats_get(gpumlib_nmf_server_handle s, gpumlib_nmf_server_stats* out) {
    if (!s || !out) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->pool_bytes = s->pool.total();
    return s->fatal;
}

void gpumlib_nmf_server_destroy(gpumlib_nmf_server_handle s) {
    if (!s) return;
    {
        LiveServers& l = live_servers();
        std::lock_guard<std::mutex> g(l.mu);
        l.set.erase(s);
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->engine.joinable()) s->engine.join();  // finishes every admitted and queued problem first
    (void)hipStreamSynchronize(s->stream);
    (void)hipStreamSynchronize(s->copy_stream);
    (void)hipHostFree(s->h_obj);
    (void)hipFree(s->d_obj);
    (void)hipStreamDestroy(s->stream);
    (void)hipStreamDestroy(s->copy_stream);
    (void)hipEventDestroy(s->ev_staged);
    (void)hipEventDestroy(s->ev_retired);
    (void)hipEventDestroy(s->ev_ready);
    gpumlib_nmf_batch_destroy(s->batch);
    (void)hipFree(s->d_pool);
    delete s;
}

}  // extern "C"
