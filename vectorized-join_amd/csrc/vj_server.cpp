// vj_server: continuous batching of join requests on one GPU.
//
// A worker thread owns two launch lanes (a context with its own arenas, a
// stream and an event each). Whenever a lane is free and requests are
// queued, it admits queued requests in arrival order while the HBM budget
// allows (resident tables + the other lane's arenas + this launch's scratch
// and kept tables) and launches them as one pipeline (every stage one launch
// over the tiles of all admitted joins, one persistent build+probe launch).
// A batch is launched while the other lane's batch is still on the device, so
// requests that arrive during a launch start before it ends. A request that
// arrives at an idle server is launched by the submitting thread itself.
// Completion is detected by the worker, which polls the lanes' events back to
// back while a batch is on the device and sleeps when nothing is.
//
// Reference counterpart: JoinProvider::probe (src/join/join_provider.cuh:792-911),
// which runs bucket pairs on up to 50 detached host threads over round-robin
// streams and admits them against free device memory (DeviceConfig,
// :124-220, reserve_memory / is_out_of_memory :140-150,847-873).
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "vj_host.hpp"
#include "vj_internal.hpp"

using namespace vj;

namespace {

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Resident {
    uint64_t id = 0;
    vj_built *b = nullptr;
    int users = 0;  // queued or running requests that probe it
    bool dropped = false;
};

struct Req {
    uint64_t ticket = 0;
    vj_table build{}, probe{};
    bool has_build = false;
    Resident *res = nullptr;
    uint32_t *out_b = nullptr, *out_p = nullptr;
    uint64_t cap = 0;
    bool keep = false;
    uint64_t need = 0;          // device bytes of this request alone
    uint64_t rows = 0;          // build + probe rows (the size of its device work)
    vj_built *kept = nullptr;   // table being filled by the running launch
    bool done = false;
    vj_response resp{};
};

struct Lane {
    vj_ctx *ctx = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    // set while a thread plans and launches this lane's batch with the lock
    // released; the lane's event is recorded only at the end of that, so until
    // then the event still reports the previous batch (or, never recorded,
    // reports complete) and must not be polled
    bool launching = false;
    std::vector<Req *> reqs;
    uint64_t rows = 0;  // build + probe rows of the running batch
    size_t count_off = 0;
    uint64_t kept_bytes = 0;
};

// Batches of more rows than this do not share the GPU with each other: two
// such joins side by side each take longer than alone (they contend for HBM
// and CUs), while overlap only pays for hiding launch and host time, which a
// join of this size (tens of microseconds of device work) already covers.
constexpr uint64_t kOverlapRows = 1ull << 22;

JoinSpec spec_of(const Req *r) {
    JoinSpec sp{r->has_build ? &r->build : nullptr, &r->probe, r->out_b, r->out_p, r->cap};
    sp.res = r->res ? r->res->b : nullptr;
    sp.keep = r->kept;
    return sp;
}

enum LaunchResult { LAUNCHED, WAITING, RESOLVED };

}  // namespace

struct vj_server {
    int device = 0;
    uint64_t budget = 0;
    vj_join_opts opts{};
    int max_batch = 64;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> pending;
    std::unordered_map<uint64_t, Req *> reqs;
    std::unordered_map<uint64_t, Resident *> tables;
    uint64_t next_ticket = 1, next_table = 1;
    uint64_t resident_bytes = 0;
    Lane lanes[2];
    // one launch at a time (from the worker or a submitting thread): admission
    // reads the other lane's arena sizes, which only a launch changes
    bool launch_in_progress = false;
    std::thread worker;
    bool stop = false;
    vj_server_stats st{};
};

namespace {

// All helpers below run with sv->m held.

void release_if_unused(vj_server *sv, Resident *r) {
    if (!r->dropped || r->users) return;
    sv->resident_bytes -= r->b->bytes;
    built_release(r->b);
    delete r->b;
    sv->tables.erase(r->id);
    delete r;
    sv->st.resident_tables = sv->tables.size();
    sv->st.resident_bytes = sv->resident_bytes;
}

void finish_req(vj_server *sv, Req *r, vj_status status, uint64_t t) {
    r->resp.status = status;
    r->resp.t_done_ns = t;
    if (r->kept) {  // a failed launch leaves no table
        if (status == VJ_OK || status == VJ_ERR_CAPACITY) {
            Resident *res = new Resident();
            res->id = sv->next_table++;
            res->b = r->kept;
            sv->tables[res->id] = res;
            sv->resident_bytes += r->kept->bytes;
            r->resp.table = res->id;
        } else {
            built_release(r->kept);
            delete r->kept;
        }
        r->kept = nullptr;
    }
    if (r->res) {
        r->res->users--;
        release_if_unused(sv, r->res);
        r->res = nullptr;
    }
    r->done = true;
    sv->st.completed++;
    if (status != VJ_OK) sv->st.failed++;
    sv->st.resident_tables = sv->tables.size();
    sv->st.resident_bytes = sv->resident_bytes;
}

void complete_lane(vj_server *sv, Lane &L, vj_status failed) {
    std::vector<uint64_t> counts;
    if (failed == VJ_OK) pipeline_counts(L.ctx, L.count_off, L.reqs.size(), counts);
    const uint64_t t = now_ns();
    for (size_t i = 0; i < L.reqs.size(); i++) {
        Req *r = L.reqs[i];
        vj_status st = failed;
        if (failed == VJ_OK) {
            r->resp.count = counts[i];
            st = counts[i] > r->cap ? VJ_ERR_CAPACITY : VJ_OK;
        }
        finish_req(sv, r, st, t);
    }
    L.reqs.clear();
    L.busy = false;
    L.kept_bytes = 0;
    sv->cv_done.notify_all();
}

uint64_t held_now(vj_server *sv) {
    uint64_t b = sv->resident_bytes;
    for (Lane &L : sv->lanes) b += ctx_bytes(L.ctx) + L.kept_bytes;
    return b;
}

// Forms a batch for lane `L` from the queue and launches it.
LaunchResult launch_batch(vj_server *sv, Lane &L, Lane &O, std::unique_lock<std::mutex> &lk) {
    uint64_t avail = 0;  // 0 = unlimited
    if (sv->budget) {
        const uint64_t used = sv->resident_bytes + ctx_bytes(O.ctx) + O.kept_bytes;
        avail = sv->budget > used ? sv->budget - used : 1;
    }
    std::vector<Req *> batch;
    uint64_t need = 0;
    bool budget_cut = false;
    const bool big_running = O.busy && O.rows > kOverlapRows;
    while (!sv->pending.empty() && (int)batch.size() < sv->max_batch) {
        Req *r = sv->pending.front();
        if (big_running && r->rows > kOverlapRows) break;  // waits for the large batch (see try_launch)
        if (avail && need + r->need > avail) {
            budget_cut = true;
            break;
        }
        need += r->need;
        batch.push_back(r);
        sv->pending.pop_front();
    }
    if (budget_cut && !batch.empty()) sv->st.budget_deferred++;
    if (batch.empty()) {
        Req *r = sv->pending.front();
        if (r->need + sv->resident_bytes > sv->budget) {  // cannot fit even with both lanes empty
            sv->pending.pop_front();
            finish_req(sv, r, VJ_ERR_BUDGET, now_ns());
            sv->cv_done.notify_all();
            return RESOLVED;
        }
        sv->st.budget_waits++;
        if (O.busy) return WAITING;
        ctx_release(O.ctx);  // the idle lane's arenas make room
        return RESOLVED;
    }
    const uint64_t t = now_ns();
    L.kept_bytes = 0;
    for (Req *r : batch) {
        r->resp.t_launch_ns = t;
        r->resp.batch_size = (uint32_t)batch.size();
        if (r->keep) r->kept = new vj_built();
    }
    L.busy = true;
    L.launching = true;
    sv->launch_in_progress = true;
    L.reqs = batch;
    L.rows = 0;
    for (Req *r : batch) L.rows += r->rows;
    sv->st.batches++;
    if (O.busy) sv->st.overlapped_batches++;
    if (batch.size() > sv->st.max_batch) sv->st.max_batch = batch.size();
    std::vector<JoinSpec> specs;
    specs.reserve(batch.size());
    for (Req *r : batch) specs.push_back(spec_of(r));
    const vj_join_opts opts = sv->opts;
    lk.unlock();  // HIP work without the lock: submitters are not held up
    Err e = pipeline_launch(L.ctx, specs, &opts, L.s, avail, &L.count_off);
    if (!e.failed()) {
        hipError_t he = hipEventRecord(L.ev, L.s);
        if (he != hipSuccess) e = Err::make(VJ_ERR_HIP, hipGetErrorString(he));
    }
    // A launch that failed part-way may have queued stages that still read the
    // requests' tables and write their outputs: they finish before the
    // requests are reported failed (and their buffers can be freed).
    if (e.failed()) (void)hipStreamSynchronize(L.s);
    lk.lock();
    L.launching = false;
    sv->launch_in_progress = false;
    for (Req *r : batch)
        if (r->kept) L.kept_bytes += r->kept->bytes;
    const uint64_t held = held_now(sv);
    if (held > sv->st.peak_bytes) sv->st.peak_bytes = held;
    if (e.failed()) complete_lane(sv, L, e.code);
    return LAUNCHED;
}

// Launches queued requests on a free lane, if any (with sv->m held; the lock
// is released during the HIP calls). Returns true if something happened.
bool try_launch(vj_server *sv, std::unique_lock<std::mutex> &lk) {
    if (sv->pending.empty() || sv->launch_in_progress) return false;
    Lane *F = !sv->lanes[0].busy ? &sv->lanes[0] : (!sv->lanes[1].busy ? &sv->lanes[1] : nullptr);
    if (!F) return false;
    Lane &O = F == &sv->lanes[0] ? sv->lanes[1] : sv->lanes[0];
    // a large request waits for a large running batch to finish (it then
    // goes out at once, batched with whatever queued behind it)
    if (O.busy && O.rows > kOverlapRows && sv->pending.front()->rows > kOverlapRows) return false;
    return launch_batch(sv, *F, O, lk) != WAITING;
}

void worker_main(vj_server *sv) {
    (void)hipSetDevice(sv->device);
    std::unique_lock<std::mutex> lk(sv->m);
    while (true) {
        bool progressed = false;
        for (Lane &L : sv->lanes) {
            if (!L.busy || L.launching) continue;
            const hipError_t q = hipEventQuery(L.ev);
            if (q == hipSuccess) {
                complete_lane(sv, L, VJ_OK);
                progressed = true;
            } else if (q != hipErrorNotReady) {
                (void)hipGetLastError();
                complete_lane(sv, L, VJ_ERR_HIP);
                progressed = true;
            }
        }
        if (try_launch(sv, lk)) progressed = true;
        const bool busy = sv->lanes[0].busy || sv->lanes[1].busy;
        if (sv->stop && sv->pending.empty() && !busy) break;
        if (progressed) continue;
        if (busy) {
            // a batch is on the device: poll its event again at once (a timed
            // wait here would add its period to every request's latency)
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        } else {
            sv->cv_work.wait(lk);
        }
    }
}

void destroy_lanes(vj_server *sv) {
    for (Lane &L : sv->lanes) {
        if (L.ctx) (void)vj_ctx_destroy(L.ctx);
        if (L.ev) (void)hipEventDestroy(L.ev);
        if (L.s) (void)hipStreamDestroy(L.s);
        L.ctx = nullptr;
        L.ev = nullptr;
        L.s = nullptr;
    }
}

}  // namespace

extern "C" {

vj_status vj_server_create(int device, uint64_t hbm_budget, const vj_join_opts *opts, int32_t max_batch,
                           vj_server **out) {
    if (!out || max_batch < 0) return VJ_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    vj_server *sv = new vj_server();
    sv->device = device;
    sv->budget = hbm_budget;
    if (opts)
        sv->opts = *opts;
    else
        vj_join_opts_default(&sv->opts);
    sv->max_batch = max_batch ? max_batch : 64;
    for (Lane &L : sv->lanes) {
        vj_status st = vj_ctx_create(device, &L.ctx);
        if (st == VJ_OK && hipStreamCreateWithFlags(&L.s, hipStreamNonBlocking) != hipSuccess) st = VJ_ERR_HIP;
        if (st == VJ_OK && hipEventCreateWithFlags(&L.ev, hipEventDisableTiming) != hipSuccess) st = VJ_ERR_HIP;
        if (st != VJ_OK) {
            destroy_lanes(sv);
            delete sv;
            return st;
        }
    }
    sv->worker = std::thread(worker_main, sv);
    *out = sv;
    return VJ_OK;
}

vj_status vj_server_submit(vj_server *sv, const vj_request *rq, uint64_t *ticket) {
    if (!sv || !rq || !ticket || !rq->probe || (!rq->build && !rq->table)) return VJ_ERR_INVALID_ARGUMENT;
    if (rq->capacity && (!rq->out_build || !rq->out_probe)) return VJ_ERR_INVALID_ARGUMENT;
    if (!rq->build && rq->keep_build) return VJ_ERR_INVALID_ARGUMENT;
    Req *r = new Req();
    r->probe = *rq->probe;
    r->has_build = rq->build != nullptr;
    if (r->has_build) r->build = *rq->build;
    r->out_b = rq->out_build;
    r->out_p = rq->out_probe;
    r->cap = rq->capacity;
    r->keep = rq->keep_build != 0;
    r->rows = (r->has_build ? r->build.n_rows : 0) + r->probe.n_rows;
    std::unique_lock<std::mutex> lk(sv->m);
    if (!r->has_build) {
        auto it = sv->tables.find(rq->table);
        if (it == sv->tables.end() || it->second->dropped) {
            delete r;
            return VJ_ERR_NOT_FOUND;
        }
        r->res = it->second;
    }
    std::vector<JoinSpec> one{spec_of(r)};
    vj_built sizing;  // a kept table is sized as part of the request
    if (r->keep) one[0].keep = &sizing;
    Err e = pipeline_bytes(one, &sv->opts, &r->need);
    if (e.failed()) {
        delete r;
        return e.code;
    }
    if (r->res) r->res->users++;
    r->ticket = sv->next_ticket++;
    r->resp.t_submit_ns = now_ns();
    sv->reqs[r->ticket] = r;
    sv->pending.push_back(r);
    sv->st.submitted++;
    *ticket = r->ticket;
    // Light load: the queue held only this request and a lane is free, so the
    // submitting thread launches it at once instead of handing it to the
    // worker (whose wake-up would add to the request's latency). The worker
    // is woken either way: it detects the completion.
    if (sv->pending.size() == 1 && !sv->stop) {
        int prev = -1;
        const bool have_prev = hipGetDevice(&prev) == hipSuccess;
        (void)try_launch(sv, lk);                                     // switches to the server's device
        if (have_prev && prev != sv->device) (void)hipSetDevice(prev);  // the caller's thread keeps its device
    }
    lk.unlock();
    sv->cv_work.notify_one();
    return VJ_OK;
}

vj_status vj_server_wait(vj_server *sv, uint64_t ticket, vj_response *resp) {
    if (!sv || !resp) return VJ_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(sv->m);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return VJ_ERR_NOT_FOUND;
    Req *r = it->second;
    sv->cv_done.wait(lk, [&] { return r->done; });
    *resp = r->resp;
    sv->reqs.erase(ticket);
    delete r;
    return VJ_OK;
}

vj_status vj_server_poll(vj_server *sv, uint64_t ticket, vj_response *resp, int32_t *done) {
    if (!sv || !resp || !done) return VJ_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(sv->m);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return VJ_ERR_NOT_FOUND;
    Req *r = it->second;
    *done = r->done ? 1 : 0;
    if (r->done) {
        *resp = r->resp;
        sv->reqs.erase(ticket);
        delete r;
    }
    return VJ_OK;
}

vj_status vj_server_drop_table(vj_server *sv, uint64_t table) {
    if (!sv) return VJ_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(sv->m);
    auto it = sv->tables.find(table);
    if (it == sv->tables.end() || it->second->dropped) return VJ_ERR_NOT_FOUND;
    it->second->dropped = true;
    release_if_unused(sv, it->second);
    return VJ_OK;
}

vj_status vj_server_stats_get(vj_server *sv, vj_server_stats *stats) {
    if (!sv || !stats) return VJ_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(sv->m);
    *stats = sv->st;
    return VJ_OK;
}

vj_status vj_server_destroy(vj_server *sv) {
    if (!sv) return VJ_ERR_INVALID_ARGUMENT;
    {
        std::unique_lock<std::mutex> lk(sv->m);
        sv->stop = true;
    }
    sv->cv_work.notify_one();
    sv->worker.join();
    (void)hipSetDevice(sv->device);
    for (auto &kv : sv->reqs) delete kv.second;
    for (auto &kv : sv->tables) {
        built_release(kv.second->b);
        delete kv.second->b;
        delete kv.second;
    }
    destroy_lanes(sv);
    delete sv;
    return VJ_OK;
}

}  // extern "C"
