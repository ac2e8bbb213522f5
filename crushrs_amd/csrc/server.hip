// Serving: continuous batching of crash solves on one GPU (crush_server_*).
//
// Requests are prepared on the submitting thread and queued.  One driver
// thread owns the GPU: between replays of the step graph it admits queued
// requests into free slots whose images fit the free part of a fixed device
// pool (the HBM budget), rebuilds the active lists, replays a few steps of
// every resident request, and retires the requests whose control blocks
// say done.  The step kernels, lists and graph are the batched engine's
// (engine.hpp), so each result is bitwise the one-model solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "common.hpp"
#include "engine.hpp"
#include "outputs.hpp"
#include "signal.hpp"

namespace crush {

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// First-fit allocator over [0, size) with 256-byte granules.
class PoolAlloc {
  public:
    explicit PoolAlloc(size_t size = 0) { reset(size); }
    void reset(size_t size) {
        free_.clear();
        if (size) free_[0] = size;
        used_ = 0;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = (bytes + 255) & ~size_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t at = it->first + bytes;
                free_.erase(it);
                if (rest) free_[at] = rest;
                used_ += bytes;
                return true;
            }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }

  private:
    std::map<size_t, size_t> free_;
    size_t used_ = 0;
};

struct Request {
    int64_t ticket = 0;
    int32_t flags = 0;
    Prepared P;
    ImageOffsets o{};
    std::vector<uint8_t> image;
    size_t bytes = 0;
    int slot = -1;
    size_t off = 0;
    Clock::time_point t_submit, t_admit, t_done;
    // result
    bool done = false;
    int status = CRUSH_OK;
    std::string error;
    StepCtl ctl{};
    std::vector<double> dv, u, v, pulse, hist;
};

}  // namespace

struct ServerImpl {
    crush_server_options opt{};
    BatchImpl eng;
    PoolAlloc alloc;
    size_t budget = 0;
    int capacity = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<int64_t, std::shared_ptr<Request>> tickets;
    std::vector<std::shared_ptr<Request>> slot_req;
    std::vector<int> free_slots;
    std::vector<uint8_t> flags_host;
    int resident = 0;
    int64_t next_ticket = 1, n_submitted = 0, n_admitted = 0, n_completed = 0, n_failed = 0, element_steps = 0;
    std::vector<double> lat, qlat;
    double busy_ms = 0.0;
    double admit_ms = 0.0, lists_ms = 0.0, replay_ms = 0.0, retire_ms = 0.0;
    int64_t captures = 0;
    bool stop = false;
    std::string fatal;
    std::thread driver;

    ~ServerImpl() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (driver.joinable()) driver.join();
    }
};

void server_defaults(crush_server_options* o) {
    std::memset(o, 0, sizeof *o);
    o->hbm_budget_bytes = 16.0 * (1ull << 30);
    o->max_resident = 4096;
    o->steps_per_graph = 16;
    o->threads_per_instance = 256;
    o->force_tile = 256;
    o->contact_cell_scale = 1.0;
    o->max_queued = 1 << 20;
}

// Admit what fits; returns the admitted requests (slots and offsets set).
static std::vector<std::shared_ptr<Request>> take_admissions(ServerImpl& S) {
    std::vector<std::shared_ptr<Request>> out;
    while (!S.queue.empty() && !S.free_slots.empty()) {
        auto& r = S.queue.front();
        size_t off;
        if (!S.alloc.alloc(r->bytes, &off)) break;  // first in, first admitted
        r->off = off;
        r->slot = S.free_slots.back();
        S.free_slots.pop_back();
        r->t_admit = Clock::now();
        S.slot_req[r->slot] = r;
        out.push_back(r);
        S.queue.pop_front();
    }
    return out;
}

static void upload(ServerImpl& S, Request& r) {
    BatchImpl& E = S.eng;
    uint8_t* base = E.pool.p + r.off;
    HIP_OK(hipMemcpyAsync(base, r.image.data(), r.image.size(), hipMemcpyHostToDevice, E.stream));
    InstDev inst = make_inst(r.P, r.o, reinterpret_cast<uintptr_t>(base));
    inst.cell_scale = cell_scale_of(E.opt);
    HIP_OK(hipMemcpyAsync(E.table.p + r.slot, &inst, sizeof inst, hipMemcpyHostToDevice, E.stream));
    const uint8_t zero = 0;
    HIP_OK(hipMemcpyAsync(E.done_flag.p + r.slot, &zero, 1, hipMemcpyHostToDevice, E.stream));
    // the copies read host memory that must stay alive until they finish
    HIP_OK(hipStreamSynchronize(E.stream));
    E.prep[r.slot] = std::move(r.P);
    E.host_done[r.slot] = 0;
    r.image.clear();
    r.image.shrink_to_fit();
}

static void read_result(ServerImpl& S, Request& r) {
    BatchImpl& E = S.eng;
    const Prepared& P = E.prep[r.slot];
    uint8_t* base = E.pool.p + r.off;
    HIP_OK(hipMemcpy(&r.ctl, base + r.o.ctl, sizeof(StepCtl), hipMemcpyDeviceToHost));
    r.dv.assign(3 * (size_t)std::max(P.n_groups, 1), 0.0);
    if (P.n_groups > 0)
        HIP_OK(hipMemcpy(r.dv.data(), base + r.o.g_dv, sizeof(double) * 3 * P.n_groups, hipMemcpyDeviceToHost));
    if (r.flags & CRUSH_SUBMIT_FIELDS) {
        r.u.resize(3 * (size_t)P.n_nodes);
        r.v.resize(3 * (size_t)P.n_nodes);
        HIP_OK(hipMemcpy(r.u.data(), base + r.o.u, sizeof(double) * r.u.size(), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(r.v.data(), base + r.o.v, sizeof(double) * r.v.size(), hipMemcpyDeviceToHost));
    }
    r.pulse = pulse_download(P, reinterpret_cast<const double*>(base + r.o.pulse));
    r.hist = history_download(P, reinterpret_cast<const double*>(base + r.o.hist), r.ctl.hist_k, E.stream);
    if (r.ctl.error) {
        r.status = CRUSH_E_NONFINITE;
        r.error = "displacement became non-finite at step " + std::to_string(r.ctl.step);
    }
}

static void driver_loop(ServerImpl* Sp) {
    ServerImpl& S = *Sp;
    BatchImpl& E = S.eng;
    try {
        HIP_OK(hipSetDevice(0));
        bool lists_dirty = false;
        for (;;) {
            std::vector<std::shared_ptr<Request>> admitted;
            {
                std::unique_lock<std::mutex> lk(S.mu);
                S.cv_work.wait(lk, [&] { return S.stop || S.resident > 0 || !S.queue.empty(); });
                if (S.stop) break;
                admitted = take_admissions(S);
                S.resident += (int)admitted.size();
                S.n_admitted += (int64_t)admitted.size();
            }
            const auto ta = Clock::now();
            for (auto& r : admitted) upload(S, *r);
            const auto tb = Clock::now();
            int captured = 0;
            if (!admitted.empty() || lists_dirty) {
                rebuild_lists(E);
                // capture again when a list outgrows its grid, or when the
                // lists have shrunk to a quarter of it
                bool again = E.need_capture || 4 * E.cnt[L_INST] < E.cap[L_INST];
                for (int L = 0; L < N_LISTS; ++L) again = again || E.cnt[L] > E.cap[L];
                if (again && E.cnt[L_INST] > 0) {
                    capture(E);
                    captured = 1;
                }
                lists_dirty = false;
            }
            const auto tc = Clock::now();
            if (E.cnt[L_INST] == 0) continue;
            const auto t0 = Clock::now();
            if (!admitted.empty()) dt_init(E);
            replay(E);
            HIP_OK(hipMemcpyAsync(S.flags_host.data(), E.done_flag.p, S.capacity, hipMemcpyDeviceToHost, E.stream));
            HIP_OK(hipStreamSynchronize(E.stream));
            E.replays += 1;
            const auto td = Clock::now();
            // retire
            std::vector<std::shared_ptr<Request>> finished;
            for (int slot = 0; slot < S.capacity; ++slot)
                if (S.slot_req[slot] && S.flags_host[slot]) finished.push_back(S.slot_req[slot]);
            for (auto& r : finished) read_result(S, *r);
            const auto t1 = Clock::now();
            {
                std::lock_guard<std::mutex> g(S.mu);
                S.admit_ms += ms_between(ta, tb);
                S.lists_ms += ms_between(tb, tc);
                S.replay_ms += ms_between(t0, td);
                S.retire_ms += ms_between(td, t1);
                S.captures += captured;
            }
            if (!finished.empty()) {
                std::lock_guard<std::mutex> g(S.mu);
                for (auto& r : finished) {
                    r->t_done = t1;
                    r->done = true;
                    S.lat.push_back(ms_between(r->t_submit, r->t_done));
                    S.qlat.push_back(ms_between(r->t_submit, r->t_admit));
                    S.element_steps += (int64_t)E.prep[r->slot].n_elems * r->ctl.step;
                    r->P = std::move(E.prep[r->slot]);
                    if (r->status == CRUSH_OK)
                        S.n_completed += 1;
                    else
                        S.n_failed += 1;
                    E.host_done[r->slot] = 1;
                    S.alloc.release(r->off, r->bytes);
                    S.slot_req[r->slot].reset();
                    S.free_slots.push_back(r->slot);
                    S.resident -= 1;
                }
                lists_dirty = true;
            }
            {
                std::lock_guard<std::mutex> g(S.mu);
                S.busy_ms += ms_between(t0, t1);
            }
            if (!finished.empty()) S.cv_done.notify_all();
        }
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> g(S.mu);
        S.fatal = e.what();
        for (auto& kv : S.tickets)
            if (!kv.second->done) {
                kv.second->done = true;
                kv.second->status = CRUSH_E_DEVICE;
                kv.second->error = S.fatal;
            }
        S.cv_done.notify_all();
    }
}

ServerImpl* server_create(const crush_server_options* opt) {
    require_device();
    auto S = std::make_unique<ServerImpl>();
    server_defaults(&S->opt);
    if (opt) {
        S->opt = *opt;
        crush_server_options d;
        server_defaults(&d);
        if (S->opt.hbm_budget_bytes <= 0) S->opt.hbm_budget_bytes = d.hbm_budget_bytes;
        if (S->opt.max_resident <= 0) S->opt.max_resident = d.max_resident;
        if (S->opt.steps_per_graph <= 0) S->opt.steps_per_graph = d.steps_per_graph;
        if (S->opt.max_queued <= 0) S->opt.max_queued = d.max_queued;
    }
    S->capacity = S->opt.max_resident;
    S->budget = (size_t)S->opt.hbm_budget_bytes;
    BatchImpl& E = S->eng;
    batch_defaults(&E.opt);
    E.opt.threads_per_instance = S->opt.threads_per_instance;
    E.opt.force_tile = S->opt.force_tile;
    E.opt.contact_cell_scale = S->opt.contact_cell_scale;
    E.opt.steps_per_graph = S->opt.steps_per_graph;
    E.headroom = 1.5;
    HIP_OK(hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking));
    E.pool.alloc(S->budget);
    E.table.alloc(S->capacity);
    E.done_flag.alloc(S->capacity);
    E.d_counts.alloc(N_LISTS);
    E.n_done.alloc(1);
    E.n_done.zero(E.stream);
    E.prep.resize(S->capacity);
    E.host_done.assign(S->capacity, 1);
    E.fresh = false;
    HIP_OK(hipStreamSynchronize(E.stream));
    S->alloc.reset(S->budget);
    S->slot_req.resize(S->capacity);
    for (int i = S->capacity - 1; i >= 0; --i) S->free_slots.push_back(i);
    S->flags_host.assign(S->capacity, 0);
    S->driver = std::thread(driver_loop, S.get());
    return S.release();
}

int64_t server_submit(ServerImpl& S, const crush_model& m, int32_t flags) {
    auto r = std::make_shared<Request>();
    r->flags = flags;
    prepare(m, r->P);
    r->o = build_image(r->P, r->image);
    r->bytes = (r->o.total + 255) & ~size_t(255);
    if (r->bytes > S.budget)
        fail(CRUSH_E_BUSY, "model needs %zu bytes of device memory, more than the budget of %zu", r->bytes, S.budget);
    std::lock_guard<std::mutex> g(S.mu);
    if (!S.fatal.empty()) fail(CRUSH_E_DEVICE, "server stopped: %s", S.fatal.c_str());
    if ((int64_t)S.queue.size() >= S.opt.max_queued) fail(CRUSH_E_BUSY, "server queue full");
    r->ticket = S.next_ticket++;
    r->t_submit = Clock::now();
    S.tickets[r->ticket] = r;
    S.queue.push_back(r);
    S.n_submitted += 1;
    S.cv_work.notify_one();
    return r->ticket;
}

static void fill(const Request& r, const Prepared& P, crush_results* out) {
    out->time = r.ctl.t;
    out->steps = r.ctl.step;
    out->eroded_elements = (int64_t)r.ctl.eroded_total;
    out->added_mass_max = r.ctl.added_total_max;
    out->separated = r.ctl.separated;
    out->wall_ms = ms_between(r.t_submit, r.t_done);
    out->device_ms = ms_between(r.t_admit, r.t_done);
    out->n_groups = P.n_groups;
    if (out->displacement && !r.u.empty()) std::memcpy(out->displacement, r.u.data(), sizeof(double) * r.u.size());
    if (out->velocity && !r.v.empty()) std::memcpy(out->velocity, r.v.data(), sizeof(double) * r.v.size());
    if (out->masses) std::memcpy(out->masses, P.mass.data(), sizeof(double) * P.n_nodes);
    if (out->groups)
        for (int g = 0; g < P.n_groups; ++g) {
            crush_group_result& gr = out->groups[g];
            gr.mass = P.g_mass[g];
            for (int d = 0; d < 3; ++d) gr.delta_v[d] = r.dv[3 * g + d];
            gr.crush_end = g < 8 ? r.ctl.crush_end[g] : 0.0;
            gr.crush_max = g < 8 ? r.ctl.crush_max[g] : 0.0;
        }
    pulse_results_host(P, r.pulse.empty() ? nullptr : r.pulse.data(), r.ctl.pulse_k, r.ctl.t, out);
    history_results(P, r.hist.data(), r.ctl.hist_k, out);
    out->frame_samples = 0;
}

// Hand a finished request to the caller and release its ticket (mu held).
static void deliver(ServerImpl& S, int64_t ticket, crush_results* out) {
    auto it = S.tickets.find(ticket);
    std::shared_ptr<Request> r = it->second;
    S.tickets.erase(it);
    if (r->status != CRUSH_OK) fail(r->status, "request %lld: %s", (long long)ticket, r->error.c_str());
    fill(*r, r->P, out);
}

int server_poll(ServerImpl& S, int64_t ticket, crush_results* out) {
    std::lock_guard<std::mutex> g(S.mu);
    auto it = S.tickets.find(ticket);
    if (it == S.tickets.end()) fail(CRUSH_E_NOTFOUND, "unknown ticket %lld", (long long)ticket);
    if (!it->second->done) return 0;
    deliver(S, ticket, out);
    return 1;
}

void server_wait(ServerImpl& S, int64_t ticket, crush_results* out, double timeout_ms) {
    std::unique_lock<std::mutex> lk(S.mu);
    auto it = S.tickets.find(ticket);
    if (it == S.tickets.end()) fail(CRUSH_E_NOTFOUND, "unknown ticket %lld", (long long)ticket);
    std::shared_ptr<Request> r = it->second;
    auto pred = [&] { return r->done; };
    if (timeout_ms > 0) {
        if (!S.cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), pred))
            fail(CRUSH_E_BUSY, "request %lld not finished after %.0f ms", (long long)ticket, timeout_ms);
    } else {
        S.cv_done.wait(lk, pred);
    }
    deliver(S, ticket, out);
}

void server_stats(ServerImpl& S, crush_server_stats* st) {
    std::lock_guard<std::mutex> g(S.mu);
    std::memset(st, 0, sizeof *st);
    st->submitted = S.n_submitted;
    st->admitted = S.n_admitted;
    st->completed = S.n_completed;
    st->failed = S.n_failed;
    st->resident = S.resident;
    st->queued = (int32_t)S.queue.size();
    st->device_bytes_in_use = (double)S.alloc.used();
    auto pct = [](std::vector<double> v, double q) {
        if (v.empty()) return 0.0;
        std::sort(v.begin(), v.end());
        return v[std::min(v.size() - 1, (size_t)std::ceil(q * v.size()) - 1)];
    };
    st->latency_p50_ms = pct(S.lat, 0.50);
    st->latency_p99_ms = pct(S.lat, 0.99);
    st->queue_p50_ms = pct(S.qlat, 0.50);
    st->busy_ms = S.busy_ms;
    st->element_steps = S.element_steps;
    st->replays = S.eng.replays;
    st->captures = S.captures;
    st->admit_ms = S.admit_ms;
    st->lists_ms = S.lists_ms;
    st->replay_ms = S.replay_ms;
    st->retire_ms = S.retire_ms;
}

}  // namespace crush

struct crush_server {
    crush::ServerImpl* impl;
};

using namespace crush;

extern "C" {

void crush_server_defaults(crush_server_options* opt) {
    if (opt) server_defaults(opt);
}

int32_t crush_server_create(const crush_server_options* opt, crush_server** out) {
    return guarded([&] {
        if (!out) fail(CRUSH_E_INPUT, "null argument");
        *out = new crush_server{server_create(opt)};
    });
}

int32_t crush_server_submit(crush_server* s, const crush_model* model, int32_t flags, int64_t* ticket) {
    return guarded([&] {
        if (!s || !model || !ticket) fail(CRUSH_E_INPUT, "null argument");
        *ticket = server_submit(*s->impl, *model, flags);
    });
}

int32_t crush_server_poll(crush_server* s, int64_t ticket, crush_results* out, int32_t* done) {
    return guarded([&] {
        if (!s || !out || !done) fail(CRUSH_E_INPUT, "null argument");
        *done = server_poll(*s->impl, ticket, out);
    });
}

int32_t crush_server_wait(crush_server* s, int64_t ticket, crush_results* out, double timeout_ms) {
    return guarded([&] {
        if (!s || !out) fail(CRUSH_E_INPUT, "null argument");
        server_wait(*s->impl, ticket, out, timeout_ms);
    });
}

int32_t crush_server_stats_get(crush_server* s, crush_server_stats* st) {
    return guarded([&] {
        if (!s || !st) fail(CRUSH_E_INPUT, "null argument");
        server_stats(*s->impl, st);
    });
}

void crush_server_destroy(crush_server* s) {
    if (!s) return;
    delete s->impl;
    delete s;
}

}  // extern "C"
