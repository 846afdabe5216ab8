// ekg_server.cpp -- continuous batching on one GPU.
//
// The reference evaluates one simulation per MPI worker and keeps per-worker
// job queues topped up by the master (ParallelFramework.h:286-430); an
// external evaluator is one process per simulation (ExternalEvaluation.h:95-151).
// Here one worker thread owns n_streams launch slots.  Requests queue FIFO;
// whenever a slot is free the worker takes the oldest request and every
// queued request with the same dipole neighbourhood and contraction path
// (the path each heart's tuning selects) -- any heart, lead count, duration
// and time step -- up to max_batch, shrinks the group until
// its device bytes fit the budget and the in-flight cap, and enqueues it on
// the slot's stream as one launch group (one launch per kernel).  Groups on
// different slots run concurrently, so a request that arrives while others
// run is admitted as soon as a slot and memory free up, and its kernels
// overlap with the running ones.  Results are bitwise independent of the
// grouping: every request's sums come from its own work items over its
// heart's fixed chunking, in a fixed order.
#include "ekg_server.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

namespace ekg {

using Clock = std::chrono::steady_clock;

namespace {

// Pinned host memory for the ECGs of one launch group.  The group's requests
// share it: each keeps it alive until its result has been copied out by
// ekg_server_wait, then it returns to the server's pool for a later group.
struct HostBuf {
    double* p = nullptr;
    size_t cap = 0;  // doubles
};

class HostPool {
public:
    // a buffer of at least n doubles: the smallest free one that fits, or a
    // newly pinned one of the next power of two (at least 64 Ki doubles), so
    // that groups of varying size soon find pinned buffers to reuse
    std::shared_ptr<HostBuf> take(size_t n)
    {
        HostBuf* b = nullptr;
        {
            std::lock_guard<std::mutex> g(mu_);
            size_t best = free_.size();
            for (size_t i = 0; i < free_.size(); ++i)
                if (free_[i]->cap >= n && (best == free_.size() || free_[i]->cap < free_[best]->cap)) best = i;
            if (best < free_.size()) {
                b = free_[best];
                free_.erase(free_.begin() + (std::ptrdiff_t)best);
            }
        }
        if (!b) {
            std::unique_ptr<HostBuf> nb(new HostBuf);
            size_t cap = (size_t)1 << 16;
            while (cap < n) cap <<= 1;
            hip_check(hipHostMalloc((void**)&nb->p, cap * 8, hipHostMallocDefault), "hipHostMalloc");
            nb->cap = cap;
            b = nb.release();
        }
        return std::shared_ptr<HostBuf>(b, [this](HostBuf* q) {
            std::lock_guard<std::mutex> g(mu_);
            free_.push_back(q);
        });
    }
    // called once no result refers to a buffer any more
    ~HostPool()
    {
        for (HostBuf* b : free_) {
            if (b->p) (void)hipHostFree(b->p);
            delete b;
        }
    }

private:
    std::mutex mu_;
    std::vector<HostBuf*> free_;
};

struct Request {
    uint64_t ticket = 0;
    int heart_id = 0;
    std::vector<double> layer_k, lead_pos;
    ekg_sim_desc desc{};
    int64_t T = 0;
    enum State { Pending, Running, Done, Failed } state = Pending;
    int err = EKG_OK;
    std::string msg;
    // the result: n_out doubles at out->p + out_off
    std::shared_ptr<HostBuf> out;
    size_t out_off = 0, n_out = 0;
    bool viewed = false;  // handed out by server_wait_view, kept until server_release
    Clock::time_point t_submit, t_done;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Workspace ws;
    bool busy = false;
    std::vector<Request*> reqs;
    std::vector<int64_t> out_off;
    std::shared_ptr<HostBuf> h_out;
    size_t bytes = 0;
};

}  // namespace

struct Server {
    HostPool pool;  // destroyed last: every result buffer has returned by then
    Ctx& ctx;
    ekg_server_opts opts{};
    std::vector<Heart*> hearts;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    std::vector<std::unique_ptr<Slot>> slots;
    std::thread worker;
    bool stop = false;
    uint64_t next_ticket = 1, completed = 0, groups = 0;
    int inflight = 0, peak_inflight = 0;
    size_t bytes_in_use = 0, peak_bytes = 0;

    explicit Server(Ctx& c) : ctx(c) {}
    // Requests that may share a launch group: the same dipole neighbourhood
    // and the same contraction path of their hearts (a group runs one path,
    // and each heart's layout is the one of its own path).
    bool same_group(const Request& a, const Request& b) const
    {
        return a.desc.nhood == b.desc.nhood &&
               heart_mfma(ctx, *hearts[a.heart_id]) == heart_mfma(ctx, *hearts[b.heart_id]);
    }
    // Requests the group of the oldest pending request could take now: its
    // group's pending requests up to max_batch and the in-flight cap.
    int group_room() const
    {
        const Request& head = *pending.front();
        int n = 0;
        const int cap = std::min(opts.max_batch, opts.max_inflight - inflight);
        for (const Request* r : pending) {
            if (n >= cap) break;
            if (same_group(head, *r)) ++n;
        }
        return cap - n;
    }
    void run();
    void retire(Slot& s);
    void fail_group(std::vector<Request*>& g, int code, const std::string& msg);
};

void Server::fail_group(std::vector<Request*>& g, int code, const std::string& msg)
{
    for (Request* r : g) {
        r->state = Request::Failed;
        r->err = code;
        r->msg = msg;
        r->t_done = Clock::now();
    }
    cv_done.notify_all();
}

// Called with mu held once the slot's completion event has fired.
void Server::retire(Slot& s)
{
    for (size_t b = 0; b < s.reqs.size(); ++b) {
        Request* r = s.reqs[b];
        r->out = s.h_out;  // no copy here: ekg_server_wait copies out of the group's buffer
        r->out_off = (size_t)s.out_off[b];
        r->n_out = (size_t)(s.out_off[b + 1] - s.out_off[b]);
        r->state = Request::Done;
        r->t_done = Clock::now();
        ++completed;
    }
    inflight -= (int)s.reqs.size();
    bytes_in_use -= s.bytes;
    s.reqs.clear();
    s.h_out.reset();
    s.busy = false;
    s.bytes = 0;
    cv_done.notify_all();
}

void Server::run()
{
    (void)hipSetDevice(ctx.device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        bool any_busy = false;
        for (auto& sp : slots) {
            Slot& s = *sp;
            if (!s.busy) continue;
            const hipError_t q = hipEventQuery(s.done);
            if (q == hipSuccess) {
                retire(s);
            } else if (q != hipErrorNotReady) {
                fail_group(s.reqs, EKG_EDEVICE, std::string("device error: ") + hipGetErrorString(q));
                inflight -= (int)s.reqs.size();
                bytes_in_use -= s.bytes;
                s.reqs.clear();
                s.h_out.reset();
                s.busy = false;
            } else {
                any_busy = true;
            }
        }
        if (stop) {
            std::vector<Request*> rest(pending.begin(), pending.end());
            pending.clear();
            fail_group(rest, EKG_ESTATE, "server stopped before the request was admitted");
            if (!any_busy) break;
        }
        Slot* free_slot = nullptr;
        for (auto& sp : slots)
            if (!sp->busy) {
                free_slot = sp.get();
                break;
            }
        bool launched = false;
        // Batching window: while another group runs (the GPU is busy anyway),
        // a group that could still take more requests (below max_batch and
        // the in-flight cap, inside the device budget) waits for them, up to
        // batch_window_us from its oldest request; an idle GPU starts at once.
        bool hold = false;
        if (free_slot && !pending.empty() && !stop && any_busy && opts.batch_window_us > 0 &&
            bytes_in_use < opts.hbm_budget_bytes && group_room() > 0) {
            const auto age = std::chrono::duration_cast<std::chrono::microseconds>(Clock::now() -
                                                                                  pending.front()->t_submit);
            hold = age.count() < opts.batch_window_us;
        }
        if (free_slot && !pending.empty() && !stop && !hold) {
            Request* head = pending.front();
            std::vector<Request*> group;
            const int cap_inflight = opts.max_inflight - inflight;
            for (Request* r : pending) {
                if ((int)group.size() >= opts.max_batch || (int)group.size() >= cap_inflight) break;
                if (same_group(*head, *r)) group.push_back(r);
            }
            GroupPlan plan;
            size_t bytes = 0;
            bool ok = !group.empty();
            if (ok) {
                try {
                    std::vector<ReqSpec> specs;
                    for (Request* r : group) {
                        Heart* h = hearts[r->heart_id];
                        std::lock_guard<std::mutex> hg(h->mu);
                        ensure_layout(*h, heart_mfma(ctx, *h));
                        specs.push_back(make_spec(*h, r->desc, true));
                    }
                    plan = plan_group(ctx, specs);
                    bytes = group_device_bytes(plan);
                    // shrink the group until it fits the budget
                    while (group.size() > 1 && bytes_in_use + bytes > opts.hbm_budget_bytes) {
                        group.pop_back();
                        specs.pop_back();
                        plan = plan_group(ctx, specs);
                        bytes = group_device_bytes(plan);
                    }
                    // requests of one heart side by side (in arrival order):
                    // the launch pairs neighbours of one heart
                    std::vector<size_t> ord(group.size());
                    for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
                    std::stable_sort(ord.begin(), ord.end(),
                                     [&](size_t x, size_t y) { return group[x]->heart_id < group[y]->heart_id; });
                    bool moved = false;
                    for (size_t i = 0; i < ord.size(); ++i) moved = moved || ord[i] != i;
                    if (moved) {
                        std::vector<Request*> g2;
                        std::vector<ReqSpec> s2;
                        for (size_t i : ord) {
                            g2.push_back(group[i]);
                            s2.push_back(specs[i]);
                        }
                        group.swap(g2);
                        specs.swap(s2);
                        plan = plan_group(ctx, specs);
                        bytes = group_device_bytes(plan);
                    }
                } catch (const Error& e) {
                    std::vector<Request*> one{head};
                    pending.pop_front();
                    fail_group(one, e.code, e.msg);
                    ok = false;
                    launched = true;
                }
            }
            if (ok && bytes_in_use + bytes > opts.hbm_budget_bytes) {
                if (inflight == 0) {
                    std::vector<Request*> one{head};
                    pending.pop_front();
                    fail_group(one, EKG_ENOMEM, "request does not fit the server's device budget");
                    launched = true;
                }
                ok = false;
            }
            if (ok) {
                for (Request* r : group) {
                    r->state = Request::Running;
                    for (auto it = pending.begin(); it != pending.end(); ++it)
                        if (*it == r) {
                            pending.erase(it);
                            break;
                        }
                }
                inflight += (int)group.size();
                bytes_in_use += bytes;
                peak_inflight = std::max(peak_inflight, inflight);
                peak_bytes = std::max(peak_bytes, bytes_in_use);
                free_slot->busy = true;
                free_slot->reqs = group;
                free_slot->bytes = bytes;
                free_slot->out_off.assign(1, 0);
                for (const ReqSpec& q : plan.reqs)
                    free_slot->out_off.push_back(free_slot->out_off.back() + (int64_t)q.L * q.T);
                ++groups;
                lk.unlock();
                std::string err;
                int code = EKG_OK;
                try {
                    // the group's hearts stay locked (in address order) while
                    // their layouts are read for the launch
                    std::vector<Heart*> hl;
                    for (const ReqSpec& q : plan.reqs)
                        if (std::find(hl.begin(), hl.end(), q.h) == hl.end()) hl.push_back(q.h);
                    std::sort(hl.begin(), hl.end());
                    std::vector<std::unique_lock<std::mutex>> locks;
                    for (Heart* h : hl) locks.emplace_back(h->mu);
                    Slot& s = *free_slot;
                    const size_t n_out = (size_t)s.out_off.back();
                    s.h_out = pool.take(n_out);
                    enqueue_group(ctx, plan, s.ws, s.stream, nullptr, nullptr, nullptr, false);
                    if (n_out)
                        hip_check(hipMemcpyAsync(s.h_out->p, s.ws.ecg.p, n_out * 8, hipMemcpyDeviceToHost,
                                                 s.stream),
                                  "download ECGs");
                    hip_check(hipEventRecord(s.done, s.stream), "hipEventRecord");
                } catch (const Error& e) {
                    code = e.code;
                    err = e.msg;
                }
                lk.lock();
                if (code != EKG_OK) {
                    fail_group(free_slot->reqs, code, err);
                    inflight -= (int)free_slot->reqs.size();
                    bytes_in_use -= free_slot->bytes;
                    free_slot->reqs.clear();
                    free_slot->h_out.reset();
                    free_slot->busy = false;
                    free_slot->bytes = 0;
                }
                launched = true;
            }
        }
        if (!launched) {
            // while groups run, their completion events are polled every
            // 10 us (a group's results and its slot are freed that much
            // sooner: same-heart serving 121k ECGs/s against 117k-118k at 50 us)
            if (any_busy)
                cv_work.wait_for(lk, std::chrono::microseconds(10));
            else if (!stop)
                cv_work.wait(lk, [&] { return stop || !pending.empty(); });
        }
    }
}

Server* server_create(Ctx& ctx, const ekg_server_opts& o)
{
    auto s = std::make_unique<Server>(ctx);
    s->opts = o;
    if (s->opts.hbm_budget_bytes == 0) s->opts.hbm_budget_bytes = (size_t)16 << 30;
    if (s->opts.max_inflight <= 0) s->opts.max_inflight = 1024;
    if (s->opts.max_batch <= 0) s->opts.max_batch = 256;
    if (s->opts.n_streams <= 0) s->opts.n_streams = 2;
    if (s->opts.n_streams > 8) fail(EKG_EINVAL, "at most 8 streams");
    if (s->opts.batch_window_us == 0) s->opts.batch_window_us = 1000;
    ctx.bind();
    // a server's requests may use any lead-slot width: resolve every kernel
    // on this device now rather than at its first launch inside a request's
    // latency (once per device and process)
    {
        static std::mutex prime_mu;
        static std::vector<int> primed;
        int dev = -1;
        if (hipGetDevice(&dev) == hipSuccess) {
            std::lock_guard<std::mutex> g(prime_mu);
            if (std::find(primed.begin(), primed.end(), dev) == primed.end()) {
                ekgk::prime_kernels();
                primed.push_back(dev);
            }
        }
    }
    for (int i = 0; i < s->opts.n_streams; ++i) {
        s->slots.push_back(std::make_unique<Slot>());
        Slot& sl = *s->slots.back();
        hip_check(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "hipEventCreate");
        sl.ws.ensure_side();
    }
    Server* raw = s.release();
    raw->worker = std::thread([raw] { raw->run(); });
    return raw;
}

void server_destroy(Server* s)
{
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->ctx.device);
    for (auto& sp : s->slots) {
        Slot& sl = *sp;
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        sl.h_out.reset();
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    s->slots.clear();
    s->reqs.clear();  // results never waited for: their buffers return to the pool
    delete s;
}

int server_add_heart(Server& s, Heart& h)
{
    if (h.ctx != &s.ctx) fail(EKG_EINVAL, "heart belongs to another context");
    std::lock_guard<std::mutex> g(s.mu);
    {
        std::lock_guard<std::mutex> hg(h.mu);
        const bool mfma = heart_mfma(s.ctx, h);
        ensure_layout(h, mfma);
        // class cells of every neighbourhood now, not in a request's latency
        if (mfma)
            for (int nh = EKG_NHOOD_2D4; nh <= EKG_NHOOD_3D8; ++nh) ensure_cells(h, nh);
    }
    s.hearts.push_back(&h);
    return (int)s.hearts.size() - 1;
}

uint64_t server_submit(Server& s, int heart_id, const ekg_sim_desc& d)
{
    auto r = std::make_unique<Request>();
    {
        std::lock_guard<std::mutex> g(s.mu);
        if (heart_id < 0 || heart_id >= (int)s.hearts.size())
            fail(EKG_ENOTFOUND, "unknown heart id " + std::to_string(heart_id));
        if (s.stop) fail(EKG_ESTATE, "server is stopping");
    }
    if (!d.layer_k || !d.lead_pos || d.n_layers < 1 || d.n_leads < 1 || d.n_leads > EKG_MAX_LEADS)
        fail(EKG_EINVAL, "invalid simulation descriptor");
    r->heart_id = heart_id;
    r->layer_k.assign(d.layer_k, d.layer_k + 9 * d.n_layers);
    r->lead_pos.assign(d.lead_pos, d.lead_pos + 3 * d.n_leads);
    r->desc = d;
    r->desc.layer_k = r->layer_k.data();
    r->desc.lead_pos = r->lead_pos.data();
    r->T = num_steps(d.length, d.dt);
    if (r->T < 0) fail(EKG_EINVAL, "too many time steps");
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> g(s.mu);
    r->ticket = s.next_ticket++;
    const uint64_t t = r->ticket;
    s.pending.push_back(r.get());
    s.reqs.emplace(t, std::move(r));
    s.cv_work.notify_one();
    return t;
}

bool server_poll(Server& s, uint64_t ticket)
{
    std::lock_guard<std::mutex> g(s.mu);
    auto it = s.reqs.find(ticket);
    if (it == s.reqs.end()) fail(EKG_ENOTFOUND, "unknown ticket");
    return it->second->state == Request::Done || it->second->state == Request::Failed;
}

void server_wait(Server& s, uint64_t ticket, double* ecg, size_t cap, double* latency_ms)
{
    std::unique_ptr<Request> own;
    {
        std::unique_lock<std::mutex> lk(s.mu);
        auto it = s.reqs.find(ticket);
        if (it == s.reqs.end()) fail(EKG_ENOTFOUND, "unknown ticket");
        Request* r = it->second.get();
        if (r->viewed) fail(EKG_ESTATE, "ticket already viewed; release it");
        s.cv_done.wait(lk, [&] { return r->state == Request::Done || r->state == Request::Failed; });
        own = std::move(it->second);
        s.reqs.erase(it);
    }
    // the copy runs without the server lock (the worker keeps admitting)
    Request* r = own.get();
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (r->state == Request::Failed) fail(r->err, r->msg);
    if (ecg) {
        if (cap < r->n_out)
            fail(EKG_EINVAL, "output buffer holds " + std::to_string(cap) + " doubles, " +
                                 std::to_string(r->n_out) + " needed");
        if (r->n_out) std::memcpy(ecg, r->out->p + r->out_off, r->n_out * sizeof(double));
    }
}

const double* server_wait_view(Server& s, uint64_t ticket, size_t* n, double* latency_ms)
{
    std::unique_lock<std::mutex> lk(s.mu);
    auto it = s.reqs.find(ticket);
    if (it == s.reqs.end()) fail(EKG_ENOTFOUND, "unknown ticket");
    Request* r = it->second.get();
    if (r->viewed) fail(EKG_ESTATE, "ticket already viewed; release it");
    s.cv_done.wait(lk, [&] { return r->state == Request::Done || r->state == Request::Failed; });
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (r->state == Request::Failed) {
        std::unique_ptr<Request> own = std::move(it->second);
        s.reqs.erase(it);
        lk.unlock();
        fail(own->err, own->msg);
    }
    // the request (and with it the pinned buffer of its group) stays until
    // server_release
    r->viewed = true;
    if (n) *n = r->n_out;
    return r->n_out ? r->out->p + r->out_off : nullptr;
}

void server_release(Server& s, uint64_t ticket)
{
    std::unique_ptr<Request> own;
    std::lock_guard<std::mutex> g(s.mu);
    auto it = s.reqs.find(ticket);
    if (it == s.reqs.end()) fail(EKG_ENOTFOUND, "unknown ticket");
    if (!it->second->viewed) fail(EKG_ESTATE, "ticket not viewed: use ekg_server_wait");
    own = std::move(it->second);
    s.reqs.erase(it);
}

void server_stats(Server& s, uint64_t* completed, uint64_t* groups, int* peak_inflight,
                  size_t* peak_bytes)
{
    std::lock_guard<std::mutex> g(s.mu);
    if (completed) *completed = s.completed;
    if (groups) *groups = s.groups;
    if (peak_inflight) *peak_inflight = s.peak_inflight;
    if (peak_bytes) *peak_bytes = s.peak_bytes;
}

}  // namespace ekg
