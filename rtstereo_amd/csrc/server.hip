// Serving loop: continuous admission of stereo requests of mixed size and
// configuration onto a fixed number of execution slots (one HIP stream and
// one host worker each), with admission by device-memory footprint against a
// byte budget.  Requests tagged with a camera stream id reuse a pipeline
// (volumes, per-pixel buffers, pinned staging) that stays resident until
// the stream is closed; one-off requests draw pipelines from a pool keyed
// by configuration.  Work from different slots overlaps on the device.
//
// The reference has no counterpart: its driver processes one image pair per
// process run (seq/Main.cpp:47-299).
#include "rts_internal.h"

#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace {

using clk = std::chrono::steady_clock;

double now_ms(clk::time_point t0) {
    return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

bool same_desc(const rts_pipeline_desc &a, const rts_pipeline_desc &b) {
    return memcmp(&a, &b, sizeof(a)) == 0;
}

struct Resident {
    rts_pipeline_desc desc{};
    rts_pipeline *pipe = nullptr;
    uint8_t *dev = nullptr;    // device inputs (rgb x2, lab x2) and outputs (dl, dr, mask)
    uint8_t *host = nullptr;   // pinned staging of the same layout
    size_t dev_bytes = 0, footprint = 0;
    bool busy = false;
    int stream_id = -1;        // camera stream owning it, or -1 for the pool
    uint64_t last_use = 0;
};

struct Request {
    uint64_t ticket = 0;
    rts_request_desc req{};
    size_t footprint = 0;
    rts_request_stats st{};
    bool done = false, started = false;
};

// Work for the allocator thread: pipelines to free (already out of the
// resident set and the budget) and the request whose pipeline to create
// (its footprint already reserved).
struct AllocJob {
    std::vector<std::unique_ptr<Resident>> victims;
    std::shared_ptr<Request> r;
};

size_t io_bytes(const rts_pipeline_desc &d) {
    const size_t P = (size_t)d.width * d.height;
    return 2 * 3 * P + 2 * 3 * P + 3 * P;   // rgb, lab, dl/dr/mask
}

}  // namespace

struct rts_server {
    size_t budget = 0, in_use = 0;
    int nslots = 0, oneoff_limit = 0;
    bool stop = false, paused = false;
    clk::time_point t0 = clk::now();
    uint64_t next_ticket = 1, use_clock = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::map<uint64_t, std::shared_ptr<Request>> all;
    std::vector<std::unique_ptr<Resident>> residents;
    std::vector<std::shared_ptr<Request>> assigned;   // per slot
    std::vector<Resident *> assigned_res;
    std::vector<std::thread> workers;
    std::thread dispatcher;
    // Pipeline creation and release run on their own thread: hipFree waits
    // for the device's outstanding work, and the dispatcher must keep
    // feeding resident camera pipelines meanwhile.
    std::deque<AllocJob> jobs;
    int preparing = 0;   // jobs queued or running
    std::condition_variable cv_alloc;
    std::thread allocator;
    uint64_t completed = 0;
};

namespace {

// Footprint a request adds to the budget: a new pipeline and its buffers,
// unless a resident one can serve it.
size_t pipeline_footprint(const rts_pipeline_desc &d) { return rts_pipeline_bytes(&d) + io_bytes(d); }

// An idle resident pipeline that can serve `r`: the camera stream's own one,
// or one from the pool with the same configuration.
Resident *find_resident(rts_server *s, const Request &r) {
    for (auto &res : s->residents) {
        if (res->busy) continue;
        if (r.req.stream_id >= 0) {
            if (res->stream_id == r.req.stream_id && same_desc(res->desc, r.req.pipe)) return res.get();
        } else if (res->stream_id < 0 && same_desc(res->desc, r.req.pipe)) {
            return res.get();
        }
    }
    return nullptr;
}

// A camera stream's frames run one at a time, in order.
bool stream_busy(const rts_server *s, int stream_id) {
    if (stream_id < 0) return false;
    for (auto &res : s->residents)
        if (res->busy && res->stream_id == stream_id) return true;
    return false;
}

// Bytes that evicting every idle pipeline would release.
size_t idle_bytes(const rts_server *s) {
    size_t n = 0;
    for (auto &res : s->residents)
        if (!res->busy) n += res->footprint;
    return n;
}

// Takes idle pipelines out of the resident set until `need` fits: pool
// pipelines first, then idle camera pipelines (which are rebuilt on the
// stream's next frame), least recently used first within each group.  The
// caller frees them (free_resident) without holding the lock.
void take_victims(rts_server *s, size_t need, std::vector<std::unique_ptr<Resident>> &out) {
    while (s->in_use + need > s->budget) {
        auto victim = s->residents.end();
        for (auto it = s->residents.begin(); it != s->residents.end(); ++it) {
            const Resident &res = **it;
            if (res.busy) continue;
            if (victim == s->residents.end()) {
                victim = it;
                continue;
            }
            const bool pool = res.stream_id < 0, vpool = (*victim)->stream_id < 0;
            if ((pool && !vpool) || (pool == vpool && res.last_use < (*victim)->last_use)) victim = it;
        }
        if (victim == s->residents.end()) return;
        s->in_use -= (*victim)->footprint;
        out.push_back(std::move(*victim));
        s->residents.erase(victim);
    }
}

void free_resident(Resident &res) {
    rts_pipeline_destroy(res.pipe);
    (void)hipFree(res.dev);
    (void)hipHostFree(res.host);
    res.pipe = nullptr;
    res.dev = res.host = nullptr;
}

// A new pipeline with its device and pinned staging buffers (no server
// state touched).
int create_resident(const rts_pipeline_desc &d, int stream_id, std::unique_ptr<Resident> &out) {
    std::unique_ptr<Resident> res(new Resident());
    res->desc = d;
    res->stream_id = stream_id;
    res->dev_bytes = io_bytes(d);
    if (int e = rts_pipeline_create(&d, &res->pipe)) return e;
    if (hipMalloc((void **)&res->dev, res->dev_bytes) != hipSuccess ||
        hipHostMalloc((void **)&res->host, res->dev_bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        free_resident(*res);
        return rts::fail(RTS_ERR_NOMEM, "server: staging allocation of %zu bytes failed", res->dev_bytes);
    }
    res->footprint = pipeline_footprint(d);
    out = std::move(res);
    return RTS_OK;
}

void allocator_main(rts_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_alloc.wait(lk, [&] { return s->stop || !s->jobs.empty(); });
        if (s->jobs.empty()) break;   // stop with nothing left to do
        AllocJob job = std::move(s->jobs.front());
        s->jobs.pop_front();
        const bool stopping = s->stop;
        lk.unlock();
        for (auto &v : job.victims) free_resident(*v);
        job.victims.clear();
        std::unique_ptr<Resident> res;
        const int e = stopping ? RTS_ERR_CANCELLED
                               : create_resident(job.r->req.pipe, job.r->req.stream_id, res);
        lk.lock();
        s->preparing--;
        std::shared_ptr<Request> r = job.r;
        if (e != RTS_OK || s->stop) {
            s->in_use -= r->footprint;
            if (res) {
                lk.unlock();
                free_resident(*res);
                lk.lock();
            }
            r->st.status = e != RTS_OK ? e : RTS_ERR_CANCELLED;
            r->done = true;
            s->cv_done.notify_all();
        } else {
            // a camera that changed configuration: its old pipeline joins the pool
            if (res->stream_id >= 0)
                for (auto &old : s->residents)
                    if (old->stream_id == res->stream_id) old->stream_id = -1;
            s->residents.push_back(std::move(res));
            s->queue.push_front(r);   // next in line for its new pipeline
        }
        s->cv_work.notify_all();
    }
}

// One request on one slot: stage inputs, run, stage outputs.
int execute(Resident *res, const rts_request_desc &q, hipStream_t st) {
    const rts_pipeline_desc &d = res->desc;
    const size_t P = (size_t)d.width * d.height;
    uint8_t *h = res->host, *g = res->dev;
    const size_t oL = 0, oR = 3 * P, oLabL = 6 * P, oLabR = 9 * P, oDL = 12 * P, oDR = 13 * P, oM = 14 * P;
    memcpy(h + oL, q.left_rgb, 3 * P);
    memcpy(h + oR, q.right_rgb, 3 * P);
    const bool lab = q.left_lab && q.right_lab;
    if (lab) {
        memcpy(h + oLabL, q.left_lab, 3 * P);
        memcpy(h + oLabR, q.right_lab, 3 * P);
    }
    RTS_HIP(hipMemcpyAsync(g, h, (lab ? 12 : 6) * P, hipMemcpyHostToDevice, st));
    if (int e = rts_pipeline_run(res->pipe, g + oL, g + oR, nullptr, nullptr, lab ? g + oLabL : nullptr,
                                 lab ? g + oLabR : nullptr, g + oDL, g + oDR, g + oM, st))
        return e;
    RTS_HIP(hipMemcpyAsync(h + oDL, g + oDL, 3 * P, hipMemcpyDeviceToHost, st));
    RTS_HIP(hipStreamSynchronize(st));
    memcpy(q.dl, h + oDL, P);
    if (q.dr) memcpy(q.dr, h + oDR, P);
    if (q.mask) memcpy(q.mask, h + oM, P);
    return RTS_OK;
}

void worker_main(rts_server *s, int slot) {
    hipStream_t st = nullptr;
    (void)hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_work.wait(lk, [&] { return s->stop || s->assigned[slot] != nullptr; });
        if (s->stop && !s->assigned[slot]) break;
        std::shared_ptr<Request> r = s->assigned[slot];
        Resident *res = s->assigned_res[slot];
        lk.unlock();
        r->st.start_ms = now_ms(s->t0);
        int e = st ? execute(res, r->req, st) : rts::fail(RTS_ERR_HIP, "server: stream creation failed");
        const double done = now_ms(s->t0);
        lk.lock();
        r->st.done_ms = done;
        r->st.status = e;
        r->done = true;
        res->busy = false;
        res->last_use = ++s->use_clock;
        s->assigned[slot] = nullptr;
        s->assigned_res[slot] = nullptr;
        s->completed++;
        s->cv_done.notify_all();
        s->cv_work.notify_all();   // wake the dispatcher: a slot and maybe memory are free
    }
    lk.unlock();
    if (st) (void)hipStreamDestroy(st);
}

// Admission, in two passes over the queue in arrival order: first the
// frames of camera streams, then one-off requests, of which at most
// `oneoff_limit` run at a time (so frames of resident streams always find a
// slot).  A request starts when a slot is free and it either finds an idle
// resident pipeline of its configuration or its footprint fits the budget
// once idle pipelines are evicted.  Once a request of a pass is waiting for
// memory, the requests behind it in that pass may still start on resident
// pipelines but may not allocate, so a large request is not starved by a
// stream of smaller ones.  A request that needs a new pipeline reserves its
// footprint and leaves the queue; the allocator thread frees the evicted
// pipelines, builds the new one and puts the request back at the head of
// the queue (one such request at a time), so that frames of resident camera
// pipelines keep starting while device memory is released and allocated.
void dispatcher_main(rts_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stop) {
        bool progressed = false;
        if (!s->paused) {
            for (int pass = 0; pass < 2; pass++) {
                bool blocked = false;
                for (auto it = s->queue.begin(); it != s->queue.end();) {
                    int slot = -1, oneoff = 0;
                    for (int i = 0; i < s->nslots; i++) {
                        if (!s->assigned[i]) {
                            if (slot < 0) slot = i;
                        } else if (s->assigned[i]->req.stream_id < 0) {
                            oneoff++;
                        }
                    }
                    if (slot < 0) break;
                    std::shared_ptr<Request> r = *it;
                    const bool is_stream = r->req.stream_id >= 0;
                    if (is_stream != (pass == 0) || (!is_stream && oneoff >= s->oneoff_limit) ||
                        stream_busy(s, r->req.stream_id)) {
                        ++it;
                        continue;
                    }
                    Resident *res = find_resident(s, *r);
                    if (!res) {
                        // one pipeline is prepared at a time; its memory is
                        // reserved now and the request returns to the head
                        // of the queue once the allocator thread has built it
                        if (blocked || s->preparing > 0 || s->in_use - idle_bytes(s) + r->footprint > s->budget) {
                            blocked = true;
                            ++it;
                            continue;
                        }
                        AllocJob job;
                        take_victims(s, r->footprint, job.victims);
                        s->in_use += r->footprint;
                        job.r = r;
                        s->jobs.push_back(std::move(job));
                        s->preparing++;
                        it = s->queue.erase(it);
                        s->cv_alloc.notify_one();
                        continue;
                    }
                    res->busy = true;
                    r->started = true;
                    r->st.footprint = res->footprint;
                    s->assigned[slot] = r;
                    s->assigned_res[slot] = res;
                    it = s->queue.erase(it);
                    progressed = true;
                }
            }
        }
        if (progressed) s->cv_work.notify_all();
        s->cv_work.wait(lk);
    }
}

}  // namespace

extern "C" {

rts_status rts_server_create(size_t budget_bytes, int slots, rts_server **out) {
    RTS_CHECK(out != nullptr, "null out pointer");
    RTS_CHECK(slots >= 1 && slots <= 16, "slots %d must be in [1, 16]", slots);
    RTS_CHECK(budget_bytes > 0, "empty budget");
    rts_server *s = new rts_server();
    s->budget = budget_bytes;
    s->nslots = slots;
    s->oneoff_limit = slots;
    s->assigned.assign(slots, nullptr);
    s->assigned_res.assign(slots, nullptr);
    for (int i = 0; i < slots; i++) s->workers.emplace_back(worker_main, s, i);
    s->dispatcher = std::thread(dispatcher_main, s);
    s->allocator = std::thread(allocator_main, s);
    *out = s;
    return RTS_OK;
}

rts_status rts_server_destroy(rts_server *s) {
    if (!s) return RTS_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        for (auto &r : s->queue) {
            r->st.status = RTS_ERR_CANCELLED;
            r->done = true;
        }
        s->queue.clear();
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    s->cv_alloc.notify_all();
    s->dispatcher.join();
    s->allocator.join();
    for (auto &t : s->workers) t.join();
    for (auto &res : s->residents) {
        rts_pipeline_destroy(res->pipe);
        (void)hipFree(res->dev);
        (void)hipHostFree(res->host);
    }
    delete s;
    return RTS_OK;
}

rts_status rts_server_submit(rts_server *s, const rts_request_desc *req, uint64_t *ticket) {
    RTS_CHECK(s && req && ticket, "null argument");
    RTS_CHECK(req->left_rgb && req->right_rgb && req->dl, "request needs both colour images and dl");
    RTS_CHECK(!req->pipe.lr_check || (req->dr && req->mask), "lr_check needs dr and mask");
    RTS_CHECK(!req->left_lab == !req->right_lab, "give both Lab images or neither");
    if (int e = rts::check_dims(req->pipe.width, req->pipe.height, req->pipe.disp)) return e;
    auto r = std::make_shared<Request>();
    r->req = *req;
    r->footprint = pipeline_footprint(req->pipe);
    std::lock_guard<std::mutex> g(s->mu);
    if (r->footprint > s->budget)
        return rts::fail(RTS_ERR_BUDGET, "request footprint %zu bytes exceeds the budget of %zu bytes", r->footprint,
                         s->budget);
    r->ticket = s->next_ticket++;
    r->st.submit_ms = now_ms(s->t0);
    r->st.footprint = r->footprint;
    s->queue.push_back(r);
    s->all[r->ticket] = r;
    *ticket = r->ticket;
    s->cv_work.notify_all();
    return RTS_OK;
}

static rts_status finish(rts_server *s, uint64_t ticket, rts_request_stats *st, bool wait) {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    RTS_CHECK(it != s->all.end(), "unknown ticket %llu", (unsigned long long)ticket);
    std::shared_ptr<Request> r = it->second;
    if (wait) s->cv_done.wait(lk, [&] { return r->done; });
    if (!r->done) return RTS_ERR_NOTREADY;
    if (st) *st = r->st;
    s->all.erase(it);
    return r->st.status;
}

rts_status rts_server_poll(rts_server *s, uint64_t ticket, rts_request_stats *st) {
    RTS_CHECK(s != nullptr, "null server");
    return finish(s, ticket, st, false);
}

rts_status rts_server_wait(rts_server *s, uint64_t ticket, rts_request_stats *st) {
    RTS_CHECK(s != nullptr, "null server");
    return finish(s, ticket, st, true);
}

rts_status rts_server_cancel(rts_server *s, uint64_t ticket) {
    RTS_CHECK(s != nullptr, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    for (auto it = s->queue.begin(); it != s->queue.end(); ++it)
        if ((*it)->ticket == ticket) {
            (*it)->st.status = RTS_ERR_CANCELLED;
            (*it)->done = true;
            s->queue.erase(it);
            s->cv_done.notify_all();
            return RTS_OK;
        }
    return rts::fail(RTS_ERR_INVALID, "ticket %llu is not waiting (running, finished or unknown)",
                     (unsigned long long)ticket);
}

rts_status rts_server_set_oneoff_limit(rts_server *s, int limit) {
    RTS_CHECK(s != nullptr, "null server");
    {
        std::lock_guard<std::mutex> g(s->mu);
        RTS_CHECK(limit >= 1 && limit <= s->nslots, "one-off limit %d must be in [1, %d]", limit, s->nslots);
        s->oneoff_limit = limit;
    }
    s->cv_work.notify_all();
    return RTS_OK;
}

rts_status rts_server_pause(rts_server *s, int paused) {
    RTS_CHECK(s != nullptr, "null server");
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->paused = paused != 0;
    }
    s->cv_work.notify_all();
    return RTS_OK;
}

rts_status rts_server_close_stream(rts_server *s, int stream_id) {
    RTS_CHECK(s != nullptr, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    for (auto &res : s->residents)
        if (res->stream_id == stream_id && !res->busy) res->stream_id = -1;   // back to the pool
    s->cv_work.notify_all();
    return RTS_OK;
}

rts_status rts_server_counters(rts_server *s, rts_server_stats *c) {
    RTS_CHECK(s && c, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    c->budget_bytes = s->budget;
    c->in_use_bytes = s->in_use;
    c->queued = (int)s->queue.size();
    int running = 0;
    for (auto &a : s->assigned) running += a != nullptr;
    c->running = running;
    c->resident = (int)s->residents.size();
    c->completed = s->completed;
    return RTS_OK;
}

}  // extern "C"
