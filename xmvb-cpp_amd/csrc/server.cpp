// Request scheduler: continuous batching of E+grad requests over resident problem slots, with the
// dispatch loop on a native thread of its own.
//
// No reference counterpart: the reference evaluates one problem per process (cli/run.cpp). Callers
// submit requests that name a loaded slot of a pool, with host orbitals; the dispatcher thread
// launches them on the pools' launch lanes as soon as a lane is free, taking every queued request
// of that pool whose slot is not already in flight into one batched launch (up to max_batch, one
// gradient flag per launch), polls the lanes' completion events, copies energies and gradients out
// and stamps each request with its queue, device and host times. The pools' lanes run on the
// context's shared streams (at most four), so however many pools are served, the device sees a
// bounded set of queues.
#include <hip/hip_runtime.h>
#include <time.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/xmvb_amd.h"
#include "noalloc.hpp"
#include "server.hpp"

namespace {

double now_s() {
    timespec ts{};
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

struct Request {
    int64_t id = 0;
    vbe_pool* pool = nullptr;
    int slot = 0;
    bool gradient = false;
    std::vector<double> C;
    double* grad_out = nullptr;
    vbe_request_info info{};
};

struct Flight {
    vbe_pool* pool = nullptr;
    int lane = -1;
    int stream = 0;  // index into the server's streams
    std::vector<Request*> reqs;
};

}  // namespace

struct vbe_server {
    vbe_context* ctx = nullptr;
    int max_batch = 64;
    int max_lanes = 4;  // in-flight launches per pool
    std::mutex mu;
    std::condition_variable wake;  // new requests, stop
    std::condition_variable done_cv;
    std::deque<Request*> queue;
    std::vector<Flight> flights;
    std::deque<Request*> done;
    std::thread thread;
    bool stop = false;
    int64_t next_id = 1;
    int64_t launches = 0, launched_requests = 0, completed = 0;
    // the context's shared streams and the server's launches in flight on each: a launch goes to
    // the stream with the fewest, so that a short launch does not queue behind long ones of other
    // pools when another stream is idle
    std::vector<void*> streams;
    std::vector<int> stream_load;
    // per pool, the running mean of its launches' device time (ms): a pool whose launches take
    // kLongLaunch times longer than the shortest pool's is kept off stream 0 and forks nothing,
    // so that short launches always find a stream with only short work ahead of them
    static constexpr double kLongLaunch = 4.0;
    static constexpr size_t kLaunchHistory = 256;  // pools remembered
    std::vector<std::pair<vbe_pool*, double>> launch_ms;
    std::string error;  // first dispatcher failure (the request carries its status)

    void run();
    void complete(Flight& f, int rc, const std::string& err);
    bool dispatch();
    double mean_launch_ms(const vbe_pool* p) const {
        for (const auto& e : launch_ms)
            if (e.first == p) return e.second;
        return 0.0;
    }
    bool long_launches(const vbe_pool* p) const {
        double least = 0.0;
        for (const auto& e : launch_ms)
            if (least == 0.0 || e.second < least) least = e.second;
        const double mine = mean_launch_ms(p);
        return streams.size() > 1 && least > 0.0 && mine > kLongLaunch * least;
    }
};

void vbe_server::complete(Flight& f, int rc, const std::string& err) {
    const size_t n = f.reqs.size();
    std::vector<double> e(n, std::numeric_limits<double>::quiet_NaN());
    std::vector<double*> g(n, nullptr);
    std::vector<int> st(n, rc);
    float device_ms = 0.f;
    std::string cerr = err;
    if (rc == VBE_OK) {
        for (size_t i = 0; i < n; ++i) g[i] = f.reqs[i]->gradient ? f.reqs[i]->grad_out : nullptr;
        rc = xmvb_amd::pool_launch_collect(f.pool, f.lane, e.data(), g.data(), st.data(), &device_ms, &cerr);
        if (rc != VBE_OK) std::fill(st.begin(), st.end(), rc);
    }
    if (rc != VBE_OK && error.empty()) error = cerr;
    if (rc == VBE_OK && device_ms > 0.f) {
        auto it = std::find_if(launch_ms.begin(), launch_ms.end(), [&](const auto& e) { return e.first == f.pool; });
        if (it != launch_ms.end()) {
            it->second = 0.8 * it->second + 0.2 * (double)device_ms;
        } else {
            if (launch_ms.size() >= kLaunchHistory) launch_ms.erase(launch_ms.begin());  // oldest pool first
            launch_ms.emplace_back(f.pool, (double)device_ms);
        }
    }
    const double t = now_s();
    for (size_t i = 0; i < n; ++i) {
        Request* r = f.reqs[i];
        r->info.status = st[i];
        r->info.energy = e[i];
        r->info.t_complete = t;
        r->info.device_ms = device_ms;
        done.push_back(r);
        ++completed;
    }
}

// One launch per pool with queued requests and a free lane; true when anything was launched.
bool vbe_server::dispatch() {
    bool launched = false;
    // pools in the order of their oldest queued request
    std::vector<vbe_pool*> order;
    for (Request* r : queue)
        if (std::find(order.begin(), order.end(), r->pool) == order.end()) order.push_back(r->pool);
    for (vbe_pool* p : order) {
        int in_flight = 0;
        for (const auto& f : flights) in_flight += f.pool == p;
        if (in_flight >= max_lanes) continue;
        std::vector<Request*> batch;
        bool gradient = false;
        for (Request* r : queue) {
            if (r->pool != p) continue;
            if (batch.empty()) gradient = r->gradient;
            if (r->gradient != gradient) continue;
            bool dup = false;
            for (Request* b : batch) dup |= b->slot == r->slot;
            if (dup || xmvb_amd::pool_slot_busy(p, r->slot)) continue;
            batch.push_back(r);
            if ((int)batch.size() >= max_batch) break;
        }
        if (batch.empty()) continue;
        std::vector<int> slots;
        std::vector<const double*> C;
        for (Request* r : batch) {
            slots.push_back(r->slot);
            C.push_back(r->C.data());
        }
        int lane = -1;
        std::string err;
        const bool long_pool = long_launches(p);
        const int first = long_pool ? 1 : 0;
        const int si = (int)(std::min_element(stream_load.begin() + first, stream_load.end()) - stream_load.begin());
        void* side = long_pool ? nullptr : streams[(size_t)(si + 1) % streams.size()];
        const double t = now_s();  // before the enqueue: the device may start while it runs
        const int rc = xmvb_amd::pool_launch(p, (int)batch.size(), slots.data(), C.data(), gradient, streams[si],
                                             side, &lane, &err);
        if (rc == VBE_OK && lane < 0) continue;  // the caller holds the pool's lanes; try again later
        for (Request* r : batch) {
            r->info.t_dispatch = t;
            r->info.batch_size = (int32_t)batch.size();
            queue.erase(std::find(queue.begin(), queue.end(), r));
        }
        Flight f{p, lane, si, batch};
        if (rc != VBE_OK) {
            complete(f, rc, err);
            done_cv.notify_all();
            continue;
        }
        ++launches;
        launched_requests += (int64_t)batch.size();
        ++stream_load[si];
        flights.push_back(std::move(f));
        launched = true;
    }
    return launched;
}

void vbe_server::run() {
    (void)hipSetDevice(xmvb_amd::context_device(ctx));
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        bool finished = false;
        for (size_t k = 0; k < flights.size();) {
            std::string err;
            const int q = xmvb_amd::pool_launch_query(flights[k].pool, flights[k].lane, &err);
            if (q == 0) {
                ++k;
                continue;
            }
            complete(flights[k], q < 0 ? -q : VBE_OK, err);
            --stream_load[flights[k].stream];
            flights.erase(flights.begin() + (long)k);
            finished = true;
        }
        if (finished) done_cv.notify_all();
        while (!queue.empty() && dispatch()) {
        }
        if (stop && queue.empty() && flights.empty()) break;
        if (flights.empty() && queue.empty())
            wake.wait(lk, [&] { return stop || !queue.empty(); });
        else
            wake.wait_for(lk, std::chrono::microseconds(flights.empty() ? 200 : 20));
    }
    done_cv.notify_all();
}

extern "C" {

int vbe_server_create(vbe_context* ctx, int32_t max_batch, int32_t max_inflight_per_pool, vbe_server** out) {
    if (!ctx || !out) return xmvb_amd::report_error(ctx, VBE_ERR_ARGUMENT, "null argument");
    if (max_batch <= 0 || max_inflight_per_pool <= 0 || max_inflight_per_pool > 4)
        return xmvb_amd::report_error(ctx, VBE_ERR_ARGUMENT,
                                      "max_batch must be positive and max_inflight_per_pool in [1, 4]");
    try {
        auto s = std::make_unique<vbe_server>();
        s->ctx = ctx;
        s->max_batch = max_batch;
        s->max_lanes = max_inflight_per_pool;
        s->streams.resize(8);
        s->streams.resize((size_t)xmvb_amd::context_streams(ctx, s->streams.data(), 8));
        s->stream_load.assign(s->streams.size(), 0);
        vbe_server* raw = s.get();
        s->thread = std::thread([raw] { raw->run(); });
        *out = s.release();
        return VBE_OK;
    } catch (const std::exception& e) {
        return xmvb_amd::report_error(ctx, VBE_ERR_DEVICE, e.what());
    }
}

void vbe_server_destroy(vbe_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->wake.notify_all();
    if (s->thread.joinable()) s->thread.join();
    for (Request* r : s->done) delete r;
    delete s;
}

int vbe_server_submit(vbe_server* s, vbe_pool* pool, int32_t slot, const double* C, int32_t want_gradient,
                      double* grad_out, double arrival, int64_t* id) {
    if (!s || !pool || !C || !id || (want_gradient && !grad_out))
        return xmvb_amd::report_error(s ? s->ctx : nullptr, VBE_ERR_ARGUMENT, "null argument");
    int n_act = 0, n_bf = 0;
    std::string err;
    const int rc = xmvb_amd::pool_slot_dims(pool, slot, &n_act, &n_bf, &err);
    if (rc != VBE_OK) return xmvb_amd::report_error(s->ctx, rc, err.c_str());
    auto r = std::make_unique<Request>();
    r->pool = pool;
    r->slot = slot;
    r->gradient = want_gradient != 0;
    r->C.assign(C, C + (size_t)n_act * n_bf);
    r->grad_out = want_gradient ? grad_out : nullptr;
    r->info.t_submit = now_s();
    r->info.t_arrival = arrival > 0.0 ? arrival : r->info.t_submit;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return xmvb_amd::report_error(s->ctx, VBE_ERR_ARGUMENT, "server is stopping");
        r->id = s->next_id++;
        r->info.id = r->id;
        *id = r->id;
        s->queue.push_back(r.release());
    }
    s->wake.notify_one();
    return VBE_OK;
}

int vbe_server_collect(vbe_server* s, int32_t max_count, int32_t wait_us, vbe_request_info* out, int32_t* count) {
    if (!s || !out || !count || max_count <= 0)
        return xmvb_amd::report_error(s ? s->ctx : nullptr, VBE_ERR_ARGUMENT, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_us > 0)
        s->done_cv.wait_for(lk, std::chrono::microseconds(wait_us), [&] { return !s->done.empty(); });
    int n = 0;
    while (n < max_count && !s->done.empty()) {
        Request* r = s->done.front();
        s->done.pop_front();
        out[n++] = r->info;
        delete r;
    }
    *count = n;
    return VBE_OK;
}

int vbe_server_stats(vbe_server* s, vbe_server_counters* c) {
    if (!s || !c) return xmvb_amd::report_error(s ? s->ctx : nullptr, VBE_ERR_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    c->queued = (int64_t)s->queue.size();
    c->in_flight_launches = (int64_t)s->flights.size();
    c->launches = s->launches;
    c->launched_requests = s->launched_requests;
    c->completed = s->completed;
    c->uncollected = (int64_t)s->done.size();
    return VBE_OK;
}

}  // extern "C"
