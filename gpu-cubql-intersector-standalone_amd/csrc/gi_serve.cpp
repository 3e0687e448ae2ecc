// gi_serve.cpp -- the serving loop declared in include/gi_intersector.h.
//
// One primary context owns the static mesh (built once). Every slot owns one
// context per registered moving mesh; those contexts borrow the static mesh's
// device state (gi_share_mesh) and keep their own moving-mesh state, stream
// and arenas, so a request costs a transform + refit + query and requests on
// different slots overlap on the device. A worker thread per slot takes the
// oldest queued request when HBM admission allows it.
#include <hip/hip_runtime_api.h>
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/gi_intersector.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket;
    int32_t mesh_id;
    double m[12];
    Clock::time_point t_submit;
    uint32_t refusals;
};

struct Done {
    bool ready = false;
    gi_result res;
    gi_request_info info;
};

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

struct gi_server {
    gi_server_config cfg;
    gi_context* primary = nullptr;
    std::vector<std::vector<gi_context*>> slot_ctx;  // [slot][mesh_id]
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request> queue;
    std::map<uint64_t, Done> done;
    std::map<uint64_t, bool> outstanding;  // submitted, not yet collected by gi_server_wait
    uint64_t next_ticket = 1;
    int in_flight = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    int n_meshes = 0;
    Clock::time_point t0 = Clock::now();
};

namespace {

bool admit(gi_server* s) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return true;
    return free_b >= s->cfg.hbm_reserve_bytes;
}

void worker(gi_server* s, int slot) {
    (void)hipSetDevice(s->cfg.device);
    for (;;) {
        Request rq;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->stop) return;
                if (s->queue.empty()) {
                    s->cv_work.wait(lk);
                    continue;
                }
                // admission: free HBM above the reserve; a request is always
                // admitted when nothing else runs, so the queue cannot stall
                if (s->in_flight == 0 || admit(s)) break;
                s->queue.front().refusals++;
                s->cv_work.wait_for(lk, std::chrono::milliseconds(1));
            }
            rq = s->queue.front();
            s->queue.pop_front();
            s->in_flight++;
        }
        Clock::time_point t_admit = Clock::now();
        gi_result res;
        memset(&res, 0, sizeof(res));
        gi_context* ctx = s->slot_ctx[slot][rq.mesh_id];
        int st = gi_set_transform_matrix(ctx, GI_MESH_B, rq.m, s->cfg.snap_to_grid);
        if (st == GI_OK) {
            // same shapes every request: replay a graph; service times come
            // from the host clock, so no stage timing events
            gi_options o = {0, s->cfg.want_segments, 0, 1, GI_TIME_NONE, 0};
            st = gi_intersect(ctx, &o, &res);
        }
        Clock::time_point t_done = Clock::now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            Done& d = s->done[rq.ticket];
            memset(&d.info, 0, sizeof(d.info));
            d.res = res;
            d.info.queue_ms = ms_between(rq.t_submit, t_admit);
            d.info.service_ms = ms_between(t_admit, t_done);
            d.info.latency_ms = ms_between(rq.t_submit, t_done);
            d.info.t_admit_ms = ms_between(s->t0, t_admit);
            d.info.t_done_ms = ms_between(s->t0, t_done);
            d.info.slot = slot;
            d.info.refusals = rq.refusals;
            d.info.status = st;
            d.ready = true;
            s->in_flight--;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

void destroy_contexts(gi_server* s) {
    for (auto& v : s->slot_ctx)
        for (gi_context* c : v)
            if (c) gi_destroy(c);
    s->slot_ctx.clear();
    if (s->primary) gi_destroy(s->primary);
    s->primary = nullptr;
}

}  // namespace

extern "C" {

int gi_server_create(const gi_server_config* cfg, const double* xyz, uint32_t nv, const uint32_t* tri, uint32_t nt,
                     gi_server** out) {
    if (!out || !xyz || !tri || nt == 0) return GI_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    gi_server* s = new gi_server();
    memset(&s->cfg, 0, sizeof(s->cfg));
    if (cfg) s->cfg = *cfg;
    if (s->cfg.slots <= 0) s->cfg.slots = 4;
    if (s->cfg.hbm_reserve_bytes == 0) s->cfg.hbm_reserve_bytes = 2ull << 30;
    gi_config c;
    memset(&c, 0, sizeof(c));
    c.device = s->cfg.device;
    int st = gi_create(&c, &s->primary);
    if (!st) st = gi_set_mesh(s->primary, GI_MESH_A, xyz, nv, tri, nt);
    if (!st) st = gi_build(s->primary);
    if (st) {
        destroy_contexts(s);
        delete s;
        return st;
    }
    s->slot_ctx.resize((size_t)s->cfg.slots);
    for (int i = 0; i < s->cfg.slots; ++i) s->workers.emplace_back(worker, s, i);
    *out = s;
    return GI_OK;
}

int gi_server_add_mesh(gi_server* s, const double* xyz, uint32_t nv, const uint32_t* tri, uint32_t nt,
                       int32_t* mesh_id) {
    if (!s || !xyz || !tri || !mesh_id || nt == 0) return GI_ERR_INVALID_ARGUMENT;
    std::vector<gi_context*> made;
    int st = GI_OK;
    for (int i = 0; i < s->cfg.slots && st == GI_OK; ++i) {
        gi_config c;
        memset(&c, 0, sizeof(c));
        c.device = s->cfg.device;
        gi_context* ctx = nullptr;
        st = gi_create(&c, &ctx);
        if (st) break;
        made.push_back(ctx);
        st = gi_share_mesh(ctx, GI_MESH_A, s->primary, GI_MESH_A);
        if (!st) st = gi_set_mesh(ctx, GI_MESH_B, xyz, nv, tri, nt);
        if (!st) st = gi_build(ctx);
    }
    if (st) {
        for (gi_context* c : made) gi_destroy(c);
        return st;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    for (int i = 0; i < s->cfg.slots; ++i) s->slot_ctx[(size_t)i].push_back(made[(size_t)i]);
    *mesh_id = s->n_meshes++;
    return GI_OK;
}

int gi_server_submit(gi_server* s, int32_t mesh_id, const double m[12], uint64_t* ticket) {
    if (!s || !m || !ticket) return GI_ERR_INVALID_ARGUMENT;
    Request rq;
    rq.mesh_id = mesh_id;
    memcpy(rq.m, m, sizeof(rq.m));
    rq.t_submit = Clock::now();
    rq.refusals = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (mesh_id < 0 || mesh_id >= s->n_meshes) return GI_ERR_INVALID_ARGUMENT;
        rq.ticket = s->next_ticket++;
        s->queue.push_back(rq);
        s->outstanding[rq.ticket] = true;
        *ticket = rq.ticket;
    }
    s->cv_work.notify_one();
    return GI_OK;
}

int gi_server_wait(gi_server* s, uint64_t ticket, gi_result* out, gi_request_info* info) {
    if (!s || !out) return GI_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->outstanding.find(ticket) == s->outstanding.end()) return GI_ERR_INVALID_ARGUMENT;
    s->cv_done.wait(lk, [&] {
        auto j = s->done.find(ticket);
        return j != s->done.end() && j->second.ready;
    });
    Done d = s->done[ticket];
    s->done.erase(ticket);
    s->outstanding.erase(ticket);
    *out = d.res;
    if (info) *info = d.info;
    return d.info.status;
}

int gi_server_destroy(gi_server* s) {
    if (!s) return GI_ERR_INVALID_ARGUMENT;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    for (auto& kv : s->done)
        if (kv.second.ready) gi_result_free(&kv.second.res);
    destroy_contexts(s);
    delete s;
    return GI_OK;
}

}  // extern "C"
