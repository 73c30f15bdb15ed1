// flip_server.cpp -- continuous-batching server of independent tank solves (vare_server_* in
// include/vare_flip.h).  Host C++ over the vare_flip C ABI; no counterpart in the reference.
//
// Scheduling: requests are admitted strictly in arrival order (FIFO, no overtaking) when (a) the
// request has arrived, (b) a slot is free and (c) its reserved bytes fit in what is left of the HBM
// budget.  Each slot is a worker thread that owns one vare_flip handle (its own HIP stream) for the
// lifetime of the request, so all resident tanks step concurrently on the device.  A finished tank
// releases its slot and budget immediately and the next arrived request is admitted.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/vare_flip.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    vare_tank_request req;
    std::vector<float> pos, vel;
    int64_t id = 0;
    size_t bytes = 0;
    std::vector<float> out_pos, out_vel;
    std::vector<double> step_ms;
    vare_request_stats st;
    bool done = false;
};

size_t request_bytes(const vare_tank_request& r) {
    const int64_t B = r.grid.block_size;
    if (B <= 0) return 0;
    const int64_t fx = (r.grid.nx + B - 1) / B, fy = (r.grid.ny + B - 1) / B, fz = (r.grid.nz + B - 1) / B;
    const int64_t nblk = fx * fy * fz;
    // particles: 9 float arrays + removal/binning scratch; cells: 14 float fields, FP64 pressure,
    // descriptors and multigrid levels; frame tables; fixed slack for scan tiles and reductions
    const int64_t per_particle = 64, per_cell = 84, per_block = 64;
    return (size_t)(r.n * per_particle + nblk * B * B * B * per_cell + nblk * per_block + (4 << 20));
}

}  // namespace

struct vare_server {
    size_t budget = 0;
    int max_slots = 1;
    std::vector<std::unique_ptr<Request>> reqs;
    std::mutex mu;
    std::condition_variable cv;
    size_t used = 0;
    size_t peak = 0;
    int resident = 0;
    int max_resident = 0;
    size_t next = 0;                  // next request (in arrival order) to admit
    std::vector<size_t> order;        // request indices sorted by arrival
    Clock::time_point t0;
    double wall_s = 0.0;
};

namespace {

double since(const Clock::time_point& t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

void run_request(vare_server* s, Request* q) {
    const double admit = since(s->t0);
    q->st.admit_s = admit;
    vare_flip* h = nullptr;
    int rc = vare_flip_create(&q->req.grid, &q->req.params, q->req.n, 0, &h);
    if (rc == VARE_OK)
        rc = vare_flip_set_particles(h, q->pos.data(), q->vel.data(), nullptr, nullptr, nullptr, q->req.n,
                                     VARE_SPACE_VOXEL, VARE_MEM_HOST);
    for (int k = 0; rc == VARE_OK && k < q->req.steps; ++k) {
        const auto a = Clock::now();
        rc = vare_flip_step(h, q->req.dt);
        if (rc == VARE_ENOCONV) rc = VARE_OK;      // results are still written; counted as a step
        q->step_ms.push_back(std::chrono::duration<double, std::milli>(Clock::now() - a).count());
    }
    if (rc == VARE_OK) {
        size_t n = 0;
        vare_flip_num_particles(h, &n);
        q->out_pos.resize(3 * n);
        q->out_vel.resize(3 * n);
        rc = vare_flip_get_particles(h, q->out_pos.data(), q->out_vel.data(), nullptr, nullptr, nullptr, n, &n,
                                     VARE_SPACE_VOXEL, VARE_MEM_HOST);
        q->st.num_particles = (int64_t)n;
    }
    if (h) vare_flip_destroy(h);
    q->st.status = rc;
    q->st.steps_done = (int)q->step_ms.size();
    double sum = 0.0, mx = 0.0;
    for (double v : q->step_ms) { sum += v; mx = std::max(mx, v); }
    q->st.step_ms_mean = q->step_ms.empty() ? 0.0 : sum / q->step_ms.size();
    q->st.step_ms_max = mx;
    q->st.finish_s = since(s->t0);
}

void worker(vare_server* s) {
    for (;;) {
        Request* q = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->next >= s->order.size()) return;
                Request* head = s->reqs[s->order[s->next]].get();
                const double now = since(s->t0);
                if (head->req.arrival_s > now) {
                    s->cv.wait_for(lk, std::chrono::duration<double>(head->req.arrival_s - now));
                    continue;
                }
                if (s->used + head->bytes > s->budget) {   // FIFO: wait for memory to free up
                    s->cv.wait(lk);
                    continue;
                }
                q = head;
                s->next++;
                s->used += q->bytes;
                s->peak = std::max(s->peak, s->used);
                s->resident++;
                s->max_resident = std::max(s->max_resident, s->resident);
                break;
            }
        }
        run_request(s, q);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->used -= q->bytes;
            s->resident--;
            q->done = true;
        }
        s->cv.notify_all();
    }
}

}  // namespace

extern "C" {

int vare_server_request_bytes(const vare_tank_request* req, size_t* bytes) {
    if (!req || !bytes) return VARE_EINVAL;
    *bytes = request_bytes(*req);
    return VARE_OK;
}

int vare_server_create(size_t budget, int max_slots, vare_server** out) {
    if (!out || budget == 0 || max_slots < 1) return VARE_EINVAL;
    vare_server* s = new vare_server();
    s->budget = budget;
    s->max_slots = max_slots;
    *out = s;
    return VARE_OK;
}

int vare_server_destroy(vare_server* s) {
    delete s;
    return VARE_OK;
}

int vare_server_submit(vare_server* s, const vare_tank_request* req, int64_t* id) {
    if (!s || !req || (req->n && (!req->pos || !req->vel)) || req->steps < 0) return VARE_EINVAL;
    std::unique_ptr<Request> q(new Request());
    q->req = *req;
    q->pos.assign(req->pos, req->pos + 3 * req->n);
    q->vel.assign(req->vel, req->vel + 3 * req->n);
    q->req.pos = nullptr;
    q->req.vel = nullptr;
    q->bytes = request_bytes(*req);
    std::memset(&q->st, 0, sizeof q->st);
    q->st.arrival_s = req->arrival_s;
    q->st.bytes_reserved = (int64_t)q->bytes;
    if (q->bytes > s->budget) return VARE_EBUDGET;
    q->id = (int64_t)s->reqs.size();
    if (id) *id = q->id;
    s->reqs.push_back(std::move(q));
    return VARE_OK;
}

int vare_server_run(vare_server* s) {
    if (!s) return VARE_EINVAL;
    s->order.clear();
    for (size_t i = 0; i < s->reqs.size(); ++i)
        if (!s->reqs[i]->done) s->order.push_back(i);
    std::stable_sort(s->order.begin(), s->order.end(),
                     [&](size_t a, size_t b) { return s->reqs[a]->req.arrival_s < s->reqs[b]->req.arrival_s; });
    s->next = 0;
    s->used = 0;
    s->peak = 0;
    s->resident = 0;
    s->max_resident = 0;
    s->t0 = Clock::now();
    std::vector<std::thread> th;
    for (int k = 0; k < s->max_slots; ++k) th.emplace_back(worker, s);
    for (auto& t : th) t.join();
    s->wall_s = since(s->t0);
    for (size_t i : s->order)
        if (s->reqs[i]->st.status != VARE_OK) return s->reqs[i]->st.status;
    return VARE_OK;
}

int vare_server_result(vare_server* s, int64_t id, float* pos, float* vel, size_t cap, size_t* n,
                       vare_request_stats* st) {
    if (!s || id < 0 || (size_t)id >= s->reqs.size()) return VARE_EINVAL;
    Request* q = s->reqs[(size_t)id].get();
    const size_t np = q->out_pos.size() / 3;
    if (n) *n = np;
    if (st) *st = q->st;
    if ((pos || vel) && cap < np) return VARE_EINVAL;
    if (pos) std::memcpy(pos, q->out_pos.data(), q->out_pos.size() * sizeof(float));
    if (vel) std::memcpy(vel, q->out_vel.data(), q->out_vel.size() * sizeof(float));
    return VARE_OK;
}

int vare_server_get_stats(vare_server* s, vare_server_stats* o) {
    if (!s || !o) return VARE_EINVAL;
    std::memset(o, 0, sizeof *o);
    std::vector<double> all;
    double psteps = 0.0;
    double first = 1e300, last = 0.0;
    for (auto& q : s->reqs) {
        if (!q->done) continue;
        o->requests++;
        all.insert(all.end(), q->step_ms.begin(), q->step_ms.end());
        psteps += (double)q->req.n * (double)q->step_ms.size();
        first = std::min(first, q->st.arrival_s);
        last = std::max(last, q->st.finish_s);
    }
    o->steps = (int64_t)all.size();
    if (!all.empty()) {
        std::sort(all.begin(), all.end());
        auto pct = [&](double p) {
            const double x = p * (double)(all.size() - 1);
            const size_t i = (size_t)x;
            const double f = x - (double)i;
            return i + 1 < all.size() ? all[i] * (1.0 - f) + all[i + 1] * f : all[i];
        };
        o->step_p50_ms = pct(0.50);
        o->step_p99_ms = pct(0.99);
        double sum = 0.0;
        for (double v : all) sum += v;
        o->step_mean_ms = sum / (double)all.size();
    }
    o->wall_s = o->requests ? last - first : 0.0;
    o->particle_steps_per_s = o->wall_s > 0 ? psteps / o->wall_s : 0.0;
    o->max_resident = s->max_resident;
    o->peak_bytes = (int64_t)s->peak;
    return VARE_OK;
}

}  // extern "C"
