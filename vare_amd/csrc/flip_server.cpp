// flip_server.cpp -- continuous-batching server of independent tank solves (vare_server_* in
// include/vare_flip.h).  Host C++ over the vare_flip C ABI; no counterpart in the reference.
//
// Scheduling: requests are admitted strictly in arrival order (FIFO, no overtaking) when (a) the
// request has arrived, (b) a slot is free and (c) its reserved bytes fit in what is left of the HBM
// budget.  Each slot is a worker thread that owns one vare_flip handle for the lifetime of the request:
// it creates (or takes from the pool) the handle, uploads the particles, hands the tank to the stepper
// and, when its steps are done, downloads the result and releases the slot, so the next arrived
// request is admitted at once.
//
// Stepping (batched, the default): one stepper thread advances every resident tank by one step per
// round with vare_flip_step_batched, so each kernel of the step is launched once for all of them; a
// tank joins at the round after its upload and leaves after its last step (iteration-level continuous
// batching).  A tank's step latency is the round's wall time.  With batching off
// (vare_server_set_batching), each worker steps its own tank on its handle's stream instead, and the
// tanks overlap only as concurrent streams.
//
// Handle pool: a finished tank's handle (device state sized for its grid) is kept, not destroyed, and
// the next request with the same grid and velocity storage takes it over, so serving does not
// allocate or free device memory per request (hipFree waits for the whole device, which would stall
// every other resident tank).  Pooled handles keep their live bytes counted against the budget; when
// the head request does not fit and destroying pooled handles alone makes room, they are destroyed,
// least recently used first, outside the server lock (their bytes stay counted until they are gone).
//
// Accounting: a request is charged the bytes vare_flip_plan_bytes reports for its grid, particle
// count and parameters (the exact arrays the library allocates for create and the first steps, plus
// the staging of the particle upload and download).  A pooled handle serves only requests whose plan
// is within its footprint bound (the largest plan it has served), so it never grows, and such a
// request is charged that bound.
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
    size_t bytes = 0;                 // planned bytes of the request (vare_flip_plan_bytes)
    size_t charge = 0;                // bytes counted against the budget while it runs
    std::vector<float> out_pos, out_vel;
    std::vector<double> step_ms;
    vare_request_stats st;
    bool done = false;
    // batched stepping: the stepper owns the tank from hand-over until stepped == true
    vare_flip* h = nullptr;
    int steps_taken = 0;
    int step_rc = VARE_OK;
    bool stepped = false;
};

// Device bytes a request reserves: what vare_flip_create and the first steps of its tank allocate,
// plus the staging of the particle upload and download (vare_flip_plan_bytes counts exactly the
// arrays the library allocates).  0 when the request is invalid.
size_t request_bytes(const vare_tank_request& r) {
    size_t b = 0;
    if (vare_flip_plan_bytes(&r.grid, &r.params, r.n, 0, &b) != VARE_OK) return 0;
    return b;
}

}  // namespace

struct Pooled {
    vare_flip* h = nullptr;
    vare_grid_desc grid;
    int fp16 = 0;
    size_t bytes = 0;                 // live device bytes of the idle handle
    size_t plan = 0;                  // the largest planned bytes of a request it served: its footprint bound
    uint64_t last_use = 0;
};

bool same_grid(const vare_grid_desc& a, const vare_grid_desc& b) {
    return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.block_size == b.block_size && a.min[0] == b.min[0] &&
           a.min[1] == b.min[1] && a.min[2] == b.min[2] && a.max[0] == b.max[0] && a.max[1] == b.max[1] &&
           a.max[2] == b.max[2];
}

struct vare_server {
    size_t budget = 0;
    int max_slots = 1;
    std::vector<Pooled> pool;         // idle handles, their bytes included in `used`
    uint64_t use_clock = 0;
    int64_t handles_created = 0;
    int64_t handles_reused = 0;
    std::vector<std::unique_ptr<Request>> reqs;
    std::mutex mu;
    std::condition_variable cv;
    size_t used = 0;                  // resident charges + pooled handles + handles being destroyed
    size_t pending = 0;               // bytes of handles being destroyed
    size_t peak = 0;
    int resident = 0;
    int max_resident = 0;
    size_t next = 0;                  // next request (in arrival order) to admit
    std::vector<size_t> order;        // request indices sorted by arrival
    Clock::time_point t0;
    double wall_s = 0.0;
    // batched stepping
    int batching = 1;
    std::vector<Request*> ready;      // uploaded tanks waiting to join the next round
    std::condition_variable cv_step;
    bool stop_stepper = false;
    int64_t rounds = 0;
    int max_batch = 0;
};

namespace {

double since(const Clock::time_point& t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

// Index of a pooled handle that can serve this request without growing, or -1: the same grid and
// velocity storage, and a request whose planned bytes are within the handle's footprint bound (a
// handle's device bytes never exceed the plan of the largest request it has served, so such a request
// allocates nothing new).  Called with s->mu held.
long find_pooled(const vare_server* s, const Request* q) {
    const int fp16 = q->req.params.velocity_fp16 ? 1 : 0;
    for (size_t i = 0; i < s->pool.size(); ++i)
        if (s->pool[i].fp16 == fp16 && same_grid(s->pool[i].grid, q->req.grid) && q->bytes <= s->pool[i].plan)
            return (long)i;
    return -1;
}

// Runs the request on h (a pooled handle) or on a new handle; returns the handle to pool afterwards
// (nullptr when it failed and was destroyed).
vare_flip* run_request(vare_server* s, Request* q, vare_flip* h) {
    int rc = VARE_OK;
    if (h) rc = vare_flip_set_params(h, &q->req.params);
    else rc = vare_flip_create(&q->req.grid, &q->req.params, q->req.n, 0, &h);
    if (rc == VARE_OK)
        rc = vare_flip_set_particles(h, q->pos.data(), q->vel.data(), nullptr, nullptr, nullptr, q->req.n,
                                     VARE_SPACE_VOXEL, VARE_MEM_HOST);
    if (rc == VARE_OK && s->batching && q->req.steps > 0) {
        // hand the tank to the stepper and wait until it has taken all of its steps
        std::unique_lock<std::mutex> lk(s->mu);
        q->h = h;
        s->ready.push_back(q);
        s->cv_step.notify_all();
        s->cv.wait(lk, [&] { return q->stepped; });
        rc = q->step_rc;
    }
    for (int k = 0; rc == VARE_OK && !s->batching && k < q->req.steps; ++k) {
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
    if (h && rc != VARE_OK) {
        vare_flip_destroy(h);
        h = nullptr;
    }
    q->st.status = rc;
    q->st.steps_done = (int)q->step_ms.size();
    double sum = 0.0, mx = 0.0;
    for (double v : q->step_ms) { sum += v; mx = std::max(mx, v); }
    q->st.step_ms_mean = q->step_ms.empty() ? 0.0 : sum / q->step_ms.size();
    q->st.step_ms_max = mx;
    q->st.finish_s = since(s->t0);
    return h;
}

void worker(vare_server* s) {
    for (;;) {
        Request* q = nullptr;
        vare_flip* h = nullptr;
        std::vector<vare_flip*> victims;          // pooled handles to destroy outside the lock
        size_t freed = 0;                         // their bytes
        size_t q_plan = 0;                        // footprint bound of the handle the request runs on
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (!victims.empty()) {
                    // destroy evicted handles without holding the lock (hipFree waits for the
                    // device); their bytes stay counted until they are gone
                    lk.unlock();
                    for (vare_flip* v : victims) vare_flip_destroy(v);
                    lk.lock();
                    s->used -= std::min(freed, s->used);
                    s->pending -= std::min(freed, s->pending);
                    victims.clear();
                    freed = 0;
                    s->cv.notify_all();
                }
                if (s->next >= s->order.size()) return;
                Request* head = s->reqs[s->order[s->next]].get();
                const double now = since(s->t0);
                if (head->req.arrival_s > now) {
                    s->cv.wait_for(lk, std::chrono::duration<double>(head->req.arrival_s - now));
                    continue;
                }
                const long m = find_pooled(s, head);
                const size_t mine = m >= 0 ? s->pool[(size_t)m].bytes : 0;        // already in `used`
                const size_t need = m >= 0 ? s->pool[(size_t)m].plan : head->bytes;
                const size_t base = s->used - mine;
                if (base + need > s->budget) {
                    // evict only when destroying pooled handles alone makes room (FIFO: otherwise
                    // wait for resident tanks to finish)
                    size_t evictable = 0;
                    for (size_t i = 0; i < s->pool.size(); ++i)
                        if ((long)i != m) evictable += s->pool[i].bytes;
                    if (base - std::min(evictable, base) + need > s->budget) {
                        s->cv.wait(lk);
                        continue;
                    }
                    size_t room = base;
                    while (room + need > s->budget) {
                        size_t lru = s->pool.size();
                        for (size_t i = 0; i < s->pool.size(); ++i) {
                            if (find_pooled(s, head) == (long)i) continue;
                            if (lru == s->pool.size() || s->pool[i].last_use < s->pool[lru].last_use) lru = i;
                        }
                        if (lru == s->pool.size()) break;
                        room -= s->pool[lru].bytes;
                        freed += s->pool[lru].bytes;
                        s->pending += s->pool[lru].bytes;
                        victims.push_back(s->pool[lru].h);
                        s->pool.erase(s->pool.begin() + (long)lru);
                    }
                    continue;
                }
                const long m2 = find_pooled(s, head);
                if (m2 >= 0) {
                    h = s->pool[(size_t)m2].h;
                    q_plan = s->pool[(size_t)m2].plan;
                    s->pool.erase(s->pool.begin() + m2);
                    s->handles_reused++;
                } else {
                    s->handles_created++;
                }
                q = head;
                q->st.admit_s = since(s->t0);     // stamped under the lock, so in admission order
                q->charge = need;
                s->next++;
                s->used = base + need;
                s->peak = std::max(s->peak, s->used);
                s->resident++;
                s->max_resident = std::max(s->max_resident, s->resident);
                break;
            }
        }
        h = run_request(s, q, h);
        vare_flip* drop = nullptr;
        size_t live = 0;
        if (h) vare_flip_get_device_bytes(h, &live, nullptr);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->used -= q->charge;
            if (h && (int)s->pool.size() < s->max_slots) {      // the handle's live bytes pass to the pool
                Pooled pd;
                pd.h = h;
                pd.grid = q->req.grid;
                pd.fp16 = q->req.params.velocity_fp16 ? 1 : 0;
                pd.bytes = live;
                pd.plan = std::max(q_plan, q->bytes);
                pd.last_use = ++s->use_clock;
                s->pool.push_back(pd);
                s->used += live;
                s->peak = std::max(s->peak, s->used);
            } else if (h) {
                drop = h;                                     // destroyed below, its bytes counted until then
                s->used += live;
                s->pending += live;
            }
            s->resident--;
            q->done = true;
        }
        if (drop) {
            vare_flip_destroy(drop);
            std::lock_guard<std::mutex> lk(s->mu);
            s->used -= std::min(live, s->used);
            s->pending -= std::min(live, s->pending);
        }
        s->cv.notify_all();
    }
}

// Batched stepping: one round steps every resident tank once with shared launches
// (vare_flip_step_batched); tanks handed over meanwhile join the next round, finished ones leave.
void stepper(vare_server* s) {
    std::vector<Request*> batch;
    std::vector<vare_flip*> hs;
    std::vector<float> dts;
    std::vector<int> st;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_step.wait(lk, [&] { return !s->ready.empty() || !batch.empty() || s->stop_stepper; });
            if (batch.empty() && s->ready.empty()) return;      // stop_stepper: every request is done
            batch.insert(batch.end(), s->ready.begin(), s->ready.end());
            s->ready.clear();
            s->max_batch = std::max(s->max_batch, (int)batch.size());
        }
        hs.clear();
        dts.clear();
        for (Request* q : batch) {
            hs.push_back(q->h);
            dts.push_back(q->req.dt);
        }
        st.assign(batch.size(), VARE_OK);
        const auto a = Clock::now();
        (void)vare_flip_step_batched(hs.data(), (int)hs.size(), dts.data(), st.data());
        const double ms = std::chrono::duration<double, std::milli>(Clock::now() - a).count();
        std::lock_guard<std::mutex> lk(s->mu);
        s->rounds++;
        size_t keep = 0;
        for (size_t i = 0; i < batch.size(); ++i) {
            Request* q = batch[i];
            int rc = st[i];
            if (rc == VARE_ENOCONV) rc = VARE_OK;      // results are still written; counted as a step
            q->step_ms.push_back(ms);
            q->steps_taken++;
            if (rc != VARE_OK || q->steps_taken >= q->req.steps) {
                q->step_rc = rc;
                q->stepped = true;
            } else {
                batch[keep++] = q;
            }
        }
        batch.resize(keep);
        s->cv.notify_all();
    }
}

}  // namespace

extern "C" {

int vare_server_set_batching(vare_server* s, int on) {
    if (!s) return VARE_EINVAL;
    s->batching = on ? 1 : 0;
    return VARE_OK;
}

int vare_server_request_bytes(const vare_tank_request* req, size_t* bytes) {
    if (!req || !bytes) return VARE_EINVAL;
    *bytes = request_bytes(*req);
    return *bytes ? VARE_OK : VARE_EINVAL;
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
    if (!s) return VARE_OK;
    for (auto& pd : s->pool) vare_flip_destroy(pd.h);
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
    if (q->bytes == 0) return VARE_EINVAL;
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
    s->handles_created = 0;
    s->handles_reused = 0;
    s->used = 0;
    s->pending = 0;
    for (const Pooled& pd : s->pool) s->used += pd.bytes;
    s->peak = s->used;
    s->resident = 0;
    s->max_resident = 0;
    s->t0 = Clock::now();
    s->ready.clear();
    s->stop_stepper = false;
    s->rounds = 0;
    s->max_batch = 0;
    std::thread step_thread;
    if (s->batching) step_thread = std::thread(stepper, s);
    std::vector<std::thread> th;
    for (int k = 0; k < s->max_slots; ++k) th.emplace_back(worker, s);
    for (auto& t : th) t.join();
    if (s->batching) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop_stepper = true;
        }
        s->cv_step.notify_all();
        step_thread.join();
    }
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
    o->handles_created = s->handles_created;
    o->handles_reused = s->handles_reused;
    o->rounds = s->rounds;
    o->max_batch = s->max_batch;
    return VARE_OK;
}

}  // extern "C"
