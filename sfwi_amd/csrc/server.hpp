// Serving loop over the per-shot solves (host code, included by sfwi.hip).
//
// Requests (forward, Born forward, gradient, Born gradient solves of any grid
// and record length, each with 1..max_batch shots) enter a FIFO queue.  A pool
// of worker threads, each driving its own context and stream, takes the
// first queued request whose device memory fits the free part of the HBM
// budget, so requests that arrive while others run are admitted as workers
// and memory free up, and their kernels overlap on the GPU.  Requests of the
// same shape and kind (every sfwi_params field equal) that are queued when a
// worker takes one join its batch, up to max_batch shots in all: they share
// every launch, each shot with its own request's model and wavelet, and
// each request's images are merged over its own shots (solve_merged), so the
// results equal the requests solved one by one.  Contexts are
// cached by shape (every sfwi_params field) and request kind and stay
// resident with their wavefield slots; an idle cached context is evicted
// (least recently used first) when a new request needs its memory.  A request's result does not depend on what else runs:
// each solve uses its own context and the batched kernels are deterministic.
#pragma once

#include <chrono>
#include <condition_variable>
#include <algorithm>
#include <deque>
#include <list>
#include <mutex>
#include <string>
#include <thread>

namespace {

size_t solve_bytes(const sfwi_params &p, int kind, int nshots) {
    const size_t f = sizeof(float);
    const size_t ncell = (size_t)(p.nx + 2 * p.pml) * (p.nz + 2 * p.pml);
    const size_t nmodel = (size_t)p.allnx * p.allnz, nrec = (size_t)p.ntr * p.nt, nint = (size_t)p.nx * p.nz;
    const int n = std::min(std::max(nshots, 1), std::max(p.max_batch, 1));
    const bool born = kind == SFWI_REQ_BORN_FORWARD || kind == SFWI_REQ_BORN_GRADIENT;
    const bool grad = kind == SFWI_REQ_GRADIENT || kind == SFWI_REQ_BORN_GRADIENT;
    const size_t sf = 16 * (size_t)(p.nx + p.nz);
    size_t fixed = p.nt * f + 12 * nmodel * f + 2 * nrec * f + p.max_batch * 1024 + 2 * p.ntr * 8 + 4096;
    // per shot: the slot, its window and, in a merged batch, its request's model
    // (a snapshot context's slot also holds the fused adjoint step's second buffer, 9 fields)
    size_t slot = (43 + 2 * (p.grad_rho != 0) + (p.recon == SFWI_RECON_SNAPSHOTS ? 9 : 0)) * ncell * f +
                  6 * nrec * f + 4 * nint * f + 20 * (size_t)p.ntr + 68 + 4 * (ncell / (size_t)(p.nz + 2 * p.pml)) +
                  5 * nmodel * f + p.nt * f;
    if (born) slot += 15 * ncell * f;
    const size_t chunk = p.strip_chunk > 0 ? (size_t)p.strip_chunk : 100;
    const size_t ring = p.cpu_mem && 2 * chunk < (size_t)p.nt ? 2 * chunk : (size_t)p.nt;   // strip steps in HBM
    if (grad && p.recon == SFWI_RECON_SNAPSHOTS) {   // velocity snapshots instead of strips
        const size_t k = p.snap_every > 0 ? (size_t)p.snap_every : 4, nsnap = (p.nt + k - 1) / k;
        const size_t rn = (size_t)(p.nx + 8) * (p.nz + 8);
        // partial maxima: 2 x max(256, workgroups of the fused step per shot) floats
        const size_t wc = p.step_cols > 0 ? (size_t)p.step_cols : 16;
        const size_t wg = (ncell / (size_t)(p.nz + 2 * p.pml) - 8 + wc - 1) / wc * ((p.nz + 2 * p.pml - 8 + 239) / 240);
        slot += ((nsnap * 8 + 255) & ~(size_t)255) + 8 * std::max<size_t>(256, wg) + nsnap * 2 * rn * (p.storage ? 2 : 4);
        slot += 2 * (size_t)p.nt * (ncell / (size_t)(p.nz + 2 * p.pml)) * f;   // the adjoint source table
    } else if (grad) {
        slot += p.storage ? ring * 2 * sf * 2 + ring * 8 + 2 * sf * f + 256 : ring * 2 * sf * f;
    }
    const size_t window = 12 * ncell * f;
    size_t ot = 0;
    if (p.misfit != SFWI_MISFIT_L2 && grad) {
        const size_t jobs = p.misfit == SFWI_MISFIT_OT_GATHER ? 2 * (size_t)n : 2 * (size_t)n * p.ntr;
        const size_t len = p.misfit == SFWI_MISFIT_OT_GATHER ? nrec : (size_t)p.nt;
        ot = std::min<size_t>(jobs, 1024) * 4 * (len + 1) * 8 + jobs * 48;
    }
    return fixed + n * (slot + window) + ot;
}

bool same_params(const sfwi_params &a, const sfwi_params &b) { return std::memcmp(&a, &b, sizeof a) == 0; }

double ms_since(std::chrono::steady_clock::time_point t0, std::chrono::steady_clock::time_point t1) {
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

}  // namespace

struct sfwi_server {
    struct Job {
        long long ticket;
        sfwi_request req;
        std::vector<sfwi_shot> shots;
        size_t bytes;
        std::chrono::steady_clock::time_point t_submit, t_start, t_end;
        int status = 1;   // 1 queued, 2 running, 0 done, < 0 error code
        std::string err;
    };
    struct Cached {
        sfwi_params p;
        int kind;
        sfwi_ctx *ctx;
        size_t bytes;
        bool busy;
        long long last_use;
    };
    size_t budget = 0, in_use = 0;   // in_use: running solves + cached idle contexts
    int device = 0;
    bool stop = false;
    bool batching = true;   // queued requests of one shape and kind join a worker's batch
    long long next_ticket = 1, use_clock = 0;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<long long> queue;
    std::map<long long, Job> jobs;
    std::list<Cached> cache;
    std::vector<std::thread> workers;

    // Context of shape p for a job of `bytes` (lock held).  Reuses an idle
    // cached context, else evicts idle ones until the new one fits.
    bool acquire(const sfwi_params &p, int kind, size_t bytes, std::list<Cached>::iterator *out) {
        for (auto it = cache.begin(); it != cache.end(); ++it)
            if (!it->busy && it->kind == kind && same_params(it->p, p)) {
                it->busy = true;
                it->last_use = ++use_clock;
                *out = it;
                return true;
            }
        while (in_use + bytes > budget) {
            auto lru = cache.end();
            for (auto it = cache.begin(); it != cache.end(); ++it)
                if (!it->busy && (lru == cache.end() || it->last_use < lru->last_use)) lru = it;
            if (lru == cache.end()) return false;
            sfwi_ctx_destroy(lru->ctx);
            in_use -= lru->bytes;
            cache.erase(lru);
        }
        cache.push_back({p, kind, nullptr, bytes, true, ++use_clock});
        in_use += bytes;
        *out = std::prev(cache.end());
        return true;
    }

    static Kind kind_of(int req) {
        switch (req) {
        case SFWI_REQ_FORWARD: return K_FWD;
        case SFWI_REQ_BORN_FORWARD: return K_BORN_FWD;
        case SFWI_REQ_GRADIENT: return K_GRAD;
        default: return K_BORN_GRAD;
        }
    }

    // Several requests of one shape and kind as one merged batch; bad[j]:
    // request j's first shot whose misfit is not finite, or -1.
    int run_group(const std::vector<Job *> &group, sfwi_ctx *c, bool fresh, std::vector<int> &bad) {
        const sfwi_request &q0 = group[0]->req;
        int total = 0;
        for (const Job *J : group) total += J->req.nshots;
        if (fresh && q0.autotune && q0.params.mode == SFWI_MODE_FAST) {
            int r = sfwi_set_wavelet(c, q0.wavelet);
            if (!r) r = sfwi_set_model(c, q0.vp, q0.vs, q0.rho);
            if (!r) r = sfwi_autotune(c, total, 100, nullptr, nullptr, nullptr, nullptr);
            if (r) return r;
        }
        std::vector<Part> parts;
        for (const Job *J : group) {
            const sfwi_request &q = J->req;
            Part p{};
            p.vp = q.vp; p.vs = q.vs; p.rho = q.rho; p.dvp = q.dvp; p.dvs = q.dvs; p.wav = q.wavelet;
            p.nshots = q.nshots;
            p.shots = J->shots.data();
            p.syn_z = q.syn_z; p.syn_x = q.syn_x;
            p.misfits = q.misfits;
            p.gvp = q.grad_vp; p.gvs = q.grad_vs; p.gil = q.illum;
            parts.push_back(p);
        }
        return solve_merged(c, kind_of(q0.kind), parts, &bad);
    }

    int run_job(Job &J, sfwi_ctx *c, bool fresh) {
        const sfwi_request &q = J.req;
        int r = sfwi_set_wavelet(c, q.wavelet);
        if (!r) r = sfwi_set_model(c, q.vp, q.vs, q.rho);
        const bool born = q.kind == SFWI_REQ_BORN_FORWARD || q.kind == SFWI_REQ_BORN_GRADIENT;
        if (!r && born) r = sfwi_set_perturbation(c, q.dvp, q.dvs);
        if (!r && fresh && q.autotune && q.params.mode == SFWI_MODE_FAST)
            r = sfwi_autotune(c, q.nshots, 100, nullptr, nullptr, nullptr, nullptr);
        if (r) return r;
        const int n = q.nshots;
        const size_t rec = (size_t)q.params.ntr * q.params.nt;
        switch (q.kind) {
        case SFWI_REQ_FORWARD:
            return sfwi_batch_forward(c, n, J.shots.data(), q.syn_z, q.syn_x);
        case SFWI_REQ_BORN_FORWARD:
            return sfwi_batch_born_forward(c, n, J.shots.data(), q.syn_z, q.syn_x);
        case SFWI_REQ_GRADIENT:
        case SFWI_REQ_BORN_GRADIENT: {
            std::vector<double> mis(n, 0.0);
            r = q.kind == SFWI_REQ_GRADIENT ? sfwi_batch_gradient(c, n, J.shots.data(), mis.data())
                                            : sfwi_batch_born_gradient(c, n, J.shots.data(), mis.data());
            // a non-finite misfit still flushes, so the cached context starts clean
            if (!r || r == SFWI_ENUMERIC) {
                const std::string keep = r ? std::string(sfwi_last_error(c)) : std::string();
                const int f = sfwi_flush(c, q.grad_vp, q.grad_vs, q.illum);
                if (q.misfits)
                    for (int k = 0; k < n; ++k) q.misfits[k] = mis[k];
                if (f) return f;
                if (r) return fail(c, r, "%s", keep.c_str());
            }
            return r;
        }
        }
        return SFWI_EINVAL;
    }

    void worker() {
        hipSetDevice(device);
        std::unique_lock<std::mutex> lk(m);
        for (;;) {
            // first queued job whose memory fits (reusing or evicting idle contexts)
            Job *J = nullptr;
            std::list<Cached>::iterator slot;
            for (;;) {
                if (stop) return;
                for (auto it = queue.begin(); it != queue.end(); ++it) {
                    Job &cand = jobs[*it];
                    if (acquire(cand.req.params, cand.req.kind, cand.bytes, &slot)) {
                        J = &cand;
                        queue.erase(it);
                        break;
                    }
                }
                if (J) break;
                cv_work.wait(lk);
            }
            J->status = 2;
            J->t_start = std::chrono::steady_clock::now();
            // queued requests of the same shape and kind join this batch
            std::vector<Job *> group{J};
            int total = J->req.nshots;
            for (auto it = queue.begin(); batching && it != queue.end();) {
                Job &K = jobs[*it];
                if (K.req.kind == J->req.kind && K.req.autotune == J->req.autotune &&
                    same_params(K.req.params, J->req.params) && total + K.req.nshots <= J->req.params.max_batch) {
                    total += K.req.nshots;
                    K.status = 2;
                    K.t_start = J->t_start;
                    group.push_back(&K);
                    it = queue.erase(it);
                } else {
                    ++it;
                }
            }
            sfwi_ctx *c = slot->ctx;
            lk.unlock();
            int r = 0;
            std::string msg;
            const bool fresh = !c;
            if (!c) {
                r = sfwi_ctx_create(&J->req.params, &c);
                if (r) msg = sfwi_last_error(nullptr);
            }
            std::vector<int> bad(group.size(), -1);
            if (!r) {
                r = group.size() == 1 ? run_job(*J, c, fresh) : run_group(group, c, fresh, bad);
                if (r) msg = sfwi_last_error(c);
            }
            lk.lock();
            if (c && !slot->ctx) slot->ctx = c;
            slot->busy = false;
            if (!slot->ctx) {   // creation failed: release its reservation
                in_use -= slot->bytes;
                cache.erase(slot);
            }
            const auto t_end = std::chrono::steady_clock::now();
            for (size_t j = 0; j < group.size(); ++j) {
                Job *G = group[j];
                G->t_end = t_end;
                G->status = r ? r : bad[j] >= 0 ? SFWI_ENUMERIC : 0;
                G->err = msg;
                if (!r && bad[j] >= 0)
                    G->err = "shot " + std::to_string(bad[j]) +
                             ": misfit is not finite (NaN or Inf in its observed traces)";
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

size_t sfwi_solve_bytes(const sfwi_params *p, int kind, int nshots) {
    return p ? solve_bytes(*p, kind, nshots) : 0;
}

int sfwi_server_create(size_t hbm_budget, int nworkers, sfwi_server **out) {
    if (!out) return fail(nullptr, SFWI_EINVAL, "NULL argument");
    *out = nullptr;
    if (nworkers < 1 || nworkers > 16) return fail(nullptr, SFWI_EINVAL, "workers=%d outside [1,16]", nworkers);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(nullptr, SFWI_EHIP, "no HIP device");
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return fail(nullptr, SFWI_EHIP, "hipMemGetInfo failed");
    if (hbm_budget == 0) hbm_budget = fr - fr / 16;
    if (hbm_budget > fr) return fail(nullptr, SFWI_ENOMEM, "budget %zu MiB exceeds free HBM (%zu MiB)",
                                     hbm_budget >> 20, fr >> 20);
    sfwi_server *s = new sfwi_server();
    s->budget = hbm_budget;
    s->device = dev;
    for (int k = 0; k < nworkers; ++k) s->workers.emplace_back([s] { s->worker(); });
    *out = s;
    return 0;
}

int sfwi_server_batching(sfwi_server *s, int on) {
    if (!s) return fail(nullptr, SFWI_EINVAL, "NULL server");
    std::lock_guard<std::mutex> lk(s->m);
    s->batching = on != 0;
    return 0;
}

int sfwi_server_submit(sfwi_server *s, const sfwi_request *q, long long *ticket) {
    if (!s || !q || !ticket) return fail(nullptr, SFWI_EINVAL, "NULL argument");
    if (q->kind < SFWI_REQ_FORWARD || q->kind > SFWI_REQ_BORN_GRADIENT)
        return fail(nullptr, SFWI_EINVAL, "unknown request kind %d", q->kind);
    if (q->nshots < 1 || !q->shots || !q->wavelet || !q->vp || !q->vs)
        return fail(nullptr, SFWI_EINVAL, "request needs shots, wavelet and model");
    if (q->nshots > q->params.max_batch)
        return fail(nullptr, SFWI_EINVAL, "request of %d shots exceeds params.max_batch=%d", q->nshots,
                    q->params.max_batch);
    const bool born = q->kind == SFWI_REQ_BORN_FORWARD || q->kind == SFWI_REQ_BORN_GRADIENT;
    if (born && (!q->dvp || !q->dvs)) return fail(nullptr, SFWI_EINVAL, "Born request without perturbation");
    // a cached context keeps the slots of up to max_batch shots: account for all of them
    const size_t bytes = solve_bytes(q->params, q->kind, q->params.max_batch);
    {
        std::lock_guard<std::mutex> lk(s->m);
        if (bytes > s->budget)
            return fail(nullptr, SFWI_ENOMEM, "request needs %zu MiB, more than the server budget (%zu MiB)",
                        bytes >> 20, s->budget >> 20);
    }
    // a bad model, wavelet or perturbation is refused here, before it can join a batch
    if (int r = sfwi_check_model(&q->params, q->vp, q->vs, q->rho)) return r;
    if (q->params.nt < 1 || bad_value(q->wavelet, q->params.nt) >= 0)
        return fail(nullptr, SFWI_EINVAL, "the wavelet holds a value that is not finite");
    const long nm = (long)q->params.allnx * q->params.allnz;
    if (born && (bad_value(q->dvp, nm) >= 0 || bad_value(q->dvs, nm) >= 0))
        return fail(nullptr, SFWI_EINVAL, "the perturbation holds a value that is not finite");
    std::lock_guard<std::mutex> lk(s->m);
    const long long t = s->next_ticket++;
    sfwi_server::Job &J = s->jobs[t];
    J.ticket = t;
    J.req = *q;
    J.shots.assign(q->shots, q->shots + q->nshots);
    J.bytes = bytes;
    J.t_submit = std::chrono::steady_clock::now();
    s->queue.push_back(t);
    s->cv_work.notify_one();
    *ticket = t;
    return 0;
}

// Waits (at most timeout_ms when >= 0) for the request; see sfwi_server_wait.
static int server_wait(sfwi_server *s, long long ticket, double timeout_ms, double *queue_ms, double *run_ms) {
    if (!s) return fail(nullptr, SFWI_EINVAL, "NULL server");
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return fail(nullptr, SFWI_EINVAL, "unknown ticket %lld", ticket);
    auto done = [&] { return it->second.status <= 0; };
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, done);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), done)) {
        return fail(nullptr, SFWI_ETIMEOUT, "request %lld not done within %g ms", ticket, timeout_ms);
    }
    sfwi_server::Job &J = it->second;
    if (queue_ms) *queue_ms = ms_since(J.t_submit, J.t_start);
    if (run_ms) *run_ms = ms_since(J.t_start, J.t_end);
    const int r = J.status;
    if (r) fail(nullptr, r, "request %lld: %s", ticket, J.err.c_str());   // this thread's error slot
    s->jobs.erase(it);
    return r;
}

int sfwi_server_wait(sfwi_server *s, long long ticket, double *queue_ms, double *run_ms) {
    return server_wait(s, ticket, -1.0, queue_ms, run_ms);
}

int sfwi_server_wait_for(sfwi_server *s, long long ticket, double timeout_ms, double *queue_ms, double *run_ms) {
    if (!(timeout_ms >= 0)) return fail(nullptr, SFWI_EINVAL, "timeout_ms must be >= 0");
    return server_wait(s, ticket, timeout_ms, queue_ms, run_ms);
}

int sfwi_server_cancel(sfwi_server *s, long long ticket) {
    if (!s) return fail(nullptr, SFWI_EINVAL, "NULL server");
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return fail(nullptr, SFWI_EINVAL, "unknown ticket %lld", ticket);
    sfwi_server::Job &J = it->second;
    const auto q = std::find(s->queue.begin(), s->queue.end(), ticket);
    if (J.status != 1 || q == s->queue.end())
        return fail(nullptr, SFWI_ESTATE, "request %lld has already started or ended", ticket);
    s->queue.erase(q);
    J.t_start = J.t_end = std::chrono::steady_clock::now();
    J.status = SFWI_ECANCELED;
    J.err = "cancelled before it started";
    s->cv_done.notify_all();
    return 0;
}

void sfwi_server_destroy(sfwi_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    for (auto &cc : s->cache)
        if (cc.ctx) sfwi_ctx_destroy(cc.ctx);
    delete s;
}

}  // extern "C"
