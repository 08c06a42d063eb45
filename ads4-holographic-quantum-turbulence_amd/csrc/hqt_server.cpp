// Serving loop: a host scheduler thread that keeps evolutions device-resident and batches work
// as requests arrive (continuous batching).
//
//   - Evolution segments (a context, its initial state and a step count) are admitted in
//     arrival order while their device footprint fits the HBM budget; all admitted
//     evolutions advance together as one ensemble (hqt_ensemble_*), quantum steps at a time.
//     When one finishes (or trips the NaN guard) its state is downloaded, its memory freed and
//     the ensemble is rebuilt over the remaining members plus the newly admitted ones.
//   - Stationary Newton solves (NewtonRaphson_vy.m) that are pending at a scheduling point
//     are solved as one batch (one work-group per problem, any mix of radial sizes) on a second
//     stream, while the ensemble's queued steps run on the first, so the two kinds of work
//     overlap on the GPU.
// Each request's result is the one it would get alone: ensemble members and Newton problems
// are computed independently of their neighbours.
#include "hqt_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <tuple>

namespace {

using Clock = std::chrono::steady_clock;

enum ReqKind { RK_EVO = 0, RK_NEWTON = 1 };

struct Request {
    int64_t id = 0;
    int kind = RK_EVO;
    Clock::time_point t_submit, t_done;
    bool done = false;
    int status = HQT_OK;
    std::string msg;
    // evolution
    hqt_params p{};
    std::vector<double> axb, ayb;                  // copies of the boundary arrays (if any)
    std::vector<double> Z1, Z2, Z3, Z4, rho;       // initial state, then the result
    int32_t nsteps = 0, steps_done = 0, nan_step = 0;
    int64_t bytes = 0;
    hqt_ctx* ctx = nullptr;
    // Newton
    int32_t nz = 0;
    double zh = 1, small = 0, D = 0, vy = 0;
    double mu_seed = std::nan(""), dmu = 0;        // mu-continuation seed and step (NaN: one solve)
    std::vector<double> n1, n2, nv;                // initial guesses, then results
    int32_t it = 0;
    double err = 0;
};

}  // namespace

struct hqt_server {
    int64_t budget = 0;
    int32_t quantum = 25;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    int64_t next_id = 1;
    std::map<int64_t, std::shared_ptr<Request>> reqs;          // until the result is taken
    std::deque<std::shared_ptr<Request>> pend_evo, pend_newton;
    std::vector<double> lat[2];                                 // completed latencies (s)
    // admitted, done evo, done newton, ensemble builds, newton batches, peak bytes, peak members, steps,
    // Newton solves (continuation steps included), continued Newton requests
    int64_t counters[HQT_SERVER_NCOUNTERS] = {};
    std::thread worker;
    // worker-thread state
    std::vector<std::shared_ptr<Request>> active;
    int64_t used = 0;
    hqt_ensemble* ens = nullptr;
    hipStream_t nstream = nullptr;
    NewtonWork nwork;
};

namespace {

double seconds(Clock::duration d) { return std::chrono::duration<double>(d).count(); }

void complete(hqt_server* s, const std::shared_ptr<Request>& r, int status, const std::string& msg) {
    std::lock_guard<std::mutex> lk(s->mu);
    r->status = status;
    r->msg = msg;
    r->t_done = Clock::now();
    r->done = true;
    s->lat[r->kind].push_back(seconds(r->t_done - r->t_submit));
    s->counters[r->kind == RK_EVO ? 1 : 2] += 1;
    s->cv_done.notify_all();
}

// Newton batches: every pending problem, whatever its radial size and operator parameters, in one
// problem list (at most two kernel launches: Jacobians in LDS, then in global memory).
void run_newton(hqt_server* s, std::vector<std::shared_ptr<Request>>& batch) {
    const int nb = (int)batch.size();
    NewtonProblems P;
    P.resize(nb);
    std::vector<std::vector<double>> o1(nb), o2(nb), ov(nb);
    std::vector<double> err(nb);
    std::vector<int32_t> it(nb), solves(nb);
    std::vector<uint8_t> conv(nb);
    for (int k = 0; k < nb; ++k) {
        const Request& r = *batch[k];
        P.nz[k] = r.nz;
        P.zh[k] = r.zh;
        P.small[k] = r.small;
        P.D[k] = r.D;
        P.vy[k] = r.vy;
        P.mu_seed[k] = r.mu_seed;
        P.dmu[k] = r.dmu;
        P.Z1[k] = r.n1.data();
        P.Z2[k] = r.n2.data();
        P.vy1[k] = r.nv.data();
        o1[k].resize(r.nz);
        o2[k].resize(r.nz);
        ov[k].resize(r.nz);
        P.z1[k] = o1[k].data();
        P.z2[k] = o2[k].data();
        P.VY1[k] = ov[k].data();
    }
    P.it = it.data();
    P.err = err.data();
    P.conv = conv.data();
    P.steps = solves.data();
    const int rc = newton_problems_run(s->nstream, s->nwork, P, 1000, 1e-6);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->counters[4] += 1;
    }
    const std::string emsg = rc ? std::string(hqt_last_error()) : std::string();
    for (int k = 0; k < nb; ++k) {
        auto& r = batch[k];
        if (rc) { complete(s, r, rc, emsg); continue; }
        r->n1.swap(o1[k]);
        r->n2.swap(o2[k]);
        r->nv.swap(ov[k]);
        r->it = it[k];
        r->err = err[k];
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->counters[8] += std::max(1, solves[k]);
            s->counters[9] += solves[k] > 0;
        }
        complete(s, r, conv[k] ? HQT_OK : HQT_ERR_NOCONV,
                 conv[k] ? "" : "NewtonRaphson_vy: no convergence (it=" + std::to_string(it[k]) + ")");
    }
}

int admit(hqt_server* s, const std::shared_ptr<Request>& r) {
    hqt_params p = r->p;
    p.ax_b = r->axb.empty() ? nullptr : r->axb.data();
    p.ay_b = r->ayb.empty() ? nullptr : r->ayb.data();
    hqt_ctx* c = nullptr;
    int rc = hqt_ctx_create(&p, &c);
    if (rc) return rc;
    rc = hqt_set_state(c, r->Z1.data(), r->Z2.data(), r->Z3.data(), r->Z4.data(), r->rho.data());
    if (rc) { hqt_ctx_destroy(c); return rc; }
    r->ctx = c;
    r->bytes = c->bytes;
    return HQT_OK;
}

void drop_ensemble(hqt_server* s) {
    if (s->ens) {
        hqt_ensemble_destroy(s->ens);
        s->ens = nullptr;
    }
}

void worker(hqt_server* s) {
    for (;;) {
        std::vector<std::shared_ptr<Request>> nbatch, admitted, rejected;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                return s->stop || !s->pend_newton.empty() || !s->pend_evo.empty() || !s->active.empty();
            });
            if (s->stop) break;
            nbatch.assign(s->pend_newton.begin(), s->pend_newton.end());
            s->pend_newton.clear();
            // admission in arrival order while the footprint fits the budget
            while (!s->pend_evo.empty()) {
                auto r = s->pend_evo.front();
                const int64_t need = ctx_footprint(r->p);
                if (need > s->budget) {             // can never fit: fail it now
                    rejected.push_back(r);
                    s->pend_evo.pop_front();
                    continue;
                }
                if (s->used + need > s->budget) break;
                admitted.push_back(r);
                s->used += need;
                s->pend_evo.pop_front();
            }
        }
        for (auto& r : rejected)
            complete(s, r, HQT_ERR_OOM,
                     "evolution needs " + std::to_string(ctx_footprint(r->p)) + " device bytes, above the budget");
        bool changed = false;
        for (auto& r : admitted) {
            const int64_t est = ctx_footprint(r->p);
            const int rc = admit(s, r);
            if (rc) {
                s->used -= est;
                complete(s, r, rc, hqt_last_error());
                continue;
            }
            s->used += r->bytes - est;
            s->active.push_back(r);
            changed = true;
            std::lock_guard<std::mutex> lk(s->mu);
            s->counters[0] += 1;
            s->counters[5] = std::max(s->counters[5], s->used);
            s->counters[6] = std::max<int64_t>(s->counters[6], (int64_t)s->active.size());
        }
        int32_t steps = 0;
        if (!s->active.empty()) {
            if (changed || !s->ens) {
                drop_ensemble(s);
                std::vector<hqt_ctx*> ms;
                for (auto& r : s->active) ms.push_back(r->ctx);
                const int rc = hqt_ensemble_create(ms.data(), (int32_t)ms.size(), 0, &s->ens);
                if (rc) {
                    const std::string m = hqt_last_error();
                    for (auto& r : s->active) {
                        hqt_ctx_destroy(r->ctx);
                        r->ctx = nullptr;
                        s->used -= r->bytes;
                        complete(s, r, rc, m);
                    }
                    s->active.clear();
                    s->ens = nullptr;
                }
                std::lock_guard<std::mutex> lk(s->mu);
                s->counters[3] += 1;
            }
            if (s->ens) {
                steps = s->quantum;
                for (auto& r : s->active) steps = std::min(steps, r->nsteps - r->steps_done);
                if (steps > 0 && hqt_ensemble_step_n(s->ens, steps) != HQT_OK) steps = -1;
            }
        }
        // Newton batches run on their own stream while the ensemble's steps execute
        if (!nbatch.empty()) run_newton(s, nbatch);
        if (s->ens) {
            std::vector<int32_t> ns(s->active.size(), 0);
            const int src = steps < 0 ? HQT_ERR_HIP : hqt_ensemble_sync(s->ens, ns.data());
            const std::string serr = src ? std::string(hqt_last_error()) : std::string();
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->counters[7] += std::max(steps, 0);
            }
            std::vector<std::shared_ptr<Request>> keep, fin;
            for (size_t k = 0; k < s->active.size(); ++k) {
                auto& r = s->active[k];
                r->steps_done += std::max(steps, 0);
                r->nan_step = ns[k];
                const bool hard = src != HQT_OK && src != HQT_ERR_NAN;
                if (hard || ns[k] != 0 || r->steps_done >= r->nsteps) fin.push_back(r);
                else keep.push_back(r);
            }
            if (!fin.empty()) {
                drop_ensemble(s);   // members leave the ensemble before they are released
                for (auto& r : fin) {
                    const int g = hqt_get_state(r->ctx, r->Z1.data(), r->Z2.data(), r->Z3.data(), r->Z4.data(),
                                                r->rho.data());
                    const std::string gm = g ? std::string(hqt_last_error()) : std::string();
                    hqt_ctx_destroy(r->ctx);
                    r->ctx = nullptr;
                    s->used -= r->bytes;
                    const bool hard = src != HQT_OK && src != HQT_ERR_NAN;
                    complete(s, r, hard ? src : g, hard ? serr : gm);
                }
                s->active = keep;
            }
        }
    }
    // shutdown: release what is still resident; unfinished requests end with HQT_ERR_STATE
    drop_ensemble(s);
    for (auto& r : s->active) {
        hqt_ctx_destroy(r->ctx);
        r->ctx = nullptr;
        complete(s, r, HQT_ERR_STATE, "server stopped");
    }
    s->active.clear();
    std::vector<std::shared_ptr<Request>> rest;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        rest.assign(s->pend_evo.begin(), s->pend_evo.end());
        rest.insert(rest.end(), s->pend_newton.begin(), s->pend_newton.end());
        s->pend_evo.clear();
        s->pend_newton.clear();
    }
    for (auto& r : rest) complete(s, r, HQT_ERR_STATE, "server stopped");
}

std::shared_ptr<Request> find(hqt_server* s, int64_t id) {
    auto it = s->reqs.find(id);
    return it == s->reqs.end() ? nullptr : it->second;
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    size_t k = (size_t)std::ceil(q * v.size());   // nearest rank
    if (k < 1) k = 1;
    if (k > v.size()) k = v.size();
    return v[k - 1];
}

}  // namespace

namespace {
int submit_newton(hqt_server* s, const char* who, int32_t nz, double zh, double small, double D, const double* Z1,
                  const double* Z2, const double* vy1, double vy, double mu_seed, double dmu, int64_t* id) {
    if (!s || !Z1 || !Z2 || !vy1 || !id || nz < 3 || nz > 64)
        return hqt_fail(HQT_ERR_ARG, std::string(who) + ": bad argument (3 <= nz <= 64)");
    if (int rc = newton_cont_args(who, nz, 1, Z2, &mu_seed, &dmu)) return rc;
    auto r = std::make_shared<Request>();
    r->kind = RK_NEWTON;
    r->nz = nz;
    r->zh = zh;
    r->small = small;
    r->D = D;
    r->vy = vy;
    r->mu_seed = mu_seed;
    r->dmu = dmu;
    r->n1.assign(Z1, Z1 + nz);
    r->n2.assign(Z2, Z2 + nz);
    r->nv.assign(vy1, vy1 + nz);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        r->t_submit = Clock::now();
        s->reqs[r->id] = r;
        s->pend_newton.push_back(r);
    }
    s->cv_work.notify_all();
    *id = r->id;
    return HQT_OK;
}
}  // namespace

extern "C" {

int hqt_server_create(int64_t hbm_budget, int32_t quantum, hqt_server** out) {
    if (!out || hbm_budget <= 0 || quantum < 1) return hqt_fail(HQT_ERR_ARG, "hqt_server_create: need budget > 0, quantum >= 1");
    *out = nullptr;
    hqt_server* s = new (std::nothrow) hqt_server();
    if (!s) return hqt_fail(HQT_ERR_OOM, "host allocation failed");
    s->budget = hbm_budget;
    s->quantum = quantum;
    if (hipStreamCreateWithFlags(&s->nstream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return hqt_fail(HQT_ERR_HIP, "hipStreamCreate");
    }
    s->worker = std::thread(worker, s);
    *out = s;
    return HQT_OK;
}

int hqt_server_destroy(hqt_server* s) {
    if (!s) return HQT_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipStreamSynchronize(s->nstream);
    (void)hipStreamDestroy(s->nstream);
    delete s;
    return HQT_OK;
}

int hqt_server_submit_evolution(hqt_server* s, const hqt_params* p, const double* Z1, const double* Z2,
                                const double* Z3, const double* Z4, const double* rhoi, int32_t nsteps, int64_t* id) {
    if (!s || !p || !Z1 || !Z2 || !Z3 || !Z4 || !rhoi || !id || nsteps < 1)
        return hqt_fail(HQT_ERR_ARG, "hqt_server_submit_evolution: bad argument");
    if (p->nx < 2 || p->ny < 2 || p->nx > 512 || p->ny > 512 || !zkernels_support_nz(p->nz))
        return hqt_fail(HQT_ERR_ARG, "hqt_server_submit_evolution: grid out of range");
    auto r = std::make_shared<Request>();
    r->kind = RK_EVO;
    r->p = *p;
    const size_t nxy = (size_t)p->nx * p->ny, n = nxy * p->nz;
    if (p->ax_b) r->axb.assign(p->ax_b, p->ax_b + nxy);
    if (p->ay_b) r->ayb.assign(p->ay_b, p->ay_b + nxy);
    r->p.ax_b = nullptr;
    r->p.ay_b = nullptr;
    r->Z1.assign(Z1, Z1 + 2 * n);
    r->Z2.assign(Z2, Z2 + n);
    r->Z3.assign(Z3, Z3 + n);
    r->Z4.assign(Z4, Z4 + n);
    r->rho.assign(rhoi, rhoi + n);
    r->nsteps = nsteps;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        r->t_submit = Clock::now();
        s->reqs[r->id] = r;
        s->pend_evo.push_back(r);
    }
    s->cv_work.notify_all();
    *id = r->id;
    return HQT_OK;
}

int hqt_server_submit_newton(hqt_server* s, int32_t nz, double zh, double small, double D, const double* Z1,
                             const double* Z2, const double* vy1, double vy, int64_t* id) {
    return submit_newton(s, "hqt_server_submit_newton", nz, zh, small, D, Z1, Z2, vy1, vy, std::nan(""), 0.0, id);
}

int hqt_server_submit_newton_continued(hqt_server* s, int32_t nz, double zh, double small, double D,
                                       const double* Z1, const double* Z2, const double* vy1, double vy,
                                       double mu_seed, double dmu, int64_t* id) {
    return submit_newton(s, "hqt_server_submit_newton_continued", nz, zh, small, D, Z1, Z2, vy1, vy, mu_seed, dmu,
                         id);
}

int hqt_server_wait(hqt_server* s, int64_t id, double timeout_s, double* latency_s) {
    if (!s) return hqt_fail(HQT_ERR_ARG, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto r = find(s, id);
    if (!r) return hqt_fail(HQT_ERR_ARG, "unknown request id " + std::to_string(id));
    auto pred = [&] { return r->done; };
    if (timeout_s < 0) s->cv_done.wait(lk, pred);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
        return hqt_fail(HQT_ERR_TIMEOUT, "request " + std::to_string(id) + " not done");
    if (latency_s) *latency_s = seconds(r->t_done - r->t_submit);
    if (r->status) return hqt_fail(r->status, r->msg);
    return HQT_OK;
}

int hqt_server_evolution_result(hqt_server* s, int64_t id, double* Z1, double* Z2, double* Z3, double* Z4,
                                double* rhoi, int32_t* nan_step) {
    if (!s) return hqt_fail(HQT_ERR_ARG, "NULL server");
    std::shared_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r = find(s, id);
        if (!r || r->kind != RK_EVO) return hqt_fail(HQT_ERR_ARG, "no evolution request " + std::to_string(id));
        if (!r->done) return hqt_fail(HQT_ERR_STATE, "request " + std::to_string(id) + " not done");
        s->reqs.erase(id);
    }
    const size_t n = r->Z2.size();
    if (Z1) std::memcpy(Z1, r->Z1.data(), 2 * n * 8);
    if (Z2) std::memcpy(Z2, r->Z2.data(), n * 8);
    if (Z3) std::memcpy(Z3, r->Z3.data(), n * 8);
    if (Z4) std::memcpy(Z4, r->Z4.data(), n * 8);
    if (rhoi) std::memcpy(rhoi, r->rho.data(), n * 8);
    if (nan_step) *nan_step = r->nan_step;
    if (r->status) return hqt_fail(r->status, r->msg);
    return HQT_OK;
}

int hqt_server_newton_result(hqt_server* s, int64_t id, double* z1, double* z2, double* VY1, int32_t* it, double* err) {
    if (!s) return hqt_fail(HQT_ERR_ARG, "NULL server");
    std::shared_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r = find(s, id);
        if (!r || r->kind != RK_NEWTON) return hqt_fail(HQT_ERR_ARG, "no Newton request " + std::to_string(id));
        if (!r->done) return hqt_fail(HQT_ERR_STATE, "request " + std::to_string(id) + " not done");
        s->reqs.erase(id);
    }
    const size_t nz = r->n1.size();
    if (z1) std::memcpy(z1, r->n1.data(), nz * 8);
    if (z2) std::memcpy(z2, r->n2.data(), nz * 8);
    if (VY1) std::memcpy(VY1, r->nv.data(), nz * 8);
    if (it) *it = r->it;
    if (err) *err = r->err;
    if (r->status) return hqt_fail(r->status, r->msg);
    return HQT_OK;
}

int hqt_server_stats(hqt_server* s, int32_t kind, double* p50, double* p99, int64_t* completed) {
    if (!s || kind < 0 || kind > 2) return hqt_fail(HQT_ERR_ARG, "bad argument");
    std::vector<double> v;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (kind != RK_NEWTON) v.insert(v.end(), s->lat[RK_EVO].begin(), s->lat[RK_EVO].end());
        if (kind != RK_EVO) v.insert(v.end(), s->lat[RK_NEWTON].begin(), s->lat[RK_NEWTON].end());
    }
    if (p50) *p50 = percentile(v, 0.50);
    if (p99) *p99 = percentile(v, 0.99);
    if (completed) *completed = (int64_t)v.size();
    return HQT_OK;
}

int hqt_server_counters(hqt_server* s, int64_t* out) {
    if (!s || !out) return hqt_fail(HQT_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    for (int k = 0; k < HQT_SERVER_NCOUNTERS; ++k) out[k] = s->counters[k];
    return HQT_OK;
}

int hqt_ctx_footprint(const hqt_params* p, int64_t* bytes) {
    if (!p || !bytes) return hqt_fail(HQT_ERR_ARG, "NULL argument");
    *bytes = ctx_footprint(*p);
    return HQT_OK;
}

int hqt_ctx_device_bytes(hqt_ctx* c, int64_t* bytes) {
    if (!c || !bytes) return hqt_fail(HQT_ERR_ARG, "NULL argument");
    *bytes = c->bytes;
    return HQT_OK;
}

}  // extern "C"
