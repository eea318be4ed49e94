// egs_serve.cpp -- request loop for likelihood evaluations and kriging over
// many spatial regions of uneven size on one GPU.
//
// Regions (locations + observations) are registered once and pinned to a
// worker.  Each worker owns a library context, i.e. its own pair of HIP
// streams, so requests for different regions run concurrently on the GPU
// while requests for one region run in order on its worker (a kriging request
// that follows a likelihood evaluation at the same theta reuses the factor that
// stays in HBM).  A region's device state is created when a request needs it
// and counts against an HBM budget; when the budget would be exceeded the least
// recently used idle regions are released, and a request waits until enough
// memory is free.  Latency is measured per request from submission to
// completion.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/exageostat_amd.h"

namespace {

double now_ms() {
    using clock = std::chrono::steady_clock;
    return std::chrono::duration<double, std::milli>(clock::now().time_since_epoch()).count();
}

struct Region {
    int kernel = 0, metric = 0, dts = 0, ntheta = 3, p = 1;
    int64_t n_loc = 0;
    std::vector<double> x, y, z, obs;
    int worker = 0;
    egs_problem_t prob = nullptr;
    int64_t footprint = 0;      // device bytes while resident (estimate until created)
    uint64_t last_use = 0;
    int busy = 0;
};

enum Kind { kLoglik = 0, kPredict = 1 };

struct Request {
    int64_t ticket = 0;
    int kind = kLoglik;
    int region = 0;
    std::vector<double> theta, xm, ym;
    double t_submit = 0.0;
};

struct Result {
    bool done = false;
    int status = EGS_SUCCESS;
    std::string error;
    double value = 0.0;
    std::vector<double> zmiss;
    double latency_ms = 0.0;
};

struct Worker {
    egs_context_t ctx = nullptr;
    std::deque<Request> queue;
    std::thread thread;
    int64_t assigned_bytes = 0;
};

// Device bytes of a region's problem (tiles, inverses, vectors), before it exists.
int64_t estimate_bytes(const Region &r) {
    const int64_t ts = std::min<int64_t>(((r.dts + 63) / 64) * 64, 2048);
    const int64_t n = r.n_loc * r.p;
    const int64_t nt = (n + ts - 1) / ts;
    return nt * (nt + 1) / 2 * ts * ts * 8 + nt * ts * 64 * 8 + 8 * ts * ts + 4 * nt * ts * 8 + r.n_loc * 24;
}

}  // namespace

struct egs_server_s {
    int device = 0;
    int64_t budget = 0;
    std::vector<Worker> workers;
    std::deque<Region> regions;   // stable addresses: workers hold pointers while regions are added
    std::unordered_map<int64_t, Result> results;
    std::mutex m;
    std::condition_variable cv_work, cv_done, cv_mem;
    int64_t next_ticket = 1;
    int64_t resident = 0;
    uint64_t clock = 0;
    bool stop = false;
    std::vector<double> latencies;
    double t_first = -1.0, t_last = 0.0;

    // Releases idle resident regions, least recently used first, until `need`
    // more bytes fit.  Called with m held.
    bool make_room(int64_t need) {
        while (resident + need > budget) {
            int victim = -1;
            for (int i = 0; i < (int) regions.size(); ++i) {
                const Region &r = regions[i];
                if (r.prob && r.busy == 0 && (victim < 0 || r.last_use < regions[victim].last_use)) victim = i;
            }
            if (victim < 0) return false;
            Region &v = regions[victim];
            egs_problem_destroy(v.prob);
            v.prob = nullptr;
            resident -= v.footprint;
        }
        return true;
    }

    void run_worker(int w) {
        Worker &wk = workers[w];
        for (;;) {
            Request rq;
            {
                std::unique_lock<std::mutex> lk(m);
                cv_work.wait(lk, [&] { return stop || !wk.queue.empty(); });
                if (stop && wk.queue.empty()) return;
                rq = std::move(wk.queue.front());
                wk.queue.pop_front();
            }
            Result res;
            int st = EGS_SUCCESS;
            Region *reg = nullptr;
            {
                std::unique_lock<std::mutex> lk(m);
                reg = &regions[rq.region];
                reg->busy++;
                reg->last_use = ++clock;
                if (!reg->prob) {
                    const int64_t need = estimate_bytes(*reg);
                    if (need > budget) {
                        st = EGS_ERR_OUT_OF_MEMORY;
                        res.error = "region does not fit the HBM budget";
                    } else {
                        cv_mem.wait(lk, [&] { return make_room(need); });
                        // creation and upload under the lock: the context's problem list is shared
                        st = egs_problem_create(wk.ctx, reg->kernel, reg->n_loc, reg->x.data(), reg->y.data(),
                                                reg->z.empty() ? nullptr : reg->z.data(), reg->metric, reg->dts,
                                                &reg->prob);
                        if (st == EGS_SUCCESS)
                            st = egs_problem_set_observations(reg->prob, reg->obs.data(), (int64_t) reg->obs.size());
                        if (st == EGS_SUCCESS) {
                            reg->footprint = need;   // the estimate bounds the problem's device bytes
                            resident += reg->footprint;
                        } else {
                            res.error = egs_last_error();
                            if (reg->prob) egs_problem_destroy(reg->prob);
                            reg->prob = nullptr;
                        }
                    }
                }
            }
            if (st == EGS_SUCCESS) {
                if (rq.kind == kLoglik) {
                    st = egs_loglik(reg->prob, rq.theta.data(), &res.value, nullptr, nullptr);
                } else {
                    res.zmiss.assign(rq.xm.size() * (size_t) reg->p, 0.0);
                    st = egs_predict(reg->prob, rq.theta.data(), (int64_t) rq.xm.size(), rq.xm.data(), rq.ym.data(),
                                     nullptr, res.zmiss.data(), nullptr, &res.value);
                }
                if (st != EGS_SUCCESS) res.error = egs_last_error();
            }
            const double t = now_ms();
            {
                std::lock_guard<std::mutex> lk(m);
                reg->busy--;
                res.status = st;
                res.done = true;
                res.latency_ms = t - rq.t_submit;
                latencies.push_back(res.latency_ms);
                t_last = std::max(t_last, t);
                results[rq.ticket] = std::move(res);
            }
            cv_done.notify_all();
            cv_mem.notify_all();
        }
    }

    // Queues a request on the region's worker; theta is copied with the
    // region's parameter count.  Returns false for an unknown region.
    bool submit(Request &&rq, const double *theta, int64_t *ticket) {
        {
            std::lock_guard<std::mutex> lk(m);
            if (rq.region < 0 || rq.region >= (int) regions.size()) return false;
            const Region &r = regions[rq.region];
            rq.theta.assign(theta, theta + r.ntheta);
            rq.ticket = next_ticket++;
            rq.t_submit = now_ms();
            if (t_first < 0.0) t_first = rq.t_submit;
            *ticket = rq.ticket;
            results[rq.ticket] = Result();
            workers[r.worker].queue.push_back(std::move(rq));
        }
        cv_work.notify_all();
        return true;
    }
};

namespace egs {
void set_last_error(const char *msg);   // egs_capi.cpp: the message egs_last_error() returns
}

namespace {
int serve_fail(int code, const char *msg) {
    egs::set_last_error(msg);
    return code;
}
}  // namespace

extern "C" {

int egs_server_create(int device, int64_t hbm_budget_bytes, int workers, egs_server_t *server) {
    if (!server || workers < 1 || workers > 16 || hbm_budget_bytes <= 0)
        return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_create: need 1..16 workers and a positive budget");
    auto *s = new egs_server_s();
    s->device = device;
    s->budget = hbm_budget_bytes;
    s->workers.resize(workers);
    for (auto &w : s->workers) {
        const int st = egs_init(device, &w.ctx);
        if (st != EGS_SUCCESS) {
            for (auto &v : s->workers)
                if (v.ctx) egs_finalize(v.ctx);
            delete s;
            return st;
        }
    }
    for (int i = 0; i < workers; ++i) s->workers[i].thread = std::thread([s, i] { s->run_worker(i); });
    *server = s;
    return EGS_SUCCESS;
}

int egs_server_destroy(egs_server_t s) {
    if (!s) return EGS_SUCCESS;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &w : s->workers)
        if (w.thread.joinable()) w.thread.join();
    for (auto &r : s->regions)
        if (r.prob) egs_problem_destroy(r.prob);
    for (auto &w : s->workers)
        if (w.ctx) egs_finalize(w.ctx);
    delete s;
    return EGS_SUCCESS;
}

int egs_server_add_region(egs_server_t s, int kernel, int64_t n_loc, const double *x, const double *y,
                          const double *z, const double *obs, int metric, int dts, int *region) {
    if (!s || !x || !y || !obs || !region || n_loc <= 0 || dts <= 0)
        return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_add_region: invalid arguments");
    if (kernel != EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY && kernel != EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS)
        return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_add_region: unknown kernel");
    Region r;
    r.kernel = kernel;
    r.p = kernel == EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS ? 2 : 1;
    r.ntheta = r.p == 2 ? 6 : 3;
    r.metric = metric;
    r.dts = dts;
    r.n_loc = n_loc;
    r.x.assign(x, x + n_loc);
    r.y.assign(y, y + n_loc);
    if (z) r.z.assign(z, z + n_loc);
    r.obs.assign(obs, obs + n_loc * r.p);
    std::lock_guard<std::mutex> lk(s->m);
    // pin to the worker with the least assigned work (n^3 of its regions)
    int best = 0;
    for (int i = 1; i < (int) s->workers.size(); ++i)
        if (s->workers[i].assigned_bytes < s->workers[best].assigned_bytes) best = i;
    r.worker = best;
    const double nn = (double) (n_loc * r.p);
    s->workers[best].assigned_bytes += (int64_t) (nn * nn * nn / 1e6);
    s->regions.push_back(std::move(r));
    *region = (int) s->regions.size() - 1;
    return EGS_SUCCESS;
}

int egs_server_submit_loglik(egs_server_t s, int region, const double *theta, int64_t *ticket) {
    if (!s || !theta || !ticket) return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_submit_loglik: null argument");
    Request rq;
    rq.kind = kLoglik;
    rq.region = region;
    if (!s->submit(std::move(rq), theta, ticket))
        return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_submit_loglik: unknown region");
    return EGS_SUCCESS;
}

int egs_server_submit_predict(egs_server_t s, int region, const double *theta, int64_t n_miss, const double *xm,
                              const double *ym, int64_t *ticket) {
    if (!s || !theta || !xm || !ym || !ticket || n_miss <= 0)
        return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_submit_predict: invalid arguments");
    Request rq;
    rq.kind = kPredict;
    rq.region = region;
    rq.xm.assign(xm, xm + n_miss);
    rq.ym.assign(ym, ym + n_miss);
    if (!s->submit(std::move(rq), theta, ticket))
        return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_submit_predict: unknown region");
    return EGS_SUCCESS;
}

int egs_server_wait(egs_server_t s, int64_t ticket, double *value, double *zmiss, int64_t zmiss_len,
                    double *latency_ms) {
    if (!s) return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_wait: null server");
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->results.find(ticket);
    if (it == s->results.end()) return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_wait: unknown ticket");
    s->cv_done.wait(lk, [&] { return s->results[ticket].done; });
    Result res = std::move(s->results[ticket]);
    s->results.erase(ticket);
    lk.unlock();
    if (res.status != EGS_SUCCESS) return serve_fail(res.status, res.error.c_str());
    if (value) *value = res.value;
    if (latency_ms) *latency_ms = res.latency_ms;
    if (zmiss) {
        if ((int64_t) res.zmiss.size() > zmiss_len)
            return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_wait: zmiss buffer too small");
        std::copy(res.zmiss.begin(), res.zmiss.end(), zmiss);
    }
    return EGS_SUCCESS;
}

int egs_server_stats(egs_server_t s, int64_t *completed, double *p50_ms, double *p99_ms, double *per_second,
                     int64_t *resident_bytes) {
    if (!s) return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_stats: null server");
    std::lock_guard<std::mutex> lk(s->m);
    std::vector<double> l = s->latencies;
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
        if (l.empty()) return 0.0;
        const size_t i = std::min(l.size() - 1, (size_t) std::ceil(q * (double) l.size()) - 1);
        return l[i];
    };
    if (completed) *completed = (int64_t) l.size();
    if (p50_ms) *p50_ms = pct(0.50);
    if (p99_ms) *p99_ms = pct(0.99);
    if (per_second)
        *per_second = (l.empty() || s->t_last <= s->t_first) ? 0.0 : 1000.0 * (double) l.size() / (s->t_last - s->t_first);
    if (resident_bytes) *resident_bytes = s->resident;
    return EGS_SUCCESS;
}

int egs_server_reset_stats(egs_server_t s) {
    if (!s) return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_reset_stats: null server");
    std::lock_guard<std::mutex> lk(s->m);
    s->latencies.clear();
    s->t_first = -1.0;
    s->t_last = 0.0;
    return EGS_SUCCESS;
}

}  // extern "C"
