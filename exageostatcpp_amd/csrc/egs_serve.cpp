// egs_serve.cpp -- request loop for likelihood evaluations and kriging over
// many spatial regions of uneven size on one GPU.
//
// Requests wait in one queue.  Each worker owns a library context (its own
// pair of HIP streams) and takes the oldest request whose region is not in
// use; a region's requests therefore run one at a time in submission order,
// while requests for different regions run concurrently on the workers'
// streams.  A worker that takes a likelihood request for a small region under
// the FP64 plan also takes the other queued small FP64 likelihood requests of
// the same tile size (up to kBatchMax) and evaluates them with
// egs_loglik_batch: one launch per kernel and factorization step for all of
// them, so their latency-bound diagonal-tile chains run side by side.
//
// A region's device state (locations, observations, tiles and, once
// evaluated, its Cholesky factor) is created when a request needs it and stays
// in HBM for later requests: a kriging request at the theta of the region's
// last evaluation reuses the factor.  Regions are admitted by their footprint
// (egs_problem_estimate_bytes, which counts a mixed plan's panel copies, FP64
// copy of Sigma and inverses) against the HBM budget; when a region does not
// fit, the least recently used idle regions are released, and the request
// waits while the regions in use leave no room.  Creation, release and upload
// happen outside the server's lock; released blocks stay in the library's
// device block cache (limit: the budget) so that releasing a region does not
// wait for the whole device.  Latency is measured per request from submission
// to completion.
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

// Batches take likelihood requests of regions with at most kBatchTiles tiles
// (small problems, whose evaluation is bound by the diagonal-tile chain), at
// most kBatchMax at a time.
constexpr int kBatchTiles = 16;
constexpr int kBatchMax = 32;

struct Region {
    int kernel = 0, metric = 0, dts = 0, ntheta = 3, p = 1, dim = 2;
    int mode = EGS_PRECISION_FP64, band64 = 0, band32 = 0;
    int ts = 0, nt = 0;
    int64_t n_loc = 0;
    std::vector<double> x, y, z, obs;
    egs_problem_t prob = nullptr;
    int64_t footprint = 0;      // device bytes while resident
    uint64_t last_use = 0;
    bool busy = false;
    bool evaluated = false;     // prob holds the factor at last_theta
    double last_theta[6] = {0, 0, 0, 0, 0, 0};
};

enum Kind { kLoglik = 0, kPredict = 1 };

struct Request {
    int64_t ticket = 0;
    int kind = kLoglik;
    int region = 0;
    double theta[6] = {0, 0, 0, 0, 0, 0};
    std::vector<double> xm, ym;
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
    std::thread thread;
};

}  // namespace

struct egs_server_s {
    int device = 0;
    int64_t budget = 0;
    std::vector<Worker> workers;
    std::deque<Region> regions;   // stable addresses while regions are added
    std::deque<Request> pending;
    std::unordered_map<int64_t, Result> results;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    int64_t next_ticket = 1;
    int64_t resident = 0;
    uint64_t clock = 0;
    bool stop = false;
    std::vector<double> latencies;
    double t_first = -1.0, t_last = 0.0;
    int64_t batches = 0, batched_requests = 0, predicts = 0, predict_hits = 0, switched_regions = 0;

    bool batchable(const Region &r) const {
        return r.mode == EGS_PRECISION_FP64 && r.nt <= kBatchTiles;
    }

    // Idle resident regions, least recently used first, released until `need`
    // more bytes fit; the released problems are appended to `victims` (m held).
    // Releases nothing and returns false when even all idle regions would not
    // make room.
    bool make_room(int64_t need, std::vector<egs_problem_t> &victims) {
        if (resident + need <= budget) return true;
        int64_t idle = 0;
        for (const Region &r : regions)
            if (r.prob && !r.busy) idle += r.footprint;
        if (resident - idle + need > budget) return false;
        while (resident + need > budget) {
            int victim = -1;
            for (int i = 0; i < (int) regions.size(); ++i) {
                const Region &r = regions[i];
                if (r.prob && !r.busy && (victim < 0 || r.last_use < regions[victim].last_use)) victim = i;
            }
            Region &v = regions[victim];
            victims.push_back(v.prob);
            v.prob = nullptr;
            v.evaluated = false;
            resident -= v.footprint;
        }
        return true;
    }

    // Takes the next work item for a worker (m held): the oldest request whose
    // region is idle and fits, plus batchable companions.  Returns false when
    // nothing can start now.
    bool take(std::vector<Request> &batch, std::vector<Region *> &regs, std::vector<egs_problem_t> &victims) {
        std::vector<char> seen(regions.size(), 0);
        for (size_t qi = 0; qi < pending.size(); ++qi) {
            const int rid = pending[qi].region;
            if (seen[rid]) continue;   // an earlier request of this region waits
            seen[rid] = 1;
            Region &r = regions[rid];
            if (r.busy) continue;
            const int64_t need = r.prob ? 0 : r.footprint;
            if (need > 0 && !make_room(need, victims)) continue;
            resident += need;
            r.busy = true;
            batch.push_back(std::move(pending[qi]));
            regs.push_back(&r);
            pending.erase(pending.begin() + (long) qi);
            if (batch[0].kind != kLoglik || !batchable(r)) return true;
            // companions: later small FP64 likelihood requests of idle regions with the same tile size
            for (size_t qj = qi; qj < pending.size() && (int) batch.size() < kBatchMax;) {
                const int cid = pending[qj].region;
                Region &c = regions[cid];
                const bool first_of_region = !seen[cid];
                seen[cid] = 1;
                if (first_of_region && !c.busy && pending[qj].kind == kLoglik && batchable(c) && c.ts == r.ts) {
                    const int64_t cneed = c.prob ? 0 : c.footprint;
                    if (cneed == 0 || make_room(cneed, victims)) {
                        resident += cneed;
                        c.busy = true;
                        batch.push_back(std::move(pending[qj]));
                        regs.push_back(&c);
                        pending.erase(pending.begin() + (long) qj);
                        continue;
                    }
                }
                ++qj;
            }
            return true;
        }
        return false;
    }

    // Makes the region's problem resident on this worker's context (outside
    // the lock; its bytes are already counted).
    int prepare(Region &r, egs_context_t ctx) {
        if (r.prob) return egs_problem_set_context(r.prob, ctx);
        egs_problem_t p = nullptr;
        int st = egs_problem_create(ctx, r.kernel, r.n_loc, r.x.data(), r.y.data(), r.z.empty() ? nullptr : r.z.data(),
                                    r.metric, r.dts, &p);
        if (st == EGS_SUCCESS) st = egs_problem_set_observations(p, r.obs.data(), (int64_t) r.obs.size());
        if (st == EGS_SUCCESS && r.mode != EGS_PRECISION_FP64)
            st = egs_problem_set_precision(p, r.mode, r.band64, r.band32);
        if (st != EGS_SUCCESS) {
            if (p) egs_problem_destroy(p);
            return st;
        }
        r.prob = p;
        return EGS_SUCCESS;
    }

    void run_worker(int w) {
        Worker &wk = workers[w];
        for (;;) {
            std::vector<Request> batch;
            std::vector<Region *> regs;   // regions of the batch (the deque may grow meanwhile)
            std::vector<egs_problem_t> victims;
            {
                std::unique_lock<std::mutex> lk(m);
                cv_work.wait(lk, [&] { return (stop && pending.empty()) || take(batch, regs, victims); });
                if (batch.empty()) return;
            }
            // released regions are idle (their work is complete): moved to this
            // worker's context first, their release waits for no other worker's stream
            for (egs_problem_t v : victims) {
                egs_problem_set_context(v, wk.ctx);
                egs_problem_destroy(v);
            }
            const int nb = (int) batch.size();
            std::vector<Result> res(nb);
            std::vector<int> ready;
            std::vector<char> hit(nb, 0);
            bool switched = false;   // a mixed region fell back and now runs FP64
            for (int b = 0; b < nb; ++b) {
                Region &r = *regs[b];
                const bool had = r.prob != nullptr;
                const int st = prepare(r, wk.ctx);
                if (st != EGS_SUCCESS) {
                    res[b].status = st;
                    res[b].error = egs_last_error();
                    continue;
                }
                hit[b] = had && r.evaluated && std::memcmp(r.last_theta, batch[b].theta, sizeof(double) * r.ntheta) == 0;
                ready.push_back(b);
            }
            if (nb > 1) {
                std::vector<egs_problem_t> probs;
                std::vector<double> th, llh(ready.size()), ld(ready.size()), q(ready.size());
                std::vector<int> status(ready.size());
                for (int b : ready) {
                    probs.push_back(regs[b]->prob);
                    th.insert(th.end(), batch[b].theta, batch[b].theta + 6);
                }
                int st = ready.empty() ? EGS_SUCCESS
                                       : egs_loglik_batch(probs.data(), th.data(), (int) ready.size(), llh.data(),
                                                          ld.data(), q.data(), status.data());
                for (size_t a = 0; a < ready.size(); ++a) {
                    Result &rr = res[ready[a]];
                    rr.status = st != EGS_SUCCESS ? st : status[a];
                    if (rr.status != EGS_SUCCESS) rr.error = egs_last_error();
                    else rr.value = llh[a];
                }
            } else if (!ready.empty()) {
                Request &rq = batch[0];
                Region &r = *regs[0];
                int st;
                if (rq.kind == kLoglik) {
                    st = egs_loglik(r.prob, rq.theta, &res[0].value, nullptr, nullptr);
                    int fell = 0;
                    if (r.mode != EGS_PRECISION_FP64 && egs_last_fallback(r.prob, &fell, nullptr, nullptr) == EGS_SUCCESS &&
                        fell) {
                        // the mixed factor was not good enough for this region's covariance:
                        // later requests use the FP64 plan directly (the problem stays
                        // within the footprint it was admitted with)
                        if (egs_problem_set_precision(r.prob, EGS_PRECISION_FP64, r.band64, r.band32) == EGS_SUCCESS)
                            switched = true;
                    }
                } else {
                    res[0].zmiss.assign(rq.xm.size() * (size_t) r.p, 0.0);
                    st = egs_predict(r.prob, rq.theta, (int64_t) rq.xm.size(), rq.xm.data(), rq.ym.data(), nullptr,
                                     res[0].zmiss.data(), nullptr, &res[0].value);
                }
                res[0].status = st;
                if (st != EGS_SUCCESS) res[0].error = egs_last_error();
            }
            const double t = now_ms();
            {
                std::lock_guard<std::mutex> lk(m);
                if (nb > 1) {
                    ++batches;
                    batched_requests += nb;
                }
                if (switched) {
                    regs[0]->mode = EGS_PRECISION_FP64;
                    ++switched_regions;
                }
                for (int b = 0; b < nb; ++b) {
                    Request &rq = batch[b];
                    Region &r = *regs[b];
                    r.busy = false;
                    r.last_use = ++clock;
                    if (!r.prob) resident -= r.footprint;   // creation failed: give the bytes back
                    if (rq.kind == kPredict) {
                        ++predicts;
                        predict_hits += hit[b];
                    }
                    // the problem holds the factor of its last successful request
                    if (r.prob && res[b].status == EGS_SUCCESS && !(switched && b == 0)) {
                        r.evaluated = true;
                        std::memcpy(r.last_theta, rq.theta, sizeof(r.last_theta));
                    } else {
                        r.evaluated = false;   // failed, or re-planned to FP64 (unfactored)
                    }
                    Result &out = res[b];
                    out.done = true;
                    out.latency_ms = t - rq.t_submit;
                    latencies.push_back(out.latency_ms);
                    results[rq.ticket] = std::move(out);
                }
                t_last = std::max(t_last, t);
            }
            cv_done.notify_all();
            cv_work.notify_all();   // regions are idle again
        }
    }

    bool submit(Request &&rq, const double *theta, int64_t *ticket) {
        rq.t_submit = now_ms();
        {
            std::lock_guard<std::mutex> lk(m);
            if (rq.region < 0 || rq.region >= (int) regions.size()) return false;
            const Region &r = regions[rq.region];
            std::memcpy(rq.theta, theta, sizeof(double) * r.ntheta);
            rq.ticket = next_ticket++;
            if (t_first < 0.0) t_first = rq.t_submit;
            *ticket = rq.ticket;
            results[rq.ticket] = Result();
            pending.push_back(std::move(rq));
        }
        cv_work.notify_one();
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
    // released regions keep their blocks for the next region (within the budget)
    egs_memory_cache(hbm_budget_bytes, nullptr, nullptr);
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
    egs_memory_cache(0, nullptr, nullptr);   // back to releasing every block
    for (auto &w : s->workers)
        if (w.ctx) egs_finalize(w.ctx);
    delete s;
    return EGS_SUCCESS;
}

int egs_server_add_region_plan(egs_server_t s, int kernel, int64_t n_loc, const double *x, const double *y,
                               const double *z, const double *obs, int metric, int dts, int mode, int band_fp64,
                               int band_fp32, int *region) {
    if (!s || !x || !y || !obs || !region || n_loc <= 0 || dts <= 0)
        return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_add_region: invalid arguments");
    if (kernel != EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY && kernel != EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS)
        return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_add_region: unknown kernel");
    Region r;
    r.kernel = kernel;
    r.p = kernel == EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS ? 2 : 1;
    r.ntheta = r.p == 2 ? 6 : 3;
    r.dim = (z && kernel == EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY) ? 3 : 2;
    r.metric = metric;
    r.dts = dts;
    r.mode = mode;
    r.band64 = band_fp64;
    r.band32 = band_fp32;
    r.n_loc = n_loc;
    r.ts = std::min(((dts + 63) / 64) * 64, 2048);
    r.nt = (int) ((n_loc * r.p + r.ts - 1) / r.ts);
    // a mixed plan whose FP64 band covers every tile stores nothing in lower
    // precision: such a region runs the FP64 plan (and may join batches)
    if (mode != EGS_PRECISION_FP64 && r.nt - 1 <= band_fp64) r.mode = EGS_PRECISION_FP64;
    int64_t fp = 0;
    int st = egs_problem_estimate_bytes(kernel, n_loc, r.dim, dts, mode, band_fp64, band_fp32, &fp);
    if (st != EGS_SUCCESS) return st;
    // plus the prediction buffer of a 100-point request
    r.footprint = fp + 8 * 100 * (r.p + 3) + 4096;
    if (r.footprint > s->budget) return serve_fail(EGS_ERR_OUT_OF_MEMORY, "egs_server_add_region: region exceeds the HBM budget");
    r.x.assign(x, x + n_loc);
    r.y.assign(y, y + n_loc);
    if (r.dim == 3) r.z.assign(z, z + n_loc);
    r.obs.assign(obs, obs + n_loc * r.p);
    std::lock_guard<std::mutex> lk(s->m);
    s->regions.push_back(std::move(r));
    *region = (int) s->regions.size() - 1;
    return EGS_SUCCESS;
}

int egs_server_add_region(egs_server_t s, int kernel, int64_t n_loc, const double *x, const double *y,
                          const double *z, const double *obs, int metric, int dts, int *region) {
    return egs_server_add_region_plan(s, kernel, n_loc, x, y, z, obs, metric, dts, EGS_PRECISION_FP64, 0, 0, region);
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

int egs_server_batch_stats(egs_server_t s, int64_t *batches, int64_t *batched_requests, int64_t *predicts,
                           int64_t *predict_hits, int64_t *fp64_regions) {
    if (!s) return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_batch_stats: null server");
    std::lock_guard<std::mutex> lk(s->m);
    if (fp64_regions) *fp64_regions = s->switched_regions;
    if (batches) *batches = s->batches;
    if (batched_requests) *batched_requests = s->batched_requests;
    if (predicts) *predicts = s->predicts;
    if (predict_hits) *predict_hits = s->predict_hits;
    return EGS_SUCCESS;
}

int egs_server_reset_stats(egs_server_t s) {
    if (!s) return serve_fail(EGS_ERR_INVALID_ARG, "egs_server_reset_stats: null server");
    std::lock_guard<std::mutex> lk(s->m);
    s->latencies.clear();
    s->t_first = -1.0;
    s->t_last = 0.0;
    s->batches = s->batched_requests = s->predicts = s->predict_hits = s->switched_regions = 0;
    return EGS_SUCCESS;
}

}  // extern "C"
