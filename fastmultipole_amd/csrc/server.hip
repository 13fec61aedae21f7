// Request loop for FMM evaluations and FastGaussSeidel solves of uneven sizes.
//
// A server owns n_slots worker threads; each has its own context (HIP stream)
// and evaluation handle, whose trees, lists, expansions and solver state stay in
// HBM and are reused (grown, never shrunk) from one request to the next.
// Requests wait in arrival order; a free worker takes the first waiting request
// whose estimated HBM footprint fits beside the requests already running
// (the budget), or the first waiting request when nothing runs, so a request
// larger than the budget still completes, alone. Requests that arrive while
// others run are admitted as slots and memory free up, and their kernels overlap
// on the device through the workers' streams. A worker that takes an FMM request
// also takes FMM requests queued behind it with the same kernel and options (its
// share of the queue spread over it and the idle workers, up to a request and body
// cap, within the budget) and runs them as one batch
// evaluation (fmmhip_eval_set_batch: one tree per request, every pass of all of
// them in the same launches); each request's results and status are those of its
// own fmm! call. The FMM passes are deterministic, so results are identical
// alone, batched and in the mix; a solve's sweep adds right-hand-side updates
// atomically and matches to rounding. Latency is measured from submit to
// completion.
//
// This layer uses the library's own C ABI, like any other client.
#include "fmmhip.h"
#include "fmmhip_internal.hpp"
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Job {
    int64_t ticket = 0;
    fmmhip_request req;
    std::vector<double> src, ext;   // copies of the caller's buffers
    double footprint = 0.0;
    std::vector<double> targets, strengths;
    int32_t status = FMMHIP_OK, sweeps = 0, pmax = 0;
    int32_t error_success = 1, converged = 1;
    std::string error;
    double t_submit = 0, t_start = 0, t_end = 0;
    bool done = false;
};

// HBM estimate of one request (bytes): body arrays and outputs, two trees with
// their expansions, lists, and for a solve the influence blocks
double footprint_bytes(const fmmhip_request& r) {
    const double n = (double)std::max<int64_t>(1, r.n);
    const int p = std::max(1, r.options.expansion_order);
    const double leaf = std::max(1, r.options.leaf_size_source);
    const double H = 0.5 * (p + 1) * (p + 2);
    const double branches = 2.0 * 1.15 * n / leaf + 64;
    double b = n * 8.0 * (fmmhip::data_per_body(r.kernel) + 3 + 13 + 16 + 24);
    b += 2.0 * branches * (sizeof(fmmhip::Branch) + 32.0 * H);
    b += branches * 400.0 * 16.0;
    if (r.kind == FMMHIP_REQUEST_SOLVE) b += n * leaf * 8.0 * 30.0 + n * 8.0 * 40.0;
    return b;
}

}  // namespace

struct fmmhip_server {
    int device = 0;
    double budget = 0.0;
    double t_created = 0.0;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job*> queue;
    std::map<int64_t, Job*> jobs;
    double in_flight = 0.0;
    int running = 0;
    bool stop = false;
    int64_t next_ticket = 1;
    std::vector<double> latencies, queue_times;
    double first_submit = -1.0, last_done = 0.0;
    int64_t completed = 0;
    std::string init_error;
    int init_status = FMMHIP_OK;
    int ready = 0;
    int32_t batch_requests = 64;        // FMM requests run together at most (1: none)
    int64_t batch_bodies = 2 << 20;     // bodies of one such batch at most
};

namespace {

void run_job(fmmhip_ctx* ctx, fmmhip_eval* ev, Job& J) {
    const fmmhip_request& r = J.req;
    int st = fmmhip_eval_set_systems(ev, r.kernel, r.n, nullptr, r.n, J.src.data(), 0);
    if (st == FMMHIP_OK && r.kind == FMMHIP_REQUEST_FMM) {
        fmmhip_stats s{};
        st = fmmhip_fmm(ev, &r.options, &s);
        J.pmax = s.pmax;
        J.error_success = s.error_success;
    } else if (st == FMMHIP_OK) {
        st = fmmhip_fgs_setup(ev, &r.options, r.influence);
        if (st == FMMHIP_OK) {
            J.strengths.resize((size_t)r.n);
            st = fmmhip_fgs_solve(ev, J.ext.data(), 0, r.max_iterations, r.tolerance, J.strengths.data(), &J.sweeps,
                                  nullptr);
            // solve! stops before the sweep of the first iteration whose residual meets the tolerance;
            // a solve that sweeps max_iterations times never met it
            if (st == FMMHIP_OK && r.n > 0) J.converged = J.sweeps < r.max_iterations;
        }
    }
    if (st != FMMHIP_OK) J.error = fmmhip_last_error(ctx);
    // the call succeeded but the tolerance was not met (reference: error_success == false for fmm!, or
    // solve! ending at max_iterations): results are still delivered
    if (st == FMMHIP_OK && (!J.error_success || !J.converged)) st = FMMHIP_ETOL;
    if (st == FMMHIP_OK || st == FMMHIP_ETOL) {   // a missed tolerance still leaves results, as in the reference
        J.targets.resize((size_t)16 * r.n);
        const int g = fmmhip_eval_get_targets(ev, J.targets.data(), 0);
        if (g != FMMHIP_OK) {
            st = g;
            J.error = fmmhip_last_error(ctx);
        }
    }
    J.status = st;
}

bool same_options(const fmmhip_options& a, const fmmhip_options& b) {
    return a.expansion_order == b.expansion_order && a.leaf_size_source == b.leaf_size_source &&
           a.leaf_size_target == b.leaf_size_target && a.multipole_acceptance == b.multipole_acceptance &&
           a.list_method == b.list_method && a.shrink_recenter == b.shrink_recenter && a.error_kind == b.error_kind &&
           a.error_relative == b.error_relative && a.error_tolerance == b.error_tolerance &&
           a.error_bonus == b.error_bonus && a.scalar_potential == b.scalar_potential && a.gradient == b.gradient &&
           a.hessian == b.hessian && a.farfield == b.farfield && a.nearfield == b.nearfield &&
           a.self_induced == b.self_induced && a.precision == b.precision;
}

// FMM requests of one kernel and options as one batch evaluation; each gets its own results and status
void run_batch(fmmhip_ctx* ctx, fmmhip_eval* ev, std::vector<Job*>& G) {
    const fmmhip_request& r = G[0]->req;
    const size_t dpb = (size_t)fmmhip::data_per_body(r.kernel);
    std::vector<int64_t> counts;
    int64_t n = 0;
    for (Job* J : G) { counts.push_back(J->req.n); n += J->req.n; }
    std::vector<double> src(dpb * (size_t)n), out;
    size_t o = 0;
    for (Job* J : G) { std::memcpy(src.data() + o, J->src.data(), J->src.size() * sizeof(double)); o += J->src.size(); }
    std::vector<int32_t> pmax(G.size(), 0), ok(G.size(), 1);
    int st = fmmhip_eval_set_batch(ev, r.kernel, (int64_t)G.size(), counts.data(), src.data(), 0);
    if (st == FMMHIP_OK) {
        fmmhip_stats s{};
        st = fmmhip_fmm(ev, &r.options, &s);
    }
    if (st == FMMHIP_OK) st = fmmhip_eval_get_batch_stats(ev, pmax.data(), ok.data());
    if (st == FMMHIP_OK) {
        out.resize((size_t)16 * n);
        st = fmmhip_eval_get_targets(ev, out.data(), 0);
    }
    const std::string err = st == FMMHIP_OK ? std::string() : std::string(fmmhip_last_error(ctx));
    o = 0;
    for (size_t i = 0; i < G.size(); ++i) {
        Job& J = *G[i];
        if (st == FMMHIP_OK) {
            J.pmax = pmax[i];
            J.error_success = ok[i];
            J.targets.assign(out.begin() + o, out.begin() + o + (size_t)16 * J.req.n);
            J.status = ok[i] ? FMMHIP_OK : FMMHIP_ETOL;
        } else {
            J.status = st;
            J.error = err;
        }
        o += (size_t)16 * J.req.n;
    }
}

void worker(fmmhip_server* S) {
    fmmhip_ctx* ctx = nullptr;
    fmmhip_eval* ev = nullptr;
    int st = fmmhip_ctx_create(S->device, &ctx);
    if (st == FMMHIP_OK) fmmhip::ctx_disable_graphs(ctx);   // workers run beside each other: no stream capture
    if (st == FMMHIP_OK) st = fmmhip_eval_create(ctx, &ev);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (st != FMMHIP_OK && S->init_status == FMMHIP_OK) {
            S->init_status = st;
            S->init_error = "worker context creation failed";
        }
        S->ready += 1;
    }
    S->cv_done.notify_all();
    while (st == FMMHIP_OK) {
        std::vector<Job*> G;   // the request taken, then the FMM requests that join it
        {
            std::unique_lock<std::mutex> lk(S->mu);
            for (;;) {
                if (S->stop) break;
                auto it = S->queue.begin();
                for (; it != S->queue.end(); ++it)
                    if (S->running == 0 || S->in_flight + (*it)->footprint <= S->budget) break;
                if (it != S->queue.end()) {
                    Job* J = *it;
                    S->queue.erase(it);
                    S->in_flight += J->footprint;
                    S->running += 1;
                    G.push_back(J);
                    if (J->req.kind == FMMHIP_REQUEST_FMM && J->req.n > 0) {
                        // this worker's share of what waits: the queue spread over it and the idle workers,
                        // so a burst still runs on every worker
                        const int64_t idle = std::max<int64_t>(0, (int64_t)S->workers.size() - S->running);
                        const int64_t share = ((int64_t)S->queue.size() + 1 + idle) / (idle + 1);
                        const int64_t most = std::min<int64_t>(S->batch_requests, share);
                        int64_t bodies = J->req.n;
                        for (auto k = S->queue.begin(); k != S->queue.end() && (int64_t)G.size() < most;) {
                            const fmmhip_request& q = (*k)->req;
                            if (q.kind == FMMHIP_REQUEST_FMM && q.kernel == J->req.kernel && q.n > 0 &&
                                bodies + q.n <= S->batch_bodies && same_options(q.options, J->req.options) &&
                                S->in_flight + (*k)->footprint <= S->budget) {
                                S->in_flight += (*k)->footprint;
                                bodies += q.n;
                                G.push_back(*k);
                                k = S->queue.erase(k);
                            } else {
                                ++k;
                            }
                        }
                    }
                    break;
                }
                S->cv_work.wait(lk);
            }
            if (G.empty()) break;
        }
        const double t0 = now_ms();
        if (G.size() == 1) run_job(ctx, ev, *G[0]);
        else run_batch(ctx, ev, G);
        const double t1 = now_ms();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->running -= 1;
            for (Job* J : G) {
                J->t_start = t0;
                J->t_end = t1;
                S->in_flight -= J->footprint;
                J->done = true;
                S->latencies.push_back(J->t_end - J->t_submit);
                S->queue_times.push_back(J->t_start - J->t_submit);
                S->completed += 1;
            }
            S->last_done = std::max(S->last_done, t1);
        }
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
    if (ev) fmmhip_eval_destroy(ev);
    if (ctx) fmmhip_ctx_destroy(ctx);
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" {

fmmhip_status fmmhip_server_create(int device, int32_t n_slots, double hbm_budget_bytes, fmmhip_server** out) {
    if (!out || n_slots < 1 || n_slots > 16 || !(hbm_budget_bytes > 0.0)) return FMMHIP_EINVAL;
    *out = nullptr;
    fmmhip_server* S = new fmmhip_server();
    if (device < 0) {
        if (hipGetDevice(&S->device) != hipSuccess) { delete S; return FMMHIP_EDEVICE; }
    } else S->device = device;
    S->budget = hbm_budget_bytes;
    S->t_created = now_ms();
    for (int i = 0; i < n_slots; ++i) S->workers.emplace_back(worker, S);
    {
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_done.wait(lk, [&] { return S->ready == n_slots; });
    }
    if (S->init_status != FMMHIP_OK) {
        const int st = S->init_status;
        fmmhip_server_destroy(S);
        return (fmmhip_status)st;
    }
    *out = S;
    return FMMHIP_OK;
}

void fmmhip_server_destroy(fmmhip_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto& t : S->workers) t.join();
    for (auto& kv : S->jobs) delete kv.second;
    for (Job* J : S->queue)
        if (!S->jobs.count(J->ticket)) delete J;
    delete S;
}

fmmhip_status fmmhip_server_set_batching(fmmhip_server* S, int32_t max_requests, int64_t max_bodies) {
    if (!S || max_requests < 1 || max_bodies < 1) return FMMHIP_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    S->batch_requests = max_requests;
    S->batch_bodies = max_bodies;
    return FMMHIP_OK;
}

fmmhip_status fmmhip_server_submit(fmmhip_server* S, const fmmhip_request* r, int64_t* ticket) {
    if (!S || !r || !ticket) return FMMHIP_EINVAL;
    if (r->kind != FMMHIP_REQUEST_FMM && r->kind != FMMHIP_REQUEST_SOLVE) return FMMHIP_EINVAL;
    if (r->n < 0 || (r->n > 0 && !r->source_buffer) || !fmmhip::valid_kernel(r->kernel)) return FMMHIP_EINVAL;
    if (r->kind == FMMHIP_REQUEST_SOLVE && r->n > 0 && !r->external) return FMMHIP_EINVAL;
    Job* J = new Job();
    J->req = *r;
    const size_t dpb = (size_t)fmmhip::data_per_body(r->kernel);
    J->src.assign(r->source_buffer, r->source_buffer + dpb * (size_t)r->n);
    if (r->kind == FMMHIP_REQUEST_SOLVE) J->ext.assign(r->external, r->external + (size_t)r->n);
    J->req.source_buffer = nullptr;
    J->req.external = nullptr;
    J->footprint = footprint_bytes(*r);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        J->ticket = S->next_ticket++;
        J->t_submit = now_ms();
        if (S->first_submit < 0) S->first_submit = J->t_submit;
        S->jobs[J->ticket] = J;
        S->queue.push_back(J);
    }
    S->cv_work.notify_all();
    *ticket = J->ticket;
    return FMMHIP_OK;
}

fmmhip_status fmmhip_server_wait(fmmhip_server* S, int64_t ticket, double* target_out, double* strengths_out,
                                 fmmhip_request_stats* stats) {
    if (!S) return FMMHIP_EINVAL;
    Job* J = nullptr;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->jobs.find(ticket);
        if (it == S->jobs.end()) return FMMHIP_EINVAL;
        J = it->second;
        S->cv_done.wait(lk, [&] { return J->done; });
        S->jobs.erase(it);
    }
    const int32_t st = J->status;
    if (st == FMMHIP_OK || st == FMMHIP_ETOL) {
        if (target_out && !J->targets.empty()) std::memcpy(target_out, J->targets.data(), J->targets.size() * sizeof(double));
        if (strengths_out && !J->strengths.empty())
            std::memcpy(strengths_out, J->strengths.data(), J->strengths.size() * sizeof(double));
    }
    if (stats) {
        stats->status = st;
        stats->sweeps = J->sweeps;
        stats->pmax = J->pmax;
        stats->error_success = J->error_success;
        stats->converged = J->converged;
        stats->queue_ms = J->t_start - J->t_submit;
        stats->run_ms = J->t_end - J->t_start;
        stats->latency_ms = J->t_end - J->t_submit;
        stats->start_ms = J->t_start - S->t_created;
        stats->end_ms = J->t_end - S->t_created;
        stats->footprint_bytes = J->footprint;
    }
    delete J;
    return (fmmhip_status)st;
}

fmmhip_status fmmhip_server_stats(fmmhip_server* S, fmmhip_server_stats_t* out) {
    if (!S || !out) return FMMHIP_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    out->completed = S->completed;
    out->queued = (int64_t)S->queue.size();
    out->running = S->running;
    out->p50_ms = percentile(S->latencies, 0.50);
    out->p99_ms = percentile(S->latencies, 0.99);
    out->mean_queue_ms = S->queue_times.empty() ? 0.0 : [&] {
        double a = 0;
        for (double q : S->queue_times) a += q;
        return a / (double)S->queue_times.size();
    }();
    out->span_ms = S->completed ? S->last_done - S->first_submit : 0.0;
    return FMMHIP_OK;
}

}  // extern "C"
