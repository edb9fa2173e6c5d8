// rpd_serve.cpp -- request loop for streams of diagram solves (the inner step
// of Lloyd relaxation or semi-discrete transport) against resident domains.
//
// No counterpart in the reference, whose VoroAlgo runs one launch at a time
// with the grid kept resident across launches (voronoi.h:126-131) and its
// per-launch buffers reset after each one (voronoi.cpp:1401).
//
// Design: a FIFO of submitted requests and `max_inflight` device slots.  A
// slot is a solver context of its own (its own HIP stream and resident
// buffers, which grow to the largest request it has served and are reused).
// A slot takes the request at the head of the queue as soon as the request's
// device footprint fits in what the HBM budget has left; requests are
// admitted strictly in arrival order, so a large request is never starved by
// small ones behind it.  Admitted requests run concurrently on their slots'
// streams, so the cell work of several requests shares the device.  Every
// request is solved by the same kernels on its own buffers, with no state
// shared between requests, so its result does not depend on what else runs.

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "rpd.h"

int rpd_ctx_device(const rpd_ctx *ctx);  // rpd_api.hip

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t ticket = 0;
    rpd_domain *dom = nullptr;
    std::vector<double> seeds;  // 4n
    int64_t n = 0;
    rpd_params prm{};
    int64_t footprint = 0;
    Clock::time_point t_submit;
    // result
    std::vector<double> out;  // 4n
    int rc = RPD_OK;
    double lat_ms = 0;
};

// Device bytes one solve of n seeds holds at its peak, per seed: the seeds
// (32) and their two sorted copies (64), permutation (4), outputs, status and
// K used (40), the pass-1 id list (4), six failure lists (72), resume slots
// and last neighbours (8), cover flags (1), binning scratch (16), kNN lists
// and cell records of pass 1 and, bounded by the same size, of the re-queue
// passes that run beside it; plus the kept states of pass 1's cells whose
// list ran out (room for 2/5 of the cells, plus 1024).
int64_t footprint_bytes(int64_t n, const rpd_params &p) {
    const int64_t K = std::max(1, p.K), P = std::max(1, p.P);
    const int64_t S = p.V <= 64 ? 1 : (p.V <= 128 ? 2 : 4);
    const int64_t rec = 16 + P * 36 + 256 * S;
    return n * (32 + 64 + 4 + 40 + 4 + 72 + 8 + 1 + 16 + 2 * (4 * K + rec)) +
           (std::min(n, n * 2 / 5 + 1024)) * rec + (int64_t(1) << 20);
}

}  // namespace

struct rpd_server {
    int device = 0;
    rpd_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;      // waiting, arrival order
    std::deque<Request *> done;       // finished, completion order
    int64_t next_ticket = 0;
    int64_t inflight_bytes = 0;
    int running = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    std::vector<rpd_ctx *> slots;

    void worker(int s) {
        rpd_ctx *ctx = slots[size_t(s)];
        for (;;) {
            Request *r = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] {
                    if (stop) return true;
                    if (queue.empty()) return false;
                    const Request *h = queue.front();
                    // admission by footprint; a request larger than the whole
                    // budget runs alone
                    return inflight_bytes + h->footprint <= cfg.hbm_budget_bytes || running == 0;
                });
                if (stop && queue.empty()) return;
                if (stop) return;
                r = queue.front();
                queue.pop_front();
                inflight_bytes += r->footprint;
                ++running;
            }
            r->out.assign(size_t(r->n) * 4, 0.0);
            int rc = rpd_set_seeds(ctx, r->seeds.data(), r->n);
            // served requests return moments only: no Hessian
            r->prm.H = 0;
            if (rc == RPD_OK) rc = rpd_solve(ctx, r->dom, &r->prm, nullptr);
            if (rc == RPD_OK) rc = rpd_read_moments(ctx, r->out.data());
            r->rc = rc;
            r->lat_ms = std::chrono::duration<double, std::milli>(Clock::now() - r->t_submit).count();
            std::vector<double>().swap(r->seeds);
            {
                std::lock_guard<std::mutex> lk(mu);
                inflight_bytes -= r->footprint;
                --running;
                done.push_back(r);
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

int rpd_server_create(rpd_ctx *ctx, const rpd_server_config *cfg, rpd_server **out) {
    if (!ctx || !out) return RPD_E_ARG;
    *out = nullptr;
    const int dev = rpd_ctx_device(ctx);  // the slots run on the device of ctx
    auto *s = new (std::nothrow) rpd_server();
    if (!s) return RPD_E_NOMEM;
    s->device = dev;
    s->cfg.hbm_budget_bytes = cfg && cfg->hbm_budget_bytes > 0 ? cfg->hbm_budget_bytes : (int64_t(64) << 30);
    s->cfg.max_inflight = cfg && cfg->max_inflight > 0 ? std::min(cfg->max_inflight, 16) : 4;
    for (int i = 0; i < s->cfg.max_inflight; ++i) {
        rpd_ctx *c = nullptr;
        const int rc = rpd_create(dev, &c);
        if (rc != RPD_OK) {
            for (rpd_ctx *q : s->slots) rpd_destroy(q);
            delete s;
            return rc;
        }
        s->slots.push_back(c);
    }
    for (int i = 0; i < s->cfg.max_inflight; ++i) s->workers.emplace_back([s, i] { s->worker(i); });
    *out = s;
    return RPD_OK;
}

void rpd_server_destroy(rpd_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    for (rpd_ctx *c : s->slots) rpd_destroy(c);
    for (Request *r : s->queue) delete r;
    for (Request *r : s->done) delete r;
    delete s;
}

int rpd_server_submit(rpd_server *s, rpd_domain *dom, const double *xyzw, int64_t n, const rpd_params *prm,
                      int64_t *ticket) {
    if (!s || !xyzw || n < 2 || !ticket) return RPD_E_ARG;
    auto *r = new (std::nothrow) Request();
    if (!r) return RPD_E_NOMEM;
    r->dom = dom;
    r->n = n;
    if (prm)
        r->prm = *prm;
    else
        rpd_default_params(&r->prm);
    r->seeds.assign(xyzw, xyzw + size_t(n) * 4);
    r->footprint = footprint_bytes(n, r->prm);
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            delete r;
            return RPD_E_STATE;
        }
        r->ticket = s->next_ticket++;
        *ticket = r->ticket;
        s->queue.push_back(r);
    }
    s->cv_work.notify_all();
    return RPD_OK;
}

int rpd_server_poll(rpd_server *s, int wait, int64_t *done_ticket, double *out4n, int64_t cap_n, double *lat_ms) {
    if (!s || !done_ticket) return RPD_E_ARG;
    *done_ticket = -1;
    Request *r = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        if (wait)
            s->cv_done.wait(lk, [&] { return !s->done.empty() || (s->queue.empty() && s->running == 0); });
        if (s->done.empty()) return RPD_OK;
        r = s->done.front();
        s->done.pop_front();
    }
    *done_ticket = r->ticket;
    if (lat_ms) *lat_ms = r->lat_ms;
    int rc = r->rc;
    if (rc == RPD_OK && out4n) {
        if (cap_n < r->n)
            rc = RPD_E_ARG;
        else
            std::memcpy(out4n, r->out.data(), size_t(r->n) * 4 * sizeof(double));
    }
    delete r;
    return rc;
}

int rpd_server_drain(rpd_server *s) {
    if (!s) return RPD_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->queue.empty() && s->running == 0; });
    return RPD_OK;
}

}  // extern "C"
