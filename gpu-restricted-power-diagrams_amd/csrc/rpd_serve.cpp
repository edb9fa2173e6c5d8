// rpd_serve.cpp -- request loop for streams of diagram solves (the inner step
// of Lloyd relaxation or semi-discrete transport) against resident domains.
//
// No counterpart in the reference, whose VoroAlgo runs one launch at a time
// with the grid kept resident across launches (voronoi.h:126-131) and its
// per-launch buffers reset after each one (voronoi.cpp:1401).
//
// Design: a FIFO of submitted requests and `max_inflight` batch slots.  A
// slot is a solver context of its own (two HIP streams and resident buffers,
// which grow to the largest batch it has served and are reused).  HIP maps
// streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default): with the
// default 2 slots every stream has a queue of its own, so no slot's kernels
// wait in order behind another slot's.  A
// free slot takes the longest run of requests at the head of the queue that
// share the head's parameters and whose device footprints fit in what the
// HBM budget has left (at most max_batch requests and kBatchMaxSeeds seeds),
// and solves them as one packed batch (rpd_solve_batch): every kernel of a
// pass is one launch over the cells of all of them.  Requests are admitted strictly in arrival order, so
// a large request is never starved by small ones behind it, and a request
// larger than the whole budget runs alone.  With two slots, the next batch
// is admitted while the previous one finishes its re-queue tail, so the
// device stays busy across batch boundaries.  Every cell of a batch is
// computed as in a solve of its request alone, so a request's result does
// not depend on what it was batched with.

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

int rpd_ctx_device(const rpd_ctx *ctx);                            // rpd_api.hip
int rpd_create_streams(int device, int n_streams, rpd_ctx **out);  // rpd_api.hip

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
// Seeds of one batch: requests are packed up to this total (a larger request
// runs alone), so that the queue spreads over the slots and each batch fills
// the device in its first pass while its re-queue tail overlaps the other
// slots' work.
constexpr int64_t kBatchMaxSeeds = int64_t(1) << 19;

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

    static bool same_params(const rpd_params &a, const rpd_params &b) {
        return std::memcmp(&a, &b, sizeof(rpd_params)) == 0;
    }

    void worker(int s) {
        rpd_ctx *ctx = slots[size_t(s)];
        std::vector<Request *> b;
        std::vector<rpd_domain *> doms;
        std::vector<const double *> xs;
        std::vector<int64_t> ns;
        for (;;) {
            int64_t bytes = 0;
            b.clear();
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
                if (stop) return;
                // the head, then the requests behind it with the same
                // parameters while the footprints fit
                int64_t n_total = 0;
                while (!queue.empty() && int(b.size()) < cfg.max_batch) {
                    Request *h = queue.front();
                    if (!b.empty()) {
                        if (!same_params(h->prm, b.front()->prm)) break;
                        if (inflight_bytes + bytes + h->footprint > cfg.hbm_budget_bytes) break;
                        if (n_total + h->n > kBatchMaxSeeds) break;
                    }
                    queue.pop_front();
                    b.push_back(h);
                    bytes += h->footprint;
                    n_total += h->n;
                }
                inflight_bytes += bytes;
                ++running;
            }
            doms.clear();
            xs.clear();
            ns.clear();
            for (Request *r : b) {
                doms.push_back(r->dom);
                xs.push_back(r->seeds.data());
                ns.push_back(r->n);
            }
            int rc = rpd_solve_batch(ctx, int32_t(b.size()), doms.data(), xs.data(), ns.data(), &b.front()->prm,
                                     nullptr);
            for (size_t i = 0; i < b.size(); ++i) {
                Request *r = b[i];
                int rr = rc;
                int64_t first = 0, count = 0;
                if (rr == RPD_OK) rr = rpd_batch_range(ctx, int32_t(i), &first, &count);
                if (rr == RPD_OK) {
                    r->out.assign(size_t(r->n) * 4, 0.0);
                    rr = rpd_read_moments_range(ctx, first, count, r->out.data());
                }
                r->rc = rr;
            }
            const Clock::time_point t_done = Clock::now();
            for (Request *r : b) {
                r->lat_ms = std::chrono::duration<double, std::milli>(t_done - r->t_submit).count();
                std::vector<double>().swap(r->seeds);
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                inflight_bytes -= bytes;
                --running;
                for (Request *r : b) done.push_back(r);
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
    s->cfg.max_inflight = cfg && cfg->max_inflight > 0 ? std::min(cfg->max_inflight, 16) : 2;
    s->cfg.max_batch = cfg && cfg->max_batch > 0 ? std::min(cfg->max_batch, RPD_BATCH_MAX_REQUESTS) : 256;
    for (int i = 0; i < s->cfg.max_inflight; ++i) {
        rpd_ctx *c = nullptr;
        // two streams per slot: the default two slots hold the device's
        // four hardware queues, one stream each
        const int rc = rpd_create_streams(dev, 2, &c);
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
    // served requests return moments only: no Hessian
    r->prm.H = 0;
    r->prm.HC = 1;
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
