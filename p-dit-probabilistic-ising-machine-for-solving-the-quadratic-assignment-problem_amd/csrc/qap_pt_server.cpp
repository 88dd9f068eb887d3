// qap_pt_server.cpp -- serving loop: continuous batching of QAP solve requests on one GPU.
//
// Requests are copied at submit time and queued.  A worker thread (the only thread that
// calls HIP) admits queued requests in arrival order into pools -- up to max_batch
// requests per pool, up to max_pools pools in flight -- as long as their device
// footprint fits the budget (explicit, or free HBM minus a reserve).  Each pool runs on
// its own stream with its whole sweep budget enqueued as graph launches, so pools
// admitted while others are running overlap on the device.  Completion is detected with
// an event query; results stay until polled.  A request's trajectory depends only on
// its own inputs and seed (Philox counters are keyed by the request's seed and replica
// index, never by slot or pool), so its result is the same alone and in any mix.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/qap_pt.h"
#include "qap_pt_internal.h"

using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    int32_t n = 0;
    std::vector<int32_t> F, D, init, best_perm;
    std::vector<double> betas;
    qap_pt_params par{};
    Clock::time_point submitted;
    int32_t status = QAP_SRV_QUEUED;
    qap_pt_result res{};
    int64_t bytes = 0;
};

struct Flight {
    qap_pt_pool *pool = nullptr;
    hipEvent_t done = nullptr;
    std::vector<uint64_t> tickets;
    int64_t bytes = 0;
};

}  // namespace

struct qap_srv {
    qap_srv_config cfg{};
    std::mutex mu;
    std::condition_variable cv;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Request> reqs;
    std::vector<Flight> flights;
    uint64_t next_ticket = 1;
    int64_t bytes_in_use = 0;
    int64_t n_done = 0;
    bool stop = false;
    std::thread worker;
};

namespace {

qap_problem problem_of(const Request &r) { return qap_problem{r.n, r.F.data(), r.D.data()}; }

// Finished pools: copy results out, release the pool.  Called with the lock held.
void collect(qap_srv *s) {
    for (size_t f = 0; f < s->flights.size();) {
        Flight &fl = s->flights[f];
        const hipError_t q = hipEventQuery(fl.done);
        if (q == hipErrorNotReady) {
            ++f;
            continue;
        }
        const auto now = Clock::now();
        for (size_t k = 0; k < fl.tickets.size(); ++k) {
            Request &r = s->reqs[fl.tickets[k]];
            r.res = qap_pt_result{};
            r.res.best_perm = r.best_perm.data();
            int rc = (q == hipSuccess) ? qap::pool_result_internal(fl.pool, (int)k, &r.res)
                                       : QAP_PT_E_DEVICE;
            r.res.best_perm = nullptr;
            r.res.seconds = std::chrono::duration<double>(now - r.submitted).count();
            r.status = rc ? rc : QAP_SRV_DONE;
            ++s->n_done;
        }
        (void)hipEventDestroy(fl.done);
        qap::pool_destroy_internal(fl.pool);
        s->bytes_in_use -= fl.bytes;
        s->flights.erase(s->flights.begin() + (long)f);
    }
}

int64_t available_bytes(qap_srv *s) {
    if (s->cfg.device_budget > 0) return s->cfg.device_budget - s->bytes_in_use;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
    return (int64_t)free_b - s->cfg.reserve_bytes;
}

// Admit queued requests (FIFO) while pools and memory allow.  Lock held.
void admit(qap_srv *s) {
    while (!s->pending.empty() && (int)s->flights.size() < s->cfg.max_pools) {
        if (s->cfg.batch_window_s > 0 && (int)s->pending.size() < s->cfg.max_batch) {
            const Request &oldest = s->reqs[s->pending.front()];
            if (std::chrono::duration<double>(Clock::now() - oldest.submitted).count() <
                s->cfg.batch_window_s)
                return;
        }
        int64_t avail = available_bytes(s);
        std::vector<uint64_t> take;
        int64_t bytes = 0;
        for (uint64_t t : s->pending) {
            if ((int)take.size() >= s->cfg.max_batch) break;
            const int64_t b = s->reqs[t].bytes;
            if (bytes + b > avail) break;
            take.push_back(t);
            bytes += b;
        }
        if (take.empty()) {
            if (s->flights.empty()) {  // nothing will free memory: this request cannot run
                Request &r = s->reqs[s->pending.front()];
                r.status = QAP_PT_E_NOMEM;
                r.res.seconds = std::chrono::duration<double>(Clock::now() - r.submitted).count();
                s->pending.pop_front();
                continue;
            }
            return;  // wait for a running pool to finish
        }
        for (size_t k = 0; k < take.size(); ++k) s->pending.pop_front();
        std::vector<qap_problem> probs;
        std::vector<qap_pt_params> pars;
        for (uint64_t t : take) {
            probs.push_back(problem_of(s->reqs[t]));
            pars.push_back(s->reqs[t].par);
        }
        Flight fl;
        fl.tickets = take;
        fl.bytes = bytes;
        int rc = qap::pool_create_internal((int32_t)take.size(), probs.data(), pars.data(), &fl.pool);
        if (rc == QAP_PT_OK && hipEventCreateWithFlags(&fl.done, hipEventDisableTiming) != hipSuccess)
            rc = QAP_PT_E_DEVICE;
        if (rc == QAP_PT_OK) rc = qap::pool_launch_async(fl.pool, fl.done);
        if (rc != QAP_PT_OK) {
            if (fl.done) (void)hipEventDestroy(fl.done);
            if (fl.pool) qap::pool_destroy_internal(fl.pool);
            for (uint64_t t : take) s->reqs[t].status = rc;
            continue;
        }
        for (uint64_t t : take) s->reqs[t].status = QAP_SRV_RUNNING;
        s->bytes_in_use += bytes;
        s->flights.push_back(fl);
    }
}

void worker_main(qap_srv *s) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        if (s->stop) {
            for (uint64_t t : s->pending) s->reqs[t].status = QAP_PT_E_BUSY;
            s->pending.clear();
        }
        collect(s);
        if (s->stop && s->flights.empty()) break;
        admit(s);
        s->cv.wait_for(lk, std::chrono::microseconds(s->flights.empty() ? 2000 : 200));
    }
}

}  // namespace

extern "C" {

int qap_srv_create(const qap_srv_config *cfg, qap_srv **out) {
    if (!out) return qap::fail(QAP_PT_E_INVAL, "out is NULL");
    *out = nullptr;
    qap_srv_config c{};
    c.max_pools = 4;
    c.max_batch = 64;
    c.reserve_bytes = (int64_t)1 << 30;
    if (cfg) c = *cfg;
    if (c.max_pools < 1 || c.max_batch < 1 || c.device_budget < 0 || c.reserve_bytes < 0 ||
        c.batch_window_s < 0)
        return qap::fail(QAP_PT_E_INVAL, "invalid server configuration");
    try {
        qap_srv *s = new qap_srv();
        s->cfg = c;
        s->worker = std::thread(worker_main, s);
        *out = s;
    } catch (...) {
        return qap::fail(QAP_PT_E_NOMEM, "could not start the server");
    }
    return QAP_PT_OK;
}

int qap_srv_submit(qap_srv *s, const qap_problem *prob, const qap_pt_params *par, uint64_t *ticket) {
    if (!s || !prob || !par || !ticket) return qap::fail(QAP_PT_E_INVAL, "NULL argument");
    if (prob->n < 1 || prob->n > QAP_PT_MAX_N || !prob->F || !prob->D || !par->beta_ln ||
        par->R < 1 || par->T < 1)
        return qap::fail(QAP_PT_E_INVAL, "invalid request");
    try {
        Request r;
        r.n = prob->n;
        const size_t nn = (size_t)prob->n * prob->n;
        r.F.assign(prob->F, prob->F + nn);
        r.D.assign(prob->D, prob->D + nn);
        r.betas.assign(par->beta_ln, par->beta_ln + par->R);
        r.par = *par;
        r.par.beta_ln = nullptr;
        if (par->init_perm) r.init.assign(par->init_perm, par->init_perm + prob->n);
        r.par.init_perm = nullptr;
        r.best_perm.assign(prob->n, 0);
        qap_problem pb{r.n, r.F.data(), r.D.data()};
        qap_pt_params pp = r.par;
        pp.beta_ln = r.betas.data();
        r.bytes = qap::pool_device_bytes(1, &pb, &pp);
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return qap::fail(QAP_PT_E_BUSY, "server is shutting down");
        if (s->cfg.device_budget > 0 && r.bytes > s->cfg.device_budget)
            return qap::fail(QAP_PT_E_NOMEM, "request needs %lld device bytes, budget is %lld",
                             (long long)r.bytes, (long long)s->cfg.device_budget);
        r.submitted = Clock::now();
        const uint64_t t = s->next_ticket++;
        auto &slot = s->reqs[t];
        slot = std::move(r);
        // pointers into the request's own storage (stable: the vectors are not resized)
        slot.par.beta_ln = slot.betas.data();
        slot.par.init_perm = slot.init.empty() ? nullptr : slot.init.data();
        s->pending.push_back(t);
        *ticket = t;
    } catch (...) {
        return qap::fail(QAP_PT_E_NOMEM, "host allocation failed");
    }
    s->cv.notify_one();
    return QAP_PT_OK;
}

static int poll_locked(qap_srv *s, uint64_t ticket, qap_pt_result *res, int32_t *status) {
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return qap::fail(QAP_PT_E_INVAL, "unknown ticket %llu",
                                              (unsigned long long)ticket);
    Request &r = it->second;
    *status = r.status;
    if (r.status == QAP_SRV_DONE || r.status < 0) {
        if (res) {
            int32_t *dst = res->best_perm;
            *res = r.res;
            res->best_perm = dst;
            if (dst && r.status == QAP_SRV_DONE)
                std::copy(r.best_perm.begin(), r.best_perm.end(), dst);
        }
        s->reqs.erase(it);
    }
    return QAP_PT_OK;
}

int qap_srv_poll(qap_srv *s, uint64_t ticket, qap_pt_result *res, int32_t *status) {
    if (!s || !status) return qap::fail(QAP_PT_E_INVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    return poll_locked(s, ticket, res, status);
}

int qap_srv_wait(qap_srv *s, uint64_t ticket, double timeout_s, qap_pt_result *res,
                 int32_t *status) {
    if (!s || !status) return qap::fail(QAP_PT_E_INVAL, "NULL argument");
    const auto t0 = Clock::now();
    while (true) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            auto it = s->reqs.find(ticket);
            if (it == s->reqs.end())
                return qap::fail(QAP_PT_E_INVAL, "unknown ticket %llu", (unsigned long long)ticket);
            if (it->second.status == QAP_SRV_DONE || it->second.status < 0)
                return poll_locked(s, ticket, res, status);
            if (timeout_s > 0 &&
                std::chrono::duration<double>(Clock::now() - t0).count() > timeout_s) {
                *status = it->second.status;
                return QAP_PT_OK;
            }
        }
        std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
}

int qap_srv_stats(qap_srv *s, int64_t *queued, int64_t *running, int64_t *done,
                  int64_t *device_bytes) {
    if (!s) return qap::fail(QAP_PT_E_INVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    int64_t run = 0;
    for (auto &f : s->flights) run += (int64_t)f.tickets.size();
    if (queued) *queued = (int64_t)s->pending.size();
    if (running) *running = run;
    if (done) *done = s->n_done;
    if (device_bytes) *device_bytes = s->bytes_in_use;
    return QAP_PT_OK;
}

int qap_srv_destroy(qap_srv *s) {
    if (!s) return QAP_PT_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_one();
    if (s->worker.joinable()) s->worker.join();
    delete s;
    return QAP_PT_OK;
}

}  // extern "C"
