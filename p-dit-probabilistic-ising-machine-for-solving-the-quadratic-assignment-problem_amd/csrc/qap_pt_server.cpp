// qap_pt_server.cpp -- serving loop: continuous batching of QAP solve requests on one GPU.
//
// Requests are copied at submit time and queued.  A worker thread (the only thread that
// calls HIP) keeps slot banks resident on the device: pools of fixed-capacity instance
// slots, one bank per size class (n <= 64, 128, 256, 512, 1024; up to 32 or 64
// replicas), created when a request of that class arrives and the device budget
// (explicit, or free HBM minus a reserve) holds the bank.  Each bank runs its loaded
// requests together, a few exchange intervals per launch (about 8 ms of work) on its own
// stream; between launches finished requests are released and queued requests are
// loaded into free slots, so requests that arrive while others run join the very next
// launch and share every kernel with them.  A request that fits no bank class, or finds
// no bank room while the budget still holds a pool of its own, runs as a separate pool.
// A request's trajectory depends only on its own inputs and seed (Philox counters are
// keyed by the request's seed and replica index, never by slot or bank), so its result
// is the same alone and in any mix.  Banks idle for a second are freed.
#include <algorithm>
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

struct Bank {
    qap_pt_pool *pool = nullptr;
    int n_cap = 0, rt_cap = 0;
    hipEvent_t done = nullptr;
    bool inflight = false;
    bool fresh = false;                // a launch ended since the state was last read
    int intervals = 4;                 // per launch, adapted to ~8 ms of device work
    Clock::time_point launched, idle_since;
    std::vector<uint64_t> ticket;      // per slot, 0 = free
    std::vector<Clock::time_point> started;
    std::vector<uint8_t> halting;      // per slot: timed out, halt requested
    int64_t bytes = 0;
};

constexpr int kBankClasses[] = {64, 128, 256, 512, 1024};
constexpr double kLaunchTarget = 0.008;  // seconds of device work per bank launch
constexpr double kBankIdleFree = 1.0;    // seconds an empty bank is kept

}  // namespace

struct qap_srv {
    qap_srv_config cfg{};
    std::mutex mu;
    std::condition_variable cv;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Request> reqs;
    std::vector<Flight> flights;
    std::vector<Bank> banks;
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

void finish(qap_srv *s, Request &r, int status) {
    r.status = status;
    r.res.seconds = std::chrono::duration<double>(Clock::now() - r.submitted).count();
    ++s->n_done;
}

// Bank class of a request: the smallest size class that holds n, and 32 or 64 replicas;
// n_cap 0 when it fits no class.
void bank_class(const Request &r, int &n_cap, int &rt_cap) {
    n_cap = 0;
    rt_cap = 0;
    const int64_t RT = (int64_t)r.par.R * r.par.T;
    if (RT > 64) return;
    rt_cap = RT <= 32 ? 32 : 64;
    for (int c : kBankClasses)
        if (r.n <= c) {
            n_cap = c;
            return;
        }
}

// Launches that ended: finished requests leave their slots with their results; requests
// past their timeout are halted (they finish after the next launch).  Lock held.
void service_banks(qap_srv *s) {
    const auto now = Clock::now();
    std::vector<int> done;
    std::vector<qap::InstState> st;
    for (Bank &b : s->banks) {
        if (b.inflight) {
            const hipError_t q = hipEventQuery(b.done);
            if (q == hipErrorNotReady) continue;
            b.inflight = false;
            b.fresh = true;
            const double t = std::chrono::duration<double>(now - b.launched).count();
            if (q == hipSuccess && t > 0)  // keep about kLaunchTarget of work per launch
                b.intervals = std::max(1, std::min(1024, (int)(b.intervals * kLaunchTarget / t + 0.5)));
        }
        if (!b.fresh) continue;
        b.fresh = false;
        int rc = qap::slots_finished(b.pool, done, st);
        for (int slot = 0; slot < (int)b.ticket.size() && rc; ++slot)
            if (b.ticket[slot]) done.push_back(slot);  // device error: fail every request
        for (int slot : done) {
            Request &r = s->reqs[b.ticket[slot]];
            r.res = qap_pt_result{};
            r.res.best_perm = r.best_perm.data();
            const int rr = rc ? rc : qap::pool_result_internal(b.pool, slot, &r.res);
            r.res.best_perm = nullptr;
            finish(s, r, rr ? rr : QAP_SRV_DONE);
            (void)qap::slots_release(b.pool, slot);
            b.ticket[slot] = 0;
            b.halting[slot] = 0;
        }
        for (int slot = 0; slot < (int)b.ticket.size() && !rc; ++slot) {
            if (!b.ticket[slot] || b.halting[slot]) continue;
            const Request &r = s->reqs[b.ticket[slot]];
            if (r.par.timeout_s > 0 &&
                std::chrono::duration<double>(now - b.started[slot]).count() > r.par.timeout_s) {
                (void)qap::slots_halt(b.pool, slot);
                b.halting[slot] = 1;
            }
        }
    }
    for (Bank &b : s->banks) {
        if (b.inflight || qap::slots_in_use(b.pool) > 0) {
            b.idle_since = Clock::time_point{};
        } else if (b.idle_since == Clock::time_point{}) {
            b.idle_since = now;
        }
    }
    // empty banks idle for a while give their memory back
    for (size_t k = 0; k < s->banks.size();) {
        Bank &b = s->banks[k];
        if (!b.inflight && b.idle_since != Clock::time_point{} &&
            (s->stop || std::chrono::duration<double>(now - b.idle_since).count() > kBankIdleFree)) {
            (void)hipEventDestroy(b.done);
            qap::pool_destroy_internal(b.pool);
            s->bytes_in_use -= b.bytes;
            s->banks.erase(s->banks.begin() + (long)k);
        } else {
            ++k;
        }
    }
}

// A bank with a free slot for this class that is between launches, creating one when
// the budget holds it.  nullptr: wait (or use a pool of its own).  Lock held.
Bank *bank_for(qap_srv *s, int n_cap, int rt_cap, int &err) {
    err = QAP_PT_OK;
    bool busy_room = false;
    for (Bank &b : s->banks) {
        if (b.n_cap != n_cap || b.rt_cap != rt_cap) continue;
        if (qap::slots_free(b.pool) < 0) continue;
        if (!b.inflight) return &b;
        busy_room = true;
    }
    if (busy_room) return nullptr;  // room opens at the end of the current launch
    if ((int)(s->banks.size() + s->flights.size()) >= s->cfg.max_pools) return nullptr;
    const int64_t per_slot = qap::slots_device_bytes(1, n_cap, rt_cap, rt_cap, rt_cap);
    const int64_t avail = available_bytes(s);
    // a bank takes at most a quarter of what is available, and at least one slot
    int slots = (int)std::min<int64_t>(s->cfg.max_batch, (avail / 4) / std::max<int64_t>(per_slot, 1));
    if (slots < 1) {
        if (per_slot > avail) return nullptr;
        slots = 1;
    }
    Bank b;
    b.n_cap = n_cap;
    b.rt_cap = rt_cap;
    b.bytes = qap::slots_device_bytes(slots, n_cap, rt_cap, rt_cap, rt_cap);
    err = qap::slots_create(slots, n_cap, rt_cap, rt_cap, rt_cap, &b.pool);
    if (err == QAP_PT_OK && hipEventCreateWithFlags(&b.done, hipEventDisableTiming) != hipSuccess)
        err = qap::fail(QAP_PT_E_DEVICE, "event creation failed");
    if (err != QAP_PT_OK) {
        if (b.pool) qap::pool_destroy_internal(b.pool);
        return nullptr;
    }
    b.ticket.assign(slots, 0);
    b.started.assign(slots, Clock::time_point{});
    b.halting.assign(slots, 0);
    s->bytes_in_use += b.bytes;
    s->banks.push_back(std::move(b));
    return &s->banks.back();
}

// A request in a pool of its own (sizes no bank class holds).  Lock held.
int start_flight(qap_srv *s, uint64_t t) {
    Request &r = s->reqs[t];
    Flight fl;
    fl.tickets = {t};
    fl.bytes = r.bytes;
    qap_problem pb = problem_of(r);
    int rc = qap::pool_create_internal(1, &pb, &r.par, &fl.pool);
    if (rc == QAP_PT_OK && hipEventCreateWithFlags(&fl.done, hipEventDisableTiming) != hipSuccess)
        rc = qap::fail(QAP_PT_E_DEVICE, "event creation failed");
    if (rc == QAP_PT_OK) rc = qap::pool_launch_async(fl.pool, fl.done);
    if (rc != QAP_PT_OK) {
        if (fl.done) (void)hipEventDestroy(fl.done);
        if (fl.pool) qap::pool_destroy_internal(fl.pool);
        return rc;
    }
    r.status = QAP_SRV_RUNNING;
    s->bytes_in_use += fl.bytes;
    s->flights.push_back(fl);
    return QAP_PT_OK;
}

// Admit queued requests in arrival order while slots and memory allow; a request that
// has to wait does not hold back later ones of other classes.  Lock held.
void admit(qap_srv *s) {
    if (s->pending.empty()) return;
    if (s->cfg.batch_window_s > 0 && (int)s->pending.size() < s->cfg.max_batch) {
        const Request &oldest = s->reqs[s->pending.front()];
        if (std::chrono::duration<double>(Clock::now() - oldest.submitted).count() <
            s->cfg.batch_window_s)
            return;
    }
    const auto now = Clock::now();
    for (auto it = s->pending.begin(); it != s->pending.end();) {
        const uint64_t t = *it;
        Request &r = s->reqs[t];
        int n_cap = 0, rt_cap = 0, err = QAP_PT_OK;
        bank_class(r, n_cap, rt_cap);
        Bank *b = n_cap ? bank_for(s, n_cap, rt_cap, err) : nullptr;
        if (b) {
            const int slot = qap::slots_free(b->pool);
            const qap_problem pb = problem_of(r);
            const int rc = qap::slots_load(b->pool, slot, pb, r.par);
            if (rc) {
                finish(s, r, rc);
            } else {
                b->ticket[slot] = t;
                b->started[slot] = now;
                r.status = QAP_SRV_RUNNING;
            }
            it = s->pending.erase(it);
            continue;
        }
        if (err != QAP_PT_OK && err != QAP_PT_E_NOMEM) {  // the device refused a bank
            finish(s, r, err);
            it = s->pending.erase(it);
            continue;
        }
        // no bank room: a pool of its own when the budget holds it (sizes no bank class
        // holds, or memory too short for a bank)
        const bool alone = n_cap == 0 || qap::slots_device_bytes(1, n_cap, rt_cap, rt_cap, rt_cap) >
                                             available_bytes(s);
        if (alone && (int)(s->banks.size() + s->flights.size()) < s->cfg.max_pools &&
            r.bytes <= available_bytes(s)) {
            const int rc = start_flight(s, t);
            if (rc) finish(s, r, rc);
            it = s->pending.erase(it);
            continue;
        }
        if (alone && s->banks.empty() && s->flights.empty()) {  // nothing will free memory
            finish(s, r, QAP_PT_E_NOMEM);
            it = s->pending.erase(it);
            continue;
        }
        ++it;
    }
}

// Banks between launches with loaded requests start their next launch.  Lock held.
void launch_banks(qap_srv *s) {
    for (Bank &b : s->banks) {
        if (b.inflight || qap::slots_in_use(b.pool) == 0) continue;
        const int rc = qap::slots_launch(b.pool, b.intervals, b.done);
        if (rc) {  // fail the bank's requests
            for (int slot = 0; slot < (int)b.ticket.size(); ++slot)
                if (b.ticket[slot]) {
                    finish(s, s->reqs[b.ticket[slot]], rc);
                    (void)qap::slots_release(b.pool, slot);
                    b.ticket[slot] = 0;
                }
            continue;
        }
        b.inflight = true;
        b.launched = Clock::now();
    }
}

void worker_main(qap_srv *s) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        if (s->stop) {
            for (uint64_t t : s->pending) finish(s, s->reqs[t], QAP_PT_E_BUSY);
            s->pending.clear();
        }
        collect(s);
        service_banks(s);
        if (s->stop && s->flights.empty() && s->banks.empty()) break;
        admit(s);
        launch_banks(s);
        bool busy = !s->flights.empty();
        for (const Bank &b : s->banks) busy = busy || b.inflight;
        s->cv.wait_for(lk, std::chrono::microseconds(busy ? 100 : 2000));
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
    for (auto &b : s->banks) run += qap::slots_in_use(b.pool);
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
