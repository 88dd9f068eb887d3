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
//
// Locking: the server mutex guards only what client threads touch -- the incoming queue,
// the per-ticket records clients poll, and the published counters.  The worker takes new
// requests in one short critical section, does every upload, launch and result copy with
// the lock released, and publishes finished results in another; a request's clock starts
// when qap_srv_submit is entered, before any lock.
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
    bool wide = false;  // needs int64 force matrices: runs in a pool of its own
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

// What a client sees of a request: its status and, once final, its result.
struct Record {
    int32_t status = QAP_SRV_QUEUED;
    qap_pt_result res{};
    std::vector<int32_t> best_perm;
};

struct qap_srv {
    qap_srv_config cfg{};
    // ---- shared with client threads, guarded by mu
    std::mutex mu;
    std::condition_variable cv;       // the worker waits here for requests
    std::condition_variable done_cv;  // waiters wait here for results
    std::deque<std::pair<uint64_t, Request>> incoming;  // submitted, not yet taken
    std::unordered_map<uint64_t, Record> recs;
    uint64_t next_ticket = 1;
    int64_t n_done = 0;
    int64_t pub_queued = 0, pub_running = 0, pub_bytes = 0;  // the worker's counts
    bool stop = false;
    std::thread worker;
    // ---- the worker's own (no lock)
    std::deque<uint64_t> queue;                     // taken, waiting for admission
    std::unordered_map<uint64_t, Request> work;     // bodies of queued and running requests
    std::vector<uint64_t> outbox;                   // final since the last publish
    std::vector<uint64_t> started_run;              // became RUNNING since the last publish
    std::vector<Flight> flights;
    std::vector<Bank> banks;
    int64_t bytes_in_use = 0;
    bool stopping = false;
};

namespace {

qap_problem problem_of(const Request &r) { return qap_problem{r.n, r.F.data(), r.D.data()}; }

void finish(qap_srv *s, uint64_t t, int status);

// Finished pools: copy results out, release the pool.  Worker only, lock not held.
void collect(qap_srv *s) {
    for (size_t f = 0; f < s->flights.size();) {
        Flight &fl = s->flights[f];
        const hipError_t q = hipEventQuery(fl.done);
        if (q == hipErrorNotReady) {
            ++f;
            continue;
        }
        for (size_t k = 0; k < fl.tickets.size(); ++k) {
            Request &r = s->work[fl.tickets[k]];
            r.res = qap_pt_result{};
            r.res.best_perm = r.best_perm.data();
            int rc = (q == hipSuccess) ? qap::pool_result_internal(fl.pool, (int)k, &r.res)
                                       : QAP_PT_E_DEVICE;
            r.res.best_perm = nullptr;
            finish(s, fl.tickets[k], rc ? rc : QAP_SRV_DONE);
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

// A request's final status; its result goes out with the next publish.  Worker only.
void finish(qap_srv *s, uint64_t t, int status) {
    Request &r = s->work[t];
    r.status = status;
    r.res.seconds = std::chrono::duration<double>(Clock::now() - r.submitted).count();
    s->outbox.push_back(t);
}

// Bank class of a request: the smallest size class that holds n, and 32 or 64 replicas;
// n_cap 0 when it fits no class.
void bank_class(const Request &r, int &n_cap, int &rt_cap) {
    n_cap = 0;
    rt_cap = 0;
    const int64_t RT = (int64_t)r.par.R * r.par.T;
    if (RT > 64 || r.wide) return;  // banks hold int32 force matrices
    rt_cap = RT <= 32 ? 32 : 64;
    for (int c : kBankClasses)
        if (r.n <= c) {
            n_cap = c;
            return;
        }
}

// Launches that ended: finished requests leave their slots with their results; requests
// past their timeout are halted (they finish after the next launch).  Worker only.
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
            Request &r = s->work[b.ticket[slot]];
            r.res = qap_pt_result{};
            r.res.best_perm = r.best_perm.data();
            const int rr = rc ? rc : qap::pool_result_internal(b.pool, slot, &r.res);
            r.res.best_perm = nullptr;
            finish(s, b.ticket[slot], rr ? rr : QAP_SRV_DONE);
            (void)qap::slots_release(b.pool, slot);
            b.ticket[slot] = 0;
            b.halting[slot] = 0;
        }
        for (int slot = 0; slot < (int)b.ticket.size() && !rc; ++slot) {
            if (!b.ticket[slot] || b.halting[slot]) continue;
            const Request &r = s->work[b.ticket[slot]];
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
            (s->stopping || std::chrono::duration<double>(now - b.idle_since).count() > kBankIdleFree)) {
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
// the budget holds it.  nullptr: wait (or use a pool of its own).  Worker only.
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

// A request in a pool of its own (sizes no bank class holds).  Worker only.
int start_flight(qap_srv *s, uint64_t t) {
    Request &r = s->work[t];
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
    s->started_run.push_back(t);
    s->bytes_in_use += fl.bytes;
    s->flights.push_back(fl);
    return QAP_PT_OK;
}

// Admit queued requests in arrival order while slots and memory allow; a request that
// has to wait does not hold back later ones of other classes.  Worker only.
void admit(qap_srv *s) {
    if (s->queue.empty()) return;
    if (s->cfg.batch_window_s > 0 && (int)s->queue.size() < s->cfg.max_batch) {
        const Request &oldest = s->work[s->queue.front()];
        if (std::chrono::duration<double>(Clock::now() - oldest.submitted).count() <
            s->cfg.batch_window_s)
            return;
    }
    const auto now = Clock::now();
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        const uint64_t t = *it;
        Request &r = s->work[t];
        int n_cap = 0, rt_cap = 0, err = QAP_PT_OK;
        bank_class(r, n_cap, rt_cap);
        Bank *b = n_cap ? bank_for(s, n_cap, rt_cap, err) : nullptr;
        if (b) {
            const int slot = qap::slots_free(b->pool);
            const qap_problem pb = problem_of(r);
            const int rc = qap::slots_load(b->pool, slot, pb, r.par);
            if (rc) {
                finish(s, t, rc);
            } else {
                b->ticket[slot] = t;
                b->started[slot] = now;
                r.status = QAP_SRV_RUNNING;
                s->started_run.push_back(t);
            }
            it = s->queue.erase(it);
            continue;
        }
        if (err != QAP_PT_OK && err != QAP_PT_E_NOMEM) {  // the device refused a bank
            finish(s, t, err);
            it = s->queue.erase(it);
            continue;
        }
        // no bank room: a pool of its own when the budget holds it (sizes no bank class
        // holds, or memory too short for a bank)
        const bool alone = n_cap == 0 || qap::slots_device_bytes(1, n_cap, rt_cap, rt_cap, rt_cap) >
                                             available_bytes(s);
        if (alone && (int)(s->banks.size() + s->flights.size()) < s->cfg.max_pools &&
            r.bytes <= available_bytes(s)) {
            const int rc = start_flight(s, t);
            if (rc) finish(s, t, rc);
            it = s->queue.erase(it);
            continue;
        }
        if (alone && s->banks.empty() && s->flights.empty()) {  // nothing will free memory
            finish(s, t, QAP_PT_E_NOMEM);
            it = s->queue.erase(it);
            continue;
        }
        ++it;
    }
}

// Banks between launches with loaded requests start their next launch.  Worker only.
void launch_banks(qap_srv *s) {
    for (Bank &b : s->banks) {
        if (b.inflight || qap::slots_in_use(b.pool) == 0) continue;
        const int rc = qap::slots_launch(b.pool, b.intervals, b.done);
        if (rc) {  // fail the bank's requests
            for (int slot = 0; slot < (int)b.ticket.size(); ++slot)
                if (b.ticket[slot]) {
                    finish(s, b.ticket[slot], rc);
                    (void)qap::slots_release(b.pool, slot);
                    b.ticket[slot] = 0;
                }
            continue;
        }
        b.inflight = true;
        b.launched = Clock::now();
    }
}

// Hands finished results and status changes to the clients, with the worker's counts.
// Lock held.
void publish_locked(qap_srv *s) {
    for (uint64_t t : s->started_run) {
        auto it = s->recs.find(t);
        if (it != s->recs.end() && it->second.status == QAP_SRV_QUEUED)
            it->second.status = QAP_SRV_RUNNING;
    }
    s->started_run.clear();
    for (uint64_t t : s->outbox) {
        Request &r = s->work[t];
        Record &rec = s->recs[t];
        rec.status = r.status;
        rec.res = r.res;
        rec.best_perm = std::move(r.best_perm);
        s->work.erase(t);
        ++s->n_done;
    }
    s->outbox.clear();
    int64_t run = 0;
    for (auto &f : s->flights) run += (int64_t)f.tickets.size();
    for (auto &b : s->banks) run += qap::slots_in_use(b.pool);
    s->pub_queued = (int64_t)s->queue.size();
    s->pub_running = run;
    s->pub_bytes = s->bytes_in_use;
}

void worker_main(qap_srv *s) {
    while (true) {
        {  // take what was submitted
            std::lock_guard<std::mutex> lk(s->mu);
            for (auto &in : s->incoming) {
                s->queue.push_back(in.first);
                s->work.emplace(in.first, std::move(in.second));
            }
            s->incoming.clear();
            s->stopping = s->stop;
        }
        if (s->stopping) {
            for (uint64_t t : s->queue) finish(s, t, QAP_PT_E_BUSY);
            s->queue.clear();
        }
        // device work, lock released
        collect(s);
        service_banks(s);
        const bool exit = s->stopping && s->flights.empty() && s->banks.empty();
        if (!exit) {
            admit(s);
            launch_banks(s);
        }
        bool busy = !s->flights.empty();
        for (const Bank &b : s->banks) busy = busy || b.inflight;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            const bool any_done = !s->outbox.empty();
            publish_locked(s);
            if (any_done) s->done_cv.notify_all();
            if (exit) break;
            if (s->incoming.empty() && !s->stop)
                s->cv.wait_for(lk, std::chrono::microseconds(busy ? 100 : 2000));
        }
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
    const auto submitted = Clock::now();  // the request's clock starts before any lock
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
        if (par->init_perm) r.init.assign(par->init_perm, par->init_perm + prob->n);
        r.best_perm.assign(prob->n, 0);
        // pointers into the request's own storage (the vectors keep their buffers when
        // the request is moved)
        r.par.beta_ln = r.betas.data();
        r.par.init_perm = r.init.empty() ? nullptr : r.init.data();
        qap_problem pb{r.n, r.F.data(), r.D.data()};
        r.bytes = qap::pool_device_bytes(1, &pb, &r.par);
        r.wide = qap::wide_instance(pb);
        if (s->cfg.device_budget > 0 && r.bytes > s->cfg.device_budget)
            return qap::fail(QAP_PT_E_NOMEM, "request needs %lld device bytes, budget is %lld",
                             (long long)r.bytes, (long long)s->cfg.device_budget);
        r.submitted = submitted;
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return qap::fail(QAP_PT_E_BUSY, "server is shutting down");
        const uint64_t t = s->next_ticket++;
        s->recs.emplace(t, Record{});
        s->incoming.emplace_back(t, std::move(r));
        *ticket = t;
    } catch (...) {
        return qap::fail(QAP_PT_E_NOMEM, "host allocation failed");
    }
    s->cv.notify_one();
    return QAP_PT_OK;
}

static int poll_locked(qap_srv *s, uint64_t ticket, qap_pt_result *res, int32_t *status) {
    auto it = s->recs.find(ticket);
    if (it == s->recs.end()) return qap::fail(QAP_PT_E_INVAL, "unknown ticket %llu",
                                              (unsigned long long)ticket);
    Record &r = it->second;
    *status = r.status;
    if (r.status == QAP_SRV_DONE || r.status < 0) {
        if (res) {
            int32_t *dst = res->best_perm;
            *res = r.res;
            res->best_perm = dst;
            if (dst && r.status == QAP_SRV_DONE)
                std::copy(r.best_perm.begin(), r.best_perm.end(), dst);
        }
        s->recs.erase(it);
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
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        auto it = s->recs.find(ticket);
        if (it == s->recs.end())
            return qap::fail(QAP_PT_E_INVAL, "unknown ticket %llu", (unsigned long long)ticket);
        if (it->second.status == QAP_SRV_DONE || it->second.status < 0)
            return poll_locked(s, ticket, res, status);
        const double left =
            timeout_s > 0 ? timeout_s - std::chrono::duration<double>(Clock::now() - t0).count()
                          : 0.05;
        if (timeout_s > 0 && left <= 0) {
            *status = it->second.status;
            return QAP_PT_OK;
        }
        s->done_cv.wait_for(lk, std::chrono::duration<double>(std::min(left, 0.05)));
    }
}

int qap_srv_stats(qap_srv *s, int64_t *queued, int64_t *running, int64_t *done,
                  int64_t *device_bytes) {
    if (!s) return qap::fail(QAP_PT_E_INVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (queued) *queued = (int64_t)s->incoming.size() + s->pub_queued;
    if (running) *running = s->pub_running;
    if (done) *done = s->n_done;
    if (device_bytes) *device_bytes = s->pub_bytes;
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
