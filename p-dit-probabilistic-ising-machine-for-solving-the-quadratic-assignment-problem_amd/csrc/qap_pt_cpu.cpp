// qap_pt_cpu.cpp -- qap_pt_solve_cpu: the PT-SA solve on the host's cores, the drop-in for
// CPU-PIM's own solve (CPU-PIM/CPU_PIM.cpp:478-685: main -> setup -> threadToRun x T).
//
// It runs the Markov chain of qap_pt_solve_hip update for update: Philox4x32-10 draws on
// the same counters, the exact integer heat-bath weights and cutoff (heat_weight,
// heat_cutoff from qap_pt_device.h, the same functions the sweep kernels call), the same
// rank-2 force update, bubble-pass exchange, kick, goal test and best tracking.  Both entry
// points therefore return identical results for identical arguments, whatever the number
// of host threads.  No device call is made; the call works on a host without a GPU.
//
// Work split.  CPU-PIM runs one thread per ladder.  Here a ladder's replicas are spread
// over the threads instead (they are independent between exchange steps), so one ladder
// of 16 replicas uses 16 cores: every exchange interval, each thread advances its
// replicas by sweeps_per_exchange sweeps; thread 0 then runs the exchange step of every
// ladder, and a spinning barrier separates the phases.
//
// Per update the host keeps, besides p and p^-1, the force matrix I (int32 when the
// instance's bounds prove every entry, delta and update term fits, else int64, the
// device's hbm64 rule) and evaluates the n candidate deltas from row i, column a and the
// permuted diagonal of I.  The rank-2 update touches only the columns where the flow
// differences are nonzero (a compacted column list when they are few) and skips rows
// whose distance differences are zero.
#include <hip/hip_runtime.h>

#include <sched.h>
#include <stdlib.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <new>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/qap_pt.h"
#include "qap_pt_device.h"
#include "qap_pt_internal.h"

namespace {

using qap::fail;

constexpr double kLog2e = 1.44269504089;  // CPU_PIM.cpp:528

// Threads wait here between the sweep and exchange phases of an interval.
class SpinBarrier {
  public:
    explicit SpinBarrier(int n) : n_(n) {}
    void wait() {
        const unsigned g = gen_.load(std::memory_order_acquire);
        if (count_.fetch_add(1, std::memory_order_acq_rel) == n_ - 1) {
            count_.store(0, std::memory_order_relaxed);
            gen_.fetch_add(1, std::memory_order_release);
            return;
        }
        for (int spins = 0; gen_.load(std::memory_order_acquire) == g; ++spins)
            if (spins > 256) std::this_thread::yield();
    }

  private:
    const int n_;
    std::atomic<int> count_{0};
    std::atomic<unsigned> gen_{0};
};

struct Replica {
    int32_t *perm, *pinv, *best;
    int64_t E = 0, bestE = INT64_MAX;
    uint64_t accepted = 0, cells = 0;
    int32_t pos = 0;
};

struct Ladder {
    std::vector<int32_t> order;  // position -> local replica
    int32_t tss = 0;             // timeSinceSwap
    int64_t best = INT64_MAX;
};

// Per-thread buffers of one update.
template <typename IT>
struct Scratch {
    std::vector<IT> dE, u1, u2, v1, v2;
    std::vector<uint32_t> w;
    std::vector<int32_t> cols, rows;
    explicit Scratch(int n)
        : dE(n), u1(n), u2(n), v1(n), v2(n), w(n), cols(n), rows(n) {}
};

template <typename IT>
struct Solver {
    int n, R, T, RT, spe;
    bool sym;
    const qap_pt_params &par;
    const int32_t *F, *D;
    std::vector<int32_t> Ft, Dt, Fd, Dd;
    std::vector<float> beta2f;  // per ladder position
    std::vector<double> c20;
    std::vector<IT> I;          // RT force matrices, n x n each
    std::vector<int32_t> perms;  // perm, pinv, best per replica
    std::vector<Replica> reps;
    std::vector<Ladder> lads;
    uint32_t k0, k1;

    Solver(const qap_problem &pb, const qap_pt_params &pr)
        : n(pb.n), R(pr.R), T(pr.T), RT(pr.R * pr.T), spe(pr.sweeps_per_exchange), sym(false),
          par(pr), F(pb.F), D(pb.D), k0((uint32_t)pr.seed), k1((uint32_t)(pr.seed >> 32)) {}

    IT *Irep(int g) { return I.data() + (size_t)g * n * n; }

    void init_perm(uint32_t stream, uint32_t tag, int32_t *p) const {
        // CPU_PIM.cpp:134-161: n random transpositions from the identity
        for (int i = 0; i < n; ++i) p[i] = i;
        for (int i = 0; i < n; ++i) {
            const qap::u32x4 x = qap::philox4x32_10(qap::u32x4{(uint32_t)i, stream, 0u, tag}, k0, k1);
            const int j = (int)(((uint64_t)x.x * (uint64_t)n) >> 32);
            std::swap(p[i], p[j]);
        }
    }

    // I[i][a] = -sum_j (D[i][j] F[a][p(j)] + D[j][i] F[p(j)][a])  (CPU_PIM.cpp:164-181).
    // The partial sums of absolute terms stay within the force bound, so the int32 path
    // accumulates in int32 without overflow.
    void build_force(const int32_t *p, IT *M) const {
        for (int i = 0; i < n; ++i) {
            IT *row = M + (size_t)i * n;
            std::fill(row, row + n, (IT)0);
            for (int j = 0; j < n; ++j) {
                const IT dij = (IT)D[(size_t)i * n + j], dji = (IT)Dt[(size_t)i * n + j];
                if (dij == 0 && dji == 0) continue;
                const int32_t *fc = Ft.data() + (size_t)p[j] * n;  // F[a][p(j)] over a
                const int32_t *fr = F + (size_t)p[j] * n;          // F[p(j)][a] over a
                for (int a = 0; a < n; ++a) row[a] -= dij * (IT)fc[a] + dji * (IT)fr[a];
            }
        }
    }

    int64_t energy(const int32_t *p) const {
        int64_t e = 0;
        for (int i = 0; i < n; ++i) {
            const int32_t *fr = F + (size_t)p[i] * n;
            for (int j = 0; j < n; ++j) e += (int64_t)D[(size_t)i * n + j] * fr[p[j]];
        }
        return e;
    }

    void setup() {
        Ft.resize((size_t)n * n);
        Dt.resize((size_t)n * n);
        Fd.resize(n);
        Dd.resize(n);
        for (int x = 0; x < n; ++x) {
            Fd[x] = F[(size_t)x * n + x];
            Dd[x] = D[(size_t)x * n + x];
            for (int y = 0; y < n; ++y) {
                Ft[(size_t)x * n + y] = F[(size_t)y * n + x];
                Dt[(size_t)x * n + y] = D[(size_t)y * n + x];
            }
        }
        sym = true;
        for (size_t x = 0; x < (size_t)n * n && sym; ++x) sym = F[x] == Ft[x] && D[x] == Dt[x];
        for (int x = 0; x < n && sym; ++x) sym = Fd[x] == 0 && Dd[x] == 0;
        beta2f.resize(R);
        c20.resize(R);
        for (int l = 0; l < R; ++l) {
            const double b2 = par.beta_ln[l] * kLog2e;
            beta2f[l] = (float)b2;
            c20[l] = 20.0 / b2;
        }
        I.resize((size_t)RT * n * n);
        perms.resize((size_t)3 * RT * n);
        reps.resize(RT);
        lads.resize(T);
        for (int g = 0; g < RT; ++g) {
            Replica &r = reps[g];
            r.perm = perms.data() + (size_t)3 * g * n;
            r.pinv = r.perm + n;
            r.best = r.pinv + n;
            r.pos = g % R;  // CPU_PIM.cpp:115-116
        }
        for (int t = 0; t < T; ++t) {
            lads[t].order.resize(R);
            for (int l = 0; l < R; ++l) lads[t].order[l] = l;
        }
    }

    // Initial states of replicas [g0, g1) of this thread: ladders share their start unless
    // per-replica starts are asked for; a force matrix is built once per distinct start
    // and copied to the other replicas of the ladder.
    void init_range(int g0, int g1) {
        for (int g = g0; g < g1; ++g) {
            Replica &r = reps[g];
            const int t = g / R, l = g % R;
            const bool per_rep = (par.flags & QAP_PT_INIT_PER_REPLICA) != 0;
            if (par.init_perm) std::memcpy(r.perm, par.init_perm, sizeof(int32_t) * n);
            else if (per_rep) init_perm((uint32_t)g, 2u, r.perm);
            else init_perm((uint32_t)t, 1u, r.perm);
            for (int i = 0; i < n; ++i) r.pinv[r.perm[i]] = i;
            std::memcpy(r.best, r.perm, sizeof(int32_t) * n);
            const bool shared = !per_rep || par.init_perm;
            if (shared && g > g0 && (g - 1) / R == t && l > 0) {
                // the previous replica of this ladder (same thread) has the same start
                std::memcpy(Irep(g), Irep(g - 1), sizeof(IT) * (size_t)n * n);
                r.E = reps[g - 1].E;
            } else {
                build_force(r.perm, Irep(g));
                r.E = energy(r.perm);
            }
        }
    }

    // One p-dit update of location i (CPU_PIM.cpp:207-287) with the exact weights.
    void attempt(Replica &r, IT *M, int i, float b2f, double c, uint32_t u, Scratch<IT> &s) {
        int32_t *perm = r.perm, *pinv = r.pinv;
        const int a = perm[i];
        const IT *rowi = M + (size_t)i * n;
        const IT Iia = rowi[a];
        const int32_t *Di = D + (size_t)i * n, *Dti = Dt.data() + (size_t)i * n;
        const int32_t *Fa = F + (size_t)a * n, *Fta = Ft.data() + (size_t)a * n;
        IT *dE = s.dE.data();
        IT maxd = 0;
        if (sym) {
            // zero diagonals, symmetric: dist * flow = 4 D[i][j] F[a][b]
            for (int b = 0; b < n; ++b) {
                const int j = pinv[b];
                const IT *rj = M + (size_t)j * n;
                const IT d = rowi[b] + rj[a] - Iia - rj[b] - (IT)4 * (IT)Di[j] * (IT)Fa[b];
                dE[b] = d;
                maxd = d > maxd ? d : maxd;
            }
        } else {
            const IT ddi = (IT)Dd[i], fda = (IT)Fd[a];
            for (int b = 0; b < n; ++b) {
                const int j = pinv[b];
                const IT *rj = M + (size_t)j * n;
                const IT dist = ddi + (IT)Dd[j] - (IT)Di[j] - (IT)Dti[j];
                const IT flow = fda + (IT)Fd[b] - (IT)Fa[b] - (IT)Fta[b];
                const IT d = rowi[b] + rj[a] - Iia - rj[b] - dist * flow;
                dE[b] = d;
                maxd = d > maxd ? d : maxd;
            }
        }
        uint32_t *w = s.w.data();
        uint32_t S = 0;
        if constexpr (sizeof(IT) == 4) {
            const int32_t cut = qap::heat_cutoff(maxd, c);
            for (int b = 0; b < n; ++b) {
                w[b] = dE[b] > cut ? qap::heat_weight(b2f, dE[b] - maxd) : 0u;
                S += w[b];
            }
        } else {
            const int64_t cut = qap::heat_cutoff64(maxd, c);
            for (int b = 0; b < n; ++b) {
                w[b] = dE[b] > cut ? qap::heat_weight64(b2f, dE[b] - maxd) : 0u;
                S += w[b];
            }
        }
        // roulette: the first candidate whose running sum passes u * S / 2^32
        const uint32_t target = (uint32_t)(((uint64_t)u * S) >> 32);
        uint32_t P = 0;
        int sel = a;
        for (int b = 0; b < n; ++b) {
            P += w[b];
            if (P > target) {
                sel = b;
                break;
            }
        }
        if (sel == a) return;
        commit(r, M, i, sel, dE[sel], s);
    }

    // E -= dE; swap; I[k][c] += u1[k] v1[c] + u2[k] v2[c] with u1 = D[k][j] - D[k][i],
    // v1 = F[c][b] - F[c][a], u2 = D[j][k] - D[i][k], v2 = F[b][c] - F[a][c]
    // (CPU_PIM.cpp:252-283); symmetric instances fold both terms into u2 (2 v2).
    void commit(Replica &r, IT *M, int i, int b, IT dEb, Scratch<IT> &s) {
        const int a = r.perm[i], j = r.pinv[b];
        r.E -= (int64_t)dEb;
        r.perm[i] = b;
        r.pinv[b] = i;
        r.perm[j] = a;
        r.pinv[a] = j;
        ++r.accepted;
        if (r.E < r.bestE) {  // CPU_PIM.cpp:271-272
            r.bestE = r.E;
            std::memcpy(r.best, r.perm, sizeof(int32_t) * n);
        }
        const int32_t *Fa = F + (size_t)a * n, *Fb = F + (size_t)b * n;
        const int32_t *Fta = Ft.data() + (size_t)a * n, *Ftb = Ft.data() + (size_t)b * n;
        const int32_t *Di = D + (size_t)i * n, *Dj = D + (size_t)j * n;
        const int32_t *Dti = Dt.data() + (size_t)i * n, *Dtj = Dt.data() + (size_t)j * n;
        IT *v1 = s.v1.data(), *v2 = s.v2.data(), *u1 = s.u1.data(), *u2 = s.u2.data();
        int32_t *cols = s.cols.data(), *rows = s.rows.data();
        int nc = 0, nr = 0;
        if (sym) {
            for (int c = 0; c < n; ++c) {
                v2[c] = (IT)2 * ((IT)Fb[c] - (IT)Fa[c]);
                if (v2[c] != 0) cols[nc++] = c;
            }
            for (int k = 0; k < n; ++k) {
                u2[k] = (IT)Dj[k] - (IT)Di[k];
                if (u2[k] != 0) rows[nr++] = k;
            }
        } else {
            for (int c = 0; c < n; ++c) {
                v1[c] = (IT)Ftb[c] - (IT)Fta[c];
                v2[c] = (IT)Fb[c] - (IT)Fa[c];
                if ((v1[c] | v2[c]) != 0) cols[nc++] = c;
            }
            for (int k = 0; k < n; ++k) {
                u1[k] = (IT)Dtj[k] - (IT)Dti[k];
                u2[k] = (IT)Dj[k] - (IT)Di[k];
                if ((u1[k] | u2[k]) != 0) rows[nr++] = k;
            }
        }
        r.cells += (uint64_t)n * (uint64_t)nc;
        if (nc == 0 || nr == 0) return;
        const bool compact = 4 * nc < n;
        for (int q = 0; q < nr; ++q) {
            const int k = rows[q];
            IT *row = M + (size_t)k * n;
            const IT x2 = u2[k];
            if (sym) {
                if (compact)
                    for (int m = 0; m < nc; ++m) row[cols[m]] += x2 * v2[cols[m]];
                else
                    for (int c = 0; c < n; ++c) row[c] += x2 * v2[c];
            } else {
                const IT x1 = u1[k];
                if (compact)
                    for (int m = 0; m < nc; ++m) {
                        const int c = cols[m];
                        row[c] += x1 * v1[c] + x2 * v2[c];
                    }
                else
                    for (int c = 0; c < n; ++c) row[c] += x1 * v1[c] + x2 * v2[c];
            }
        }
    }

    // spe sweeps of replica g in exchange interval m (sweeps m*spe ... m*spe + spe - 1)
    void sweeps(int g, int64_t m, Scratch<IT> &s) {
        Replica &r = reps[g];
        IT *M = Irep(g);
        const float b2f = beta2f[r.pos];
        const double c = c20[r.pos];
        for (int q = 0; q < spe; ++q) {
            const uint32_t sweep = (uint32_t)(m * spe + q);
            for (int i = 0; i < n; ++i) {
                const qap::u32x4 x =
                    qap::philox4x32_10(qap::u32x4{(uint32_t)i, sweep, (uint32_t)g, 0u}, k0, k1);
                attempt(r, M, i, b2f, c, x.x, s);
            }
        }
    }

    // Fold, bubble pass and kick of ladder t (CPU_PIM.cpp:295-344); returns the ladder best.
    int64_t exchange(int t, uint64_t &nex, uint64_t &nkick) {
        Ladder &L = lads[t];
        Replica *rp = reps.data() + (size_t)t * R;
        int32_t *ord = L.order.data();
        ++L.tss;
        for (int l = 0; l < R; ++l)
            if (rp[l].bestE < L.best) {
                L.best = rp[l].bestE;
                L.tss = 0;
            }
        for (int k = 0; k + 1 < R; ++k) {
            const int r1 = ord[k], r2 = ord[k + 1];
            if (rp[r1].E < rp[r2].E) {  // the colder position takes the lower cost
                ord[k] = r2;
                ord[k + 1] = r1;
                rp[r1].pos = k + 1;
                rp[r2].pos = k;
                ++nex;
                if (r2 == R - 1) L.tss = 0;  // CPU_PIM.cpp:333-335 tests a replica index
            }
        }
        if (L.tss > par.kick_after) {  // hottest and coldest positions trade (not reset)
            const int r1 = ord[R - 1], r2 = ord[0];
            ord[R - 1] = r2;
            ord[0] = r1;
            rp[r1].pos = 0;
            rp[r2].pos = R - 1;
            ++nkick;
        }
        return L.best;
    }

    int run(qap_pt_result *res, int nthreads) {
        setup();
        const int64_t megas = par.sweeps / spe;
        nthreads = std::max(1, std::min(nthreads, RT));
        // replicas are split into contiguous ranges (a ladder's shared start is built once
        // per range)
        std::vector<int> lo(nthreads + 1);
        for (int k = 0; k <= nthreads; ++k) lo[k] = (int)((int64_t)RT * k / nthreads);
        SpinBarrier bar(nthreads);
        std::atomic<bool> alloc_failed{false};
        // shared control, written by thread 0 before a barrier and read after it
        int64_t m = 0, megas_done = 0;
        bool stop = false;
        uint64_t exchanges = 0, kicks = 0;
        int64_t global_best = INT64_MAX;
        auto t0 = std::chrono::steady_clock::now();  // reset when the sweeps start
        auto elapsed = [&] {
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        };
        auto body = [&](int tid) {
            Scratch<IT> *s = nullptr;
            try {
                s = new Scratch<IT>(n);
            } catch (const std::bad_alloc &) {
                alloc_failed = true;
            }
            if (s) init_range(lo[tid], lo[tid + 1]);
            bar.wait();
            if (tid == 0) t0 = std::chrono::steady_clock::now();  // the sweep phase's clock
            for (;;) {
                if (tid == 0) {
                    // the exchange step that opens interval m (CPU_PIM.cpp:292-357)
                    stop = alloc_failed || m >= megas ||
                           (par.timeout_s > 0 && m > 0 && elapsed() > par.timeout_s);
                    if (!stop) {
                        int64_t gb = global_best;
                        for (int t = 0; t < T; ++t) gb = std::min(gb, exchange(t, exchanges, kicks));
                        global_best = gb;
                        if (par.has_goal && gb <= par.goal) stop = true;  // :346-348
                    }
                }
                bar.wait();
                if (stop) break;
                for (int g = lo[tid]; g < lo[tid + 1]; ++g) sweeps(g, m, *s);
                bar.wait();
                if (tid == 0) megas_done = ++m;
            }
            delete s;
        };
        try {
            std::vector<std::thread> th;
            th.reserve(nthreads - 1);
            for (int k = 1; k < nthreads; ++k) th.emplace_back(body, k);
            body(0);
            for (auto &x : th) x.join();
        } catch (const std::system_error &) {
            return fail(QAP_PT_E_NOMEM, "could not start %d host threads", nthreads);
        }
        if (alloc_failed) return fail(QAP_PT_E_NOMEM, "host allocation failed");
        const double secs = elapsed();

        // best over every replica's best, ties to the lowest replica index; with no
        // accepted move anywhere, the best initial state
        int bi = -1;
        int64_t best = INT64_MAX;
        for (int g = 0; g < RT; ++g)
            if (reps[g].bestE < best) {
                best = reps[g].bestE;
                bi = g;
            }
        if (bi < 0)
            for (int g = 0; g < RT; ++g)
                if (bi < 0 || reps[g].E < best) {
                    best = reps[g].E;
                    bi = g;
                }
        res->best_cost = best;
        if (res->best_perm) std::memcpy(res->best_perm, reps[bi].best, sizeof(int32_t) * n);
        res->sweeps_done = megas_done * spe;
        res->attempts = (uint64_t)res->sweeps_done * (uint64_t)n * (uint64_t)RT;
        res->accepted = 0;
        res->update_cells = 0;
        res->flush_cells = 0;  // updates are applied move by move here
        for (const Replica &r : reps) {
            res->accepted += r.accepted;
            res->update_cells += r.cells;
        }
        res->exchanges = exchanges;
        res->kicks = kicks;
        res->seconds = secs;
        return QAP_PT_OK;
    }
};

int host_threads() {
    if (const char *env = std::getenv("QAP_PT_CPU_THREADS")) {
        const int v = std::atoi(env);
        if (v > 0) return v;
    }
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) return std::max(1, CPU_COUNT(&set));
    return std::max(1u, std::thread::hardware_concurrency());
}

}  // namespace

extern "C" int qap_pt_solve_cpu(const qap_problem *prob, const qap_pt_params *par,
                                qap_pt_result *res) {
    if (!prob || !par || !res) return fail(QAP_PT_E_INVAL, "NULL argument");
    int rc = qap::validate(*prob, *par, 0);
    if (rc) return rc;
    if (par->flags & QAP_PT_ONE_SHOT)
        return fail(QAP_PT_E_INVAL, "one-shot dynamics run on the device only (qap_pt_solve_hip)");
    try {
        if (qap::wide_instance(*prob)) {
            Solver<int64_t> s(*prob, *par);
            return s.run(res, host_threads());
        }
        Solver<int32_t> s(*prob, *par);
        return s.run(res, host_threads());
    } catch (const std::bad_alloc &) {
        return fail(QAP_PT_E_NOMEM, "host allocation failed");
    }
}
