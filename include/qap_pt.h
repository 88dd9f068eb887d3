/*
 * qap_pt.h -- C-ABI of the MI355X parallel-tempering QAP annealer.
 *
 * Drop-in for the reference's PT-SA solve path:
 *   CPU-PIM/CPU_PIM.cpp:478-685   main(): readProblem -> setup -> threadToRun x T
 *   GPU-PIM/GPU_PIM.cu:742-1057   main(): setup -> persistent_sim_kernel launch
 *   GPU-CUPIM/perl-gpu/src/lib/qap/qap.cu:28-303  qap::ReplicaPool<F,Z> (staged pool API)
 * The reference exposes no library API (state lives in globals, inputs come from a
 * file and flags, results go to stdout/log.txt).  These entry points carry exactly
 * what its flags and globals carry:
 *   -b beta0 / -m ratio  (CPU_PIM.cpp:526-561)  -> qap_pt_params.beta_ln (explicit ladder)
 *   -r replicas          (CPU_PIM.cpp:569-580)  -> qap_pt_params.R
 *   -t threads (ladders) (CPU_PIM.cpp:588-599)  -> qap_pt_params.T
 *   -w megas, -i NUMITERA(CPU_PIM.cpp:548-606)  -> sweeps = megas * NUMITERA, sweeps_per_exchange
 *   -g goal, -s timeout  (CPU_PIM.cpp:534-587)  -> has_goal/goal, timeout_s
 *   "Single best energy found" (CPU_PIM.cpp:678) -> qap_pt_result.best_cost (+ best_perm,
 *                                                 which the reference only prints with -G)
 *
 * Conventions: every call returns QAP_PT_OK (0) or a negative status and leaves a
 * message for qap_pt_last_error() (thread-local).  The caller owns every buffer;
 * inputs are only read during the call.  Solve calls are reentrant; one pool or
 * server handle must not be used from two threads at once.  No C++ exception
 * crosses this boundary.  All costs are exact integers (int64).  Force matrices are
 * int32 when a permutation-independent bound proves the deltas fit (4 B + 16 maxD maxF <
 * 2^30, B the force bound), else int64 in HBM (GPU-CUPIM's ReplicaPool<F, int64_t>,
 * qap.cu:369-411, --large-energies at qap_gpu.cu:789-793) when 4 B + 16 maxD maxF < 2^62
 * and n^2 maxD maxF < 2^62; beyond that, QAP_PT_E_OVERFLOW before any launch.
 */
#ifndef QAP_PT_H
#define QAP_PT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QAP_PT_OK 0
#define QAP_PT_E_INVAL (-1)
#define QAP_PT_E_NOMEM (-2)
#define QAP_PT_E_DEVICE (-3)
#define QAP_PT_E_OVERFLOW (-4)
#define QAP_PT_E_BUSY (-5)

/* Largest problem size the sweep kernels accept. */
#define QAP_PT_MAX_N 1024

/* qap_pt_params.flags */
#define QAP_PT_INIT_PER_REPLICA 1u /* own initial shuffle per replica (GPU_PIM.cu:151-167)
                                      instead of one per ladder (CPU_PIM.cpp:143-161)     */
#define QAP_PT_ONE_SHOT 2u         /* CUPIM's one-shot parallel-proposal dynamics
                                      (GPU-CUPIM/perl-gpu/src/lib/qap/qap.cu:1037-1609,
                                      ReplicaPool::perform_step_one_shot) instead of p-dit
                                      sweeps: each "sweep" is one step in which every
                                      location proposes a swap from the same snapshot,
                                      conflicts resolve by the largest delta, and the
                                      disjoint winners are applied together.  A different
                                      Markov chain; attempts count n proposals per step.
                                      Needs n <= 256, 2 max|D| <= 32639, 4 max|F| <= 32639. */

/* One instance: n, then flow and distance, both n*n row-major int32, exactly as
 * readProblem stores them (CPU_PIM.cpp:79-107). */
typedef struct {
    int32_t n;
    const int32_t *F;
    const int32_t *D;
} qap_problem;

typedef struct {
    int32_t R;                   /* replicas per ladder (>= 1)                            */
    int32_t T;                   /* independent ladders (>= 1)                            */
    const double *beta_ln;       /* R inverse temperatures in natural-log units, > 0 and
                                    non-decreasing (index 0 = hottest); converted
                                    to exp2 units with the reference's 1.44269504089      */
    int64_t sweeps;              /* sweeps per replica, a multiple of sweeps_per_exchange */
    int32_t sweeps_per_exchange; /* NUMITERA (CPU_PIM.cpp:19)                             */
    uint64_t seed;               /* Philox4x32-10 key                                     */
    int32_t kick_after;          /* hot/cold kick threshold, 100 in the reference         */
    int32_t has_goal;            /* stop at the first exchange step with best <= goal     */
    int64_t goal;
    double timeout_s;            /* 0 = none; checked between device batches of sweeps    */
    uint32_t flags;
    const int32_t *init_perm;    /* optional n: start every replica here (location->facility) */
} qap_pt_params;

typedef struct {
    int64_t best_cost;           /* min over every replica's best cost                    */
    int32_t *best_perm;          /* caller-allocated n (may be NULL)                      */
    uint64_t attempts;           /* p-dit updates performed = sweeps_done * n * R * T      */
    uint64_t accepted;           /* updates that moved a facility                          */
    uint64_t exchanges;          /* ladder swaps by the exchange pass                     */
    uint64_t kicks;              /* hot/cold kicks                                         */
    uint64_t update_cells;       /* force-matrix entries rewritten by rank-2 updates
                                    (columns with a zero flow difference are skipped)      */
    uint64_t flush_cells;        /* force-matrix entries written by flushes that apply up to
                                    64 pending rank-2 updates in one pass (placement of
                                    symmetric instances whose matrix stays in HBM); 0 where
                                    every move updates the matrix itself                   */
    int64_t sweeps_done;         /* per replica                                            */
    double seconds;              /* wall time of the sweep phase                          */
} qap_pt_result;

/* Message of the last failing call on this thread ("" if none). */
const char *qap_pt_last_error(void);

/* Number of visible HIP devices (0 when none). */
int qap_pt_device_count(void);

/* --- one-shot solve: replaces CPU_PIM/GPU_PIM main()'s solve path ------------- */
int qap_pt_solve_hip(const qap_problem *prob, const qap_pt_params *par, qap_pt_result *res);

/* --- the same solve on the host's cores: replaces CPU-PIM's own solve path
 *     (CPU-PIM/CPU_PIM.cpp:478-685, setup -> threadToRun x T).  Same arguments, errors
 *     and results as qap_pt_solve_hip (the same chain, update for update, bit-identical
 *     for any thread count); makes no device call, so it runs on hosts without a GPU.
 *     A ladder's replicas are spread over the threads (QAP_PT_CPU_THREADS, default every
 *     core the process may run on, at most R*T).  QAP_PT_ONE_SHOT is a device-only
 *     dynamics: QAP_PT_E_INVAL. ----------------------------------------------------- */
int qap_pt_solve_cpu(const qap_problem *prob, const qap_pt_params *par, qap_pt_result *res);

/* --- allocation-free solve.  No counterpart in the reference: it is for callers that
 *     cannot allocate device memory (tight serving loops, harnesses that forbid hipMalloc).
 *     The same solve as qap_pt_solve_hip -- same arguments, results (bit-identical) and
 *     errors -- but every device buffer of the solve is a piece of one fixed region of
 *     QAP_PT_NOALLOC_SCRATCH_BYTES per GPU, which the library holds from its first call on
 *     that GPU until the process ends.  It calls no hipMalloc/hipFree, creates no stream,
 *     event or graph: the kernels are launched one by one on `stream` (a hipStream_t;
 *     NULL = the null stream) and the call returns once the results are copied out.  A
 *     request whose device state exceeds the region fails with QAP_PT_E_NOMEM before any
 *     device work (the message gives the bytes it needs).  Calls on one GPU share its
 *     region, so two at once would race: the library serialises them with a per-device
 *     lock.  The region keeps the last call's contents; each call writes everything it
 *     reads first. ------------------------------------------------------------------- */
#define QAP_PT_NOALLOC_SCRATCH_BYTES (64ll << 20)
int qap_pt_solve_hip_noalloc(const qap_problem *prob, const qap_pt_params *par,
                             qap_pt_result *res, void *stream);

/* --- batched solve: many independent instances (uneven n, own ladders) in the same
 *     device launches; results[k] belongs to problems[k]. ------------------------- */
int qap_pt_solve_batch_hip(int32_t count, const qap_problem *problems,
                           const qap_pt_params *params, qap_pt_result *results);

/* --- staged pool (qap::ReplicaPool, qap.cu:28-303: from_file/cu_copy/
 *     recompute_force_mats/recompute_energies -> create; perform_step +
 *     percolate_betas -> run; thunk_scores -> result; cu_free -> destroy) ------- */
typedef struct qap_pt_pool qap_pt_pool;

/* Uploads count instances, builds every initial permutation, force matrix and cost
 * on the device.  params[k].sweeps is the default length for qap_pt_pool_run. */
int qap_pt_pool_create(int32_t count, const qap_problem *problems, const qap_pt_params *params,
                       qap_pt_pool **out);
/* Advances every unfinished instance by `sweeps` sweeps (a multiple of each
 * instance's sweeps_per_exchange; <0 = each instance's params.sweeps) and waits.
 * With kernel_ms != NULL the sweep kernels run without a graph and the summed
 * device time of every sweep-kernel launch (HIP events on the pool's stream) and
 * the number of such launches are returned. */
int qap_pt_pool_run(qap_pt_pool *pool, int64_t sweeps, double *kernel_ms, int64_t *launches);
int qap_pt_pool_result(qap_pt_pool *pool, int32_t inst, qap_pt_result *res);
/* best_costs[count]: every instance's best cost so far, as qap_pt_pool_result reports
 * it, with one device copy for the whole pool (CUPIM's thunk_scores, qap.cu:28-303). */
int qap_pt_pool_best_costs(qap_pt_pool *pool, int64_t *best_costs);
/* Copies instance inst's replica state (each pointer may be NULL). Sizes use
 * RT = R*T: perm RT*n, E RT, bestE RT (INT64_MAX = no accepted move yet), pos RT
 * (ladder position, 0 = hottest), accepted RT, best_perms RT*n. */
int qap_pt_pool_state(qap_pt_pool *pool, int32_t inst, int32_t *perm, int64_t *E, int64_t *bestE,
                      int32_t *pos, uint64_t *accepted, int32_t *best_perms);
/* Rebuilds every force matrix and cost from the permutations on the device and
 * counts disagreeing entries with the incrementally tracked ones (0 = exact). */
int qap_pt_pool_verify(qap_pt_pool *pool, int64_t *force_mismatches, int64_t *cost_mismatches);
int qap_pt_pool_destroy(qap_pt_pool *pool);

/* --- serving loop: continuous batching (no counterpart in the reference, which solves
 *     one instance per process; this is the request loop around the same solve path).
 *     Requests are copied at submit.  A worker thread keeps slot banks resident on the
 *     device -- pools of fixed-capacity instance slots per size class (n <= 64, 128, 256,
 *     512, 1024; up to 64 replicas), created when the device budget holds them -- and runs
 *     every request loaded in a bank in the same kernel launches; between launches it
 *     releases finished requests and loads queued ones into free slots, so requests that
 *     arrive while others run join the next launch.  Requests no bank class holds run as
 *     pools of their own.  A request's result does not depend on what it ran with. ----- */
typedef struct qap_srv qap_srv;

typedef struct {
    int32_t max_pools;       /* slot banks + separate pools at once (>= 1), default 4     */
    int32_t max_batch;       /* slots per bank (>= 1), default 64                          */
    int64_t device_budget;   /* device bytes the server may hold; 0 = free HBM - reserve  */
    int64_t reserve_bytes;   /* kept free when device_budget == 0, default 1 GiB          */
    double batch_window_s;   /* hold a partial batch up to this long for more requests    */
} qap_srv_config;

/* request status reported by qap_srv_poll / qap_srv_wait (negative = failed with that
 * QAP_PT_E_* code, e.g. QAP_PT_E_NOMEM when it can never fit the budget) */
#define QAP_SRV_QUEUED 0
#define QAP_SRV_RUNNING 1
#define QAP_SRV_DONE 2

int qap_srv_create(const qap_srv_config *cfg, qap_srv **out); /* cfg NULL = defaults */
/* Queues one request; QAP_PT_E_NOMEM at once if it alone exceeds an explicit budget. */
int qap_srv_submit(qap_srv *srv, const qap_problem *prob, const qap_pt_params *par,
                   uint64_t *ticket);
/* Non-blocking.  On a final status (DONE or failed) res is filled -- res->best_perm is
 * caller-allocated n or NULL; res->seconds is the submit-to-completion latency -- and
 * the ticket is released. */
int qap_srv_poll(qap_srv *srv, uint64_t ticket, qap_pt_result *res, int32_t *status);
/* As poll, waiting up to timeout_s (<= 0: no limit) for a final status. */
int qap_srv_wait(qap_srv *srv, uint64_t ticket, double timeout_s, qap_pt_result *res,
                 int32_t *status);
int qap_srv_stats(qap_srv *srv, int64_t *queued, int64_t *running, int64_t *done,
                  int64_t *device_bytes);
/* Finishes running pools, fails queued requests with QAP_PT_E_BUSY, frees everything. */
int qap_srv_destroy(qap_srv *srv);

/* --- automatic ladder (CPU_PIM.cpp:366-476: autoSelectParameters) -----------------
 * Exponents b = QAP_PT_AUTO_MIN_EXP .. QAP_PT_AUTO_MIN_EXP + QAP_PT_AUTO_STEPS - 1; at
 * each, QAP_PT_AUTO_TRIALS runs of base->T ladders x base->R replicas, all at exp2-unit
 * beta 2^b, for QAP_PT_AUTO_MEGAS exchange intervals of base->sweeps_per_exchange
 * sweeps.  avg_energy[i] (may be NULL, else QAP_PT_AUTO_STEPS doubles) receives the mean
 * final cost at exponent i; beta0_ln and ratio the ladder for qap_pt_params as the
 * reference's -b / -m would give it (largest beta, multiplicative step).  base->beta_ln,
 * sweeps, has_goal, goal, timeout_s and init_perm are ignored; base->flags chooses the
 * trials' dynamics (QAP_PT_ONE_SHOT: one-shot steps, as GPU-CUPIM's automatic mode runs
 * its own dynamics, qap_gpu.cu:502-720). */
#define QAP_PT_AUTO_MIN_EXP (-30)
#define QAP_PT_AUTO_STEPS 32
#define QAP_PT_AUTO_TRIALS 5
#define QAP_PT_AUTO_MEGAS 10
int qap_pt_auto_beta(const qap_problem *prob, const qap_pt_params *base, double lower_prop,
                     double higher_prop, double *avg_energy, double *beta0_ln, double *ratio);
/* The selection rule alone (CPU_PIM.cpp:431-473) over avg_energy[steps] at exponents
 * QAP_PT_AUTO_MIN_EXP + i; lower_prop / higher_prop are the reference's -al / -ah
 * (defaults 0.1 / 0.9).  Host only, no device needed. */
int qap_pt_auto_select(const double *avg_energy, int32_t steps, double lower_prop,
                       double higher_prop, int32_t R, double *beta0_ln, double *ratio);

/* Sweep-kernel shape chosen for a symmetric instance of size n with a large force bound
 * in a pool of `replicas` replicas in all ("threads x candidates per thread,
 * lds32|lds16|lazy|hbm"), for reports.  Pools with more replicas than the device has CUs
 * take the placement with the best tuned full-chip rate; smaller ones the lowest-latency
 * placement that fits. */
int qap_pt_kernel_shape(int32_t n, int64_t replicas, char *buf, int32_t buflen);
/* Shape actually used for instance inst of a pool ("NTxCPL,placement,sym|general,
 * dense-flow|sparse-flow"). */
int qap_pt_pool_shape(qap_pt_pool *pool, int32_t inst, char *buf, int32_t buflen);
/* Kernel bucket b of a pool (the replicas of one sweep-kernel instantiation, launched
 * together every exchange interval; b = 0, 1, ... until QAP_PT_E_INVAL): its shape
 * ("NTxCPL,placement,sym|general[,sparse-flow]", sparse-flow when the kernel is built
 * without the dense update path), its replicas, and the summed device time and number of
 * its sweep launches over the pool's timed runs (qap_pt_pool_run with kernel_ms: HIP
 * events on the stream each launch goes to, buckets concurrent as in the graphs).  A
 * report with no counterpart in the reference; each pointer may be NULL. */
int qap_pt_pool_bucket(qap_pt_pool *pool, int32_t b, char *shape, int32_t buflen,
                       int64_t *replicas, double *kernel_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif
