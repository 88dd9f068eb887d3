/*
 * qap_oracle.h -- CPU restatement of the reference's parallel-tempering (PT-SA)
 * solve path for the quadratic assignment problem.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker (or, for the
 * baseline, as the thing timed on the host cores).  The product library never
 * links or calls it.
 *
 * Reference files followed (paths relative to the reference checkout):
 *   CPU-PIM/CPU_PIM.cpp:109-193  setup(): ladder, initial permutations, force build, energy
 *   CPU-PIM/CPU_PIM.cpp:195-290  doMegaItera(): the p-dit (heat-bath) sweep
 *   CPU-PIM/CPU_PIM.cpp:292-364  threadToRun(): best fold, replica exchange, kick, goal
 *   GPU-CUPIM/perl-gpu/src/lib/qap/qap.cu:609-646  calc_energy_delta (same delta)
 *
 * Two selection modes share every other line of the driver:
 *   ORACLE_MODE_EXACT   integer heat-bath weights (2^20-scaled, own exp2), integer
 *                       prefix sums and a 32-bit fixed-point threshold.  This is the
 *                       definition the HIP kernels must reproduce bit for bit.
 *   ORACLE_MODE_REFFP   the reference's own floating-point roulette: long double
 *                       exp2 weights, double sums, ">= r*S" threshold with a 53-bit
 *                       uniform (CPU_PIM.cpp:212,234-251).  Used for the statistical
 *                       comparison with the compiled reference and as the CPU
 *                       baseline that is timed.
 * Random numbers come from Philox4x32-10 in both modes (the reference uses an
 * unseeded mt19937, CPU_PIM.cpp:110,130, so no run of it is reproducible).
 *
 * Pins (see tests/test_oracle.py): Random123 known-answer vectors and rocRAND's
 * philox4x32_10 engine for Philox; closed forms and brute-force recomputation for
 * the force matrix, cost and O(1) delta; exact Gibbs probabilities (chi-square) and
 * the beta->0 / beta->inf limits for the selection; brute-force optimum for n <= 8;
 * hand-built vectors for the exchange/kick quirks; and a best-cost comparison
 * against the compiled reference binary (oracle/_ref) at equal sweeps.
 */
#ifndef QAP_ORACLE_H
#define QAP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_MODE_EXACT = 0, ORACLE_MODE_REFFP = 1 };

enum {
    ORACLE_FLAG_INIT_PER_REPLICA = 1, /* each replica gets its own shuffle (GPU_PIM.cu:151-167) */
    ORACLE_FLAG_ONE_SHOT = 2,         /* one-shot parallel-proposal steps instead of p-dit
                                         sweeps (qap.cu:1037-1609), exact mode only       */
};

typedef struct {
    int32_t n;
    const int32_t *F; /* n*n row-major flow     (CPU_PIM.cpp:67) */
    const int32_t *D; /* n*n row-major distance (CPU_PIM.cpp:66) */
} oracle_problem;

typedef struct {
    int32_t R;               /* replicas per ladder                              */
    int32_t T;               /* independent ladders (CPU_PIM's threads)          */
    const double *beta_ln;   /* R values, ascending: position 0 = hottest        */
    int64_t sweeps;          /* sweeps per replica, multiple of spe             */
    int32_t spe;             /* sweeps between exchanges (NUMITERA)             */
    uint64_t seed;
    int32_t kick_after;      /* 100 in the reference                            */
    int32_t has_goal;
    int64_t goal;
    int32_t flags;
    const int32_t *init_perm; /* optional n: same start for every replica        */
} oracle_params;

typedef struct {
    int64_t best_cost;
    int32_t *best_perm;      /* caller-allocated n                               */
    uint64_t attempts, accepted, exchanges, kicks;
    int64_t megas_done;
    /* optional outputs, each may be NULL (sizes use RT = R*T) */
    int32_t *final_perm;     /* RT*n                                             */
    int64_t *final_E;        /* RT                                               */
    int64_t *final_bestE;    /* RT (INT64_MAX = no accepted move yet)           */
    int32_t *final_pos;      /* RT: ladder position of each replica              */
    uint64_t *final_accepted;/* RT                                               */
    int32_t *best_perms;     /* RT*n: permutation at each replica's best         */
    int64_t *traj_E;         /* traj_cap*RT: energies after each mega-iteration  */
    int64_t traj_cap;
} oracle_result;

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): out = bijection(ctr, key). */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Integer heat-bath weight floor(2^20 * 2^(beta2f*d)) for d <= 0, own exp2. */
uint32_t oracle_weight(float beta2f, int32_t d);
/* the same weight for a 64-bit difference d (instances beyond the int32 bounds) */
uint32_t oracle_weight64(float beta2f, int64_t d);

/* I[i][a] = -sum_j (D[i][j] F[a][p(j)] + D[j][i] F[p(j)][a])   (CPU_PIM.cpp:164-181) */
void oracle_force_matrix(int32_t n, const int32_t *F, const int32_t *D, const int32_t *perm,
                         int32_t *I);
/* E = sum_ij D[i][j] F[p(i)][p(j)]   (CPU_PIM.cpp:184-192) */
int64_t oracle_energy(int32_t n, const int32_t *F, const int32_t *D, const int32_t *perm);
/* O(1) delta E_old - E_new for moving facility b to location i (CPU_PIM.cpp:225) */
int64_t oracle_delta(int32_t n, const int32_t *F, const int32_t *D, const int32_t *perm,
                     const int32_t *pinv, const int32_t *I, int32_t i, int32_t b);
/* rank-2 force update after swapping the facilities at locations i and j
 * (a = p(i), b = p(j) before the swap)   (CPU_PIM.cpp:263-269) */
void oracle_force_update(int32_t n, const int32_t *F, const int32_t *D, int32_t *I, int32_t i,
                         int32_t j, int32_t a, int32_t b);

/* One exact-mode p-dit update of location i; returns the chosen facility b (b == p(i)
 * means stay) and fills dE (n) / w (n).  State arrays are updated on a commit. */
int32_t oracle_attempt_exact(int32_t n, const int32_t *F, const int32_t *D, int32_t *perm,
                             int32_t *pinv, int32_t *I, int64_t *E, int32_t i, float beta2f,
                             double c20, uint32_t u, int32_t *dE, uint32_t *w);

/* One one-shot step of one replica (GPU-CUPIM/perl-gpu/src/lib/qap/qap.cu:1037-1460):
 * every location pi proposes, from the same snapshot, a swap with the facility at one
 * location pj, drawn by the exact heat-bath rule over pj = 0..n-1 in location order
 * (pj == pi is the stay; uniform from ctr = {pi, step, stream, 3}).  A proposal claims
 * both of its locations with the key (~(delta ^ 2^31) << 32) | pi -- largest delta
 * first, then smallest pi (qap.cu:1170-1199) -- and is applied when its key holds both
 * (qap.cu:1203-1221).  Applied swaps are disjoint, so the force matrix moves by the sum
 * of their rank-2 updates taken from the snapshot (qap.cu:1301-1376); the cost is then
 * recomputed exactly as -1/2 sum_i I[i][p(i)] (qap.cu:1418-1457).  Scratch: work 4n,
 * owner n, w n.  Returns the number of applied swaps; tgt (n, may be NULL) receives
 * every location's proposal and acc (n, may be NULL) whether it was applied. */
int32_t oracle_one_shot_step(int32_t n, const int32_t *F, const int32_t *D, int32_t *perm,
                             int32_t *pinv, int32_t *I, int64_t *E, float beta2f, double c20,
                             uint64_t seed, uint32_t step, uint32_t stream, int32_t *work,
                             uint64_t *owner, uint32_t *w, int32_t *tgt, int32_t *acc);

/* One exchange step of one ladder (CPU_PIM.cpp:295-344): best fold, bubble pass from
 * the hottest position, hot/cold kick.  E/bestE/pos are indexed by replica, order by
 * ladder position.  Returns the number of exchanges; *kicked is set to 0/1. */
int32_t oracle_exchange_step(int32_t R, const int64_t *E, const int64_t *bestE, int32_t *order,
                             int32_t *pos, int32_t *tss, int64_t *ladder_best,
                             int32_t kick_after, int32_t *kicked);

/* beta (natural-log units) -> exp2 units with the reference's constant
 * (CPU_PIM.cpp:528: "* 1.44269504089"). */
double oracle_beta2(double beta_ln);

/* Initial permutation of ladder t (or replica t*R+l with INIT_PER_REPLICA):
 * identity followed by n transpositions (i, j) with j uniform in [0,n)
 * (CPU_PIM.cpp:143-161). */
void oracle_init_perm(int32_t n, uint64_t seed, uint32_t stream, int32_t *perm);

/* Full solve.  nthreads > 1 runs ladders on POSIX threads (only without a goal,
 * where ladders never interact).  Returns 0 or a negative error (-4: the entries are
 * too large even for 64-bit force matrices, or too large for one-shot steps).
 * Force matrices are int32 when the permutation-independent bounds allow it
 * (4 B + 16 maxD maxF < 2^30, B the force bound), else int64 with int64 deltas
 * (the wide path: 4 B + 16 maxD maxF < 2^62 and n^2 maxD maxF < 2^62). */
int oracle_solve(const oracle_problem *prob, const oracle_params *par, oracle_result *res,
                 int32_t mode, int32_t nthreads);

/* Largest |I| any permutation can produce and the largest |delta E|; used to prove
 * that int32 force matrices and deltas cannot overflow. */
void oracle_bounds(int32_t n, const int32_t *F, const int32_t *D, int64_t *force_bound,
                   int64_t *delta_bound);

#ifdef __cplusplus
}
#endif
#endif
