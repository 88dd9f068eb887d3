/*
 * qap_oracle.c -- CPU restatement of the reference PT-SA solve path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline); the product
 * library never links it.  The header lists the reference lines each part follows
 * and the pins that check it.  Plain C99 + POSIX threads, runtime sizes.
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off, so that float arithmetic is
 * exactly the IEEE sequence written here).
 */
#include "qap_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Philox4x32-10, Salmon et al., "Parallel random numbers: as easy as 1, 2, 3"  */
/* (SC'11).  Multipliers and Weyl key increments are the published constants.  */
/* ------------------------------------------------------------------------- */
#define PH_M0 0xD2511F53u
#define PH_M1 0xCD9E8D57u
#define PH_W0 0x9E3779B9u
#define PH_W1 0xBB67AE85u

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += PH_W0;
            k1 += PH_W1;
        }
        uint64_t p0 = (uint64_t)PH_M0 * c0;
        uint64_t p1 = (uint64_t)PH_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Counter layout shared with the device code (stated, not shared source):
 *   sweep attempt : ctr = {i, sweep, replica_stream, 0}
 *   ladder init   : ctr = {i, ladder, 0, 1}
 *   replica init  : ctr = {i, replica_stream, 0, 2}
 *   key           = {seed & 0xffffffff, seed >> 32}                            */
static void ph_draw(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                    uint32_t out[4]) {
    const uint32_t ctr[4] = {c0, c1, c2, c3};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    oracle_philox4x32_10(ctr, key, out);
}

/* ------------------------------------------------------------------------- */
/* Exact heat-bath weight.  w = floor(2^20 * 2^y), y = beta2f * d <= 0.        */
/* 2^y = 2^k * 2^f with k = rint(y), f = y - k in [-1/2, 1/2]; 2^f = e^(f ln 2) */
/* by its Taylor series truncated after degree 7 (Horner, fused multiply-add); */
/* coefficients (ln 2)^m / m! rounded to float.  Truncation error < 2e-8,      */
/* far below the 2^-20 weight resolution.                                      */
/* ------------------------------------------------------------------------- */
static const float EXP2_C[8] = {
    0x1p+0f,         0x1.62e430p-1f,  0x1.ebfbe0p-3f,  0x1.c6b08ep-5f,
    0x1.3b2ab6p-7f,  0x1.5d87fep-10f, 0x1.430912p-13f, 0x1.ffcbfcp-17f,
};

/* floor(2^20 * 2^y) for y = beta2f * d: the float steps after the product */
static uint32_t oracle_weight_y(float y) {
    if (!(y > -32.0f)) return 0u;
    float kf = rintf(y);
    float f = y - kf;
    float p = EXP2_C[7];
    for (int m = 6; m >= 0; m--) p = fmaf(p, f, EXP2_C[m]);
    int e = (int)kf + 20;
    if (e < 0) return 0u;
    float w = p * (float)(1u << e);
    return (uint32_t)w;
}

uint32_t oracle_weight(float beta2f, int32_t d) { return oracle_weight_y(beta2f * (float)d); }

double oracle_beta2(double beta_ln) { return beta_ln * 1.44269504089; }

/* ------------------------------------------------------------------------- */
/* Force matrix, energy, delta, rank-2 update                                 */
/* ------------------------------------------------------------------------- */
void oracle_force_matrix(int32_t n, const int32_t *F, const int32_t *D, const int32_t *perm,
                         int32_t *I) {
    /* I[i][a] = -sum_j (D[i][j] F[a][p(j)] + D[j][i] F[p(j)][a]) (CPU_PIM.cpp:164-181),
     * accumulated in int64 one row i at a time: for each j, the flow column p(j) read as
     * row p(j) of F^T and the flow row p(j), so every inner loop is unit-stride. */
    int32_t *Ft = malloc(sizeof(int32_t) * (size_t)n * n);
    int64_t *acc = malloc(sizeof(int64_t) * (size_t)n);
    for (int x = 0; x < n; x++)
        for (int y = 0; y < n; y++) Ft[(size_t)x * n + y] = F[(size_t)y * n + x];
    for (int i = 0; i < n; i++) {
        for (int a = 0; a < n; a++) acc[a] = 0;
        for (int j = 0; j < n; j++) {
            const int64_t dij = D[(size_t)i * n + j], dji = D[(size_t)j * n + i];
            const int32_t *fcol = Ft + (size_t)perm[j] * n; /* F[a][p(j)] over a */
            const int32_t *frow = F + (size_t)perm[j] * n;  /* F[p(j)][a] over a */
            for (int a = 0; a < n; a++) acc[a] -= dij * fcol[a] + dji * frow[a];
        }
        for (int a = 0; a < n; a++) I[(size_t)i * n + a] = (int32_t)acc[a];
    }
    free(acc);
    free(Ft);
}

int64_t oracle_energy(int32_t n, const int32_t *F, const int32_t *D, const int32_t *perm) {
    int64_t e = 0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) e += (int64_t)D[i * n + j] * F[perm[i] * n + perm[j]];
    return e;
}

int64_t oracle_delta(int32_t n, const int32_t *F, const int32_t *D, const int32_t *perm,
                     const int32_t *pinv, const int32_t *I, int32_t i, int32_t b) {
    int32_t a = perm[i];
    int32_t j = pinv[b];
    int64_t force = (int64_t)I[i * n + b] + I[j * n + a] - I[i * n + a] - I[j * n + b];
    int64_t dist = (int64_t)D[i * n + i] + D[j * n + j] - D[i * n + j] - D[j * n + i];
    int64_t flow = (int64_t)F[a * n + a] + F[b * n + b] - F[a * n + b] - F[b * n + a];
    return force - dist * flow;
}

void oracle_force_update(int32_t n, const int32_t *F, const int32_t *D, int32_t *I, int32_t i,
                         int32_t j, int32_t a, int32_t b) {
    /* I[k][c] += (D[k][j]-D[k][i]) (F[c][b]-F[c][a]) + (D[j][k]-D[i][k]) (F[b][c]-F[a][c])
     * (CPU_PIM.cpp:263-269); the flow differences depend on c only and are formed once. */
    int64_t *v = malloc(sizeof(int64_t) * 2 * (size_t)n);
    for (int c = 0; c < n; c++) {
        v[c] = (int64_t)F[(size_t)c * n + b] - F[(size_t)c * n + a];
        v[n + c] = (int64_t)F[(size_t)b * n + c] - F[(size_t)a * n + c];
    }
    for (int k = 0; k < n; k++) {
        const int64_t u1 = (int64_t)D[(size_t)k * n + j] - D[(size_t)k * n + i];
        const int64_t u2 = (int64_t)D[(size_t)j * n + k] - D[(size_t)i * n + k];
        int32_t *row = I + (size_t)k * n;
        for (int c = 0; c < n; c++) row[c] = (int32_t)(row[c] + u1 * v[c] + u2 * v[n + c]);
    }
    free(v);
}

void oracle_bounds(int32_t n, const int32_t *F, const int32_t *D, int64_t *force_bound,
                   int64_t *delta_bound) {
    int64_t maxD = 0, maxF = 0, sD = 0, sF = 0;
    for (int x = 0; x < n * n; x++) {
        int64_t d = llabs((long long)D[x]), f = llabs((long long)F[x]);
        if (d > maxD) maxD = d;
        if (f > maxF) maxF = f;
    }
    for (int i = 0; i < n; i++) {
        int64_t rd = 0, rf = 0;
        for (int j = 0; j < n; j++) {
            rd += llabs((long long)D[i * n + j]) + llabs((long long)D[j * n + i]);
            rf += llabs((long long)F[i * n + j]) + llabs((long long)F[j * n + i]);
        }
        if (rd > sD) sD = rd;
        if (rf > sF) sF = rf;
    }
    /* products in 128 bits: with entries near 2^31 the bounds pass 2^63; they saturate
     * at INT64_MAX (such an instance is refused) */
    __int128 b1 = (__int128)maxF * sD, b2 = (__int128)maxD * sF;
    __int128 B = b1 < b2 ? b1 : b2;
    __int128 db = 4 * B + 16 * (__int128)maxD * maxF;
    *force_bound = B > INT64_MAX ? INT64_MAX : (int64_t)B;
    *delta_bound = db > INT64_MAX ? INT64_MAX : (int64_t)db;
}

/* 0: the force matrix, deltas and update terms provably fit int32 (the int32 path);
 * 1: they fit int64 only (the wide path: 4 B + 16 maxD maxF < 2^62 and the cost
 *    n^2 maxD maxF < 2^62); -1: not even that. */
static int32_t width_class(int32_t n, const int32_t *F, const int32_t *D) {
    int64_t fb, db, maxD = 0, maxF = 0;
    oracle_bounds(n, F, D, &fb, &db);
    for (int64_t x = 0; x < (int64_t)n * n; x++) {
        int64_t d = llabs((long long)D[x]), f = llabs((long long)F[x]);
        if (d > maxD) maxD = d;
        if (f > maxF) maxF = f;
    }
    if (db < ((int64_t)1 << 30) && fb + 8 * maxD * maxF < ((int64_t)1 << 31)) return 0;
    if (db < ((int64_t)1 << 62) && (__int128)n * n * maxD * maxF < ((__int128)1 << 62)) return 1;
    return -1;
}

/* ------------------------------------------------------------------------- */
/* One p-dit update (CPU_PIM.cpp:207-287).                                    */
/* ------------------------------------------------------------------------- */
static int32_t cutoff_of(int32_t maxd, double c20) {
    /* CPU_PIM.cpp:234-235: (int)(maxDeltaE - 20.0 / beta), saturated below. */
    double cd = (double)maxd - c20;
    if (cd < -2147483648.0) return INT32_MIN;
    return (int32_t)cd;
}

static void commit(int32_t n, const int32_t *F, const int32_t *D, int32_t *perm, int32_t *pinv,
                   int32_t *I, int64_t *E, int32_t i, int32_t b, int32_t dEb) {
    int32_t a = perm[i];
    int32_t j = pinv[b];
    *E -= dEb;                       /* CPU_PIM.cpp:256 */
    perm[i] = b;                     /* :257-261 */
    pinv[b] = i;
    perm[j] = a;
    pinv[a] = j;
    oracle_force_update(n, F, D, I, i, j, a, b); /* :263-269 */
}

int32_t oracle_attempt_exact(int32_t n, const int32_t *F, const int32_t *D, int32_t *perm,
                             int32_t *pinv, int32_t *I, int64_t *E, int32_t i, float beta2f,
                             double c20, uint32_t u, int32_t *dE, uint32_t *w) {
    int32_t a = perm[i];
    int32_t maxd = 0;
    for (int b = 0; b < n; b++) {
        int64_t d = oracle_delta(n, F, D, perm, pinv, I, i, b);
        dE[b] = (int32_t)d;
        if (dE[b] > maxd) maxd = dE[b];
    }
    int32_t cut = cutoff_of(maxd, c20);
    uint32_t S = 0;
    for (int b = 0; b < n; b++) {
        w[b] = dE[b] > cut ? oracle_weight(beta2f, dE[b] - maxd) : 0u;
        S += w[b];
    }
    uint32_t T = (uint32_t)(((uint64_t)u * S) >> 32);
    uint32_t P = 0;
    int32_t sel = a;
    for (int b = 0; b < n; b++) {
        P += w[b];
        if (P > T) {
            sel = b;
            break;
        }
    }
    if (sel != a) commit(n, F, D, perm, pinv, I, E, i, sel, dE[sel]);
    return sel;
}

/* ------------------------------------------------------------------------- */
/* One-shot parallel-proposal step (qap.cu:1037-1460), exact weights.         */
/* ------------------------------------------------------------------------- */
int32_t oracle_one_shot_step(int32_t n, const int32_t *F, const int32_t *D, int32_t *perm,
                             int32_t *pinv, int32_t *I, int64_t *E, float beta2f, double c20,
                             uint64_t seed, uint32_t step, uint32_t stream, int32_t *work,
                             uint64_t *owner, uint32_t *w, int32_t *tgt, int32_t *acc) {
    int32_t *prop = work, *del = work + n, *dE = work + 2 * n, *snap = work + 3 * n;
    /* proposals, every one from the snapshot (qap.cu:1037-1140) */
    for (int pi = 0; pi < n; pi++) {
        int32_t maxd = 0;
        for (int pj = 0; pj < n; pj++) {
            dE[pj] = (int32_t)oracle_delta(n, F, D, perm, pinv, I, pi, perm[pj]);
            if (dE[pj] > maxd) maxd = dE[pj];
        }
        int32_t cut = cutoff_of(maxd, c20);
        uint32_t S = 0;
        for (int pj = 0; pj < n; pj++) {
            w[pj] = dE[pj] > cut ? oracle_weight(beta2f, dE[pj] - maxd) : 0u;
            S += w[pj];
        }
        uint32_t x[4];
        ph_draw(seed, (uint32_t)pi, step, stream, 3u, x);
        uint32_t T = (uint32_t)(((uint64_t)x[0] * S) >> 32);
        uint32_t P = 0;
        int32_t sel = pi; /* every weight truncated to zero: stay */
        for (int pj = 0; pj < n; pj++) {
            P += w[pj];
            if (P > T) {
                sel = pj;
                break;
            }
        }
        prop[pi] = sel;
        del[pi] = sel == pi ? 0 : dE[sel];
    }
    /* claims on both locations (qap.cu:1157-1199) */
    for (int x = 0; x < n; x++) owner[x] = UINT64_MAX;
    for (int pi = 0; pi < n; pi++) {
        int32_t pj = prop[pi];
        if (pj == pi) continue;
        uint64_t key = ((uint64_t)(~((uint32_t)del[pi] ^ 0x80000000u)) << 32) | (uint32_t)pi;
        if (key < owner[pi]) owner[pi] = key;
        if (key < owner[pj]) owner[pj] = key;
    }
    /* winners: rank-2 updates from the snapshot, then the swaps (qap.cu:1203-1376) */
    memcpy(snap, perm, sizeof(int32_t) * (size_t)n);
    int32_t m = 0;
    for (int pi = 0; pi < n; pi++) {
        int32_t pj = prop[pi];
        int32_t won = pj != pi && (uint32_t)owner[pi] == (uint32_t)pi &&
                      (uint32_t)owner[pj] == (uint32_t)pi;
        if (tgt) tgt[pi] = pj;
        if (acc) acc[pi] = won;
        if (!won) continue;
        oracle_force_update(n, F, D, I, pi, pj, snap[pi], snap[pj]);
        m++;
    }
    for (int pi = 0; pi < n; pi++) {
        int32_t pj = prop[pi];
        if (pj == pi || (uint32_t)owner[pi] != (uint32_t)pi || (uint32_t)owner[pj] != (uint32_t)pi)
            continue;
        int32_t a = snap[pi], b = snap[pj];
        perm[pi] = b;
        perm[pj] = a;
        pinv[b] = pi;
        pinv[a] = pj;
    }
    /* exact cost from the updated force matrix (qap.cu:1418-1457) */
    int64_t sum = 0;
    for (int i = 0; i < n; i++) sum += I[i * n + perm[i]];
    *E = -sum / 2;
    return m;
}

/* Reference floating-point roulette (CPU_PIM.cpp:212-287), Philox-driven. */
static int32_t attempt_reffp(int32_t n, const int32_t *F, const int32_t *D, int32_t *perm,
                             int32_t *pinv, int32_t *I, int64_t *E, int32_t i, double beta2,
                             double r, int32_t *dE, long double *probs) {
    int32_t a = perm[i];
    int32_t maxd = 0;
    for (int b = 0; b < n; b++) {
        dE[b] = (int32_t)oracle_delta(n, F, D, perm, pinv, I, i, b);
        if (dE[b] > maxd) maxd = dE[b];
    }
    int32_t cut = cutoff_of(maxd, 20.0 / beta2);
    double probSum = 0.0;
    for (int b = 0; b < n; b++) {
        probs[b] = 0.0L;
        if (dE[b] > cut) {
            long double expr = beta2 * dE[b];
            probs[b] = exp2l(expr);
            probSum += probs[b];
        }
    }
    double scaledR = r * probSum;
    double cum = 0.0;
    for (int b = 0; b < n; b++) {
        cum += probs[b];
        if (cum >= scaledR) {
            if (b != a) commit(n, F, D, perm, pinv, I, E, i, b, dE[b]);
            return b;
        }
    }
    return a;
}

/* ------------------------------------------------------------------------- */
/* The wide path: the same update with an int64 force matrix, for instances whose   */
/* bounds exceed int32 (CUPIM's ReplicaPool<F, int64_t>, qap.cu:369-411 with        */
/* Z = int64_t).  Deltas, cutoff and the weight's input are int64; the weight      */
/* itself is the same float sequence from y = beta2f * (float)d on.                */
/* ------------------------------------------------------------------------- */
static void force_matrix64(int32_t n, const int32_t *F, const int32_t *D, const int32_t *perm,
                           int64_t *I) {
    int32_t *Ft = malloc(sizeof(int32_t) * (size_t)n * n);
    for (int x = 0; x < n; x++)
        for (int y = 0; y < n; y++) Ft[(size_t)x * n + y] = F[(size_t)y * n + x];
    for (int i = 0; i < n; i++) {
        int64_t *row = I + (size_t)i * n;
        for (int a = 0; a < n; a++) row[a] = 0;
        for (int j = 0; j < n; j++) {
            const int64_t dij = D[(size_t)i * n + j], dji = D[(size_t)j * n + i];
            const int32_t *fcol = Ft + (size_t)perm[j] * n, *frow = F + (size_t)perm[j] * n;
            for (int a = 0; a < n; a++) row[a] -= dij * fcol[a] + dji * frow[a];
        }
    }
    free(Ft);
}

static int64_t delta64(int32_t n, const int32_t *F, const int32_t *D, const int32_t *perm,
                       const int32_t *pinv, const int64_t *I, int32_t i, int32_t b) {
    const int32_t a = perm[i], j = pinv[b];
    const int64_t force = I[(size_t)i * n + b] + I[(size_t)j * n + a] - I[(size_t)i * n + a] -
                          I[(size_t)j * n + b];
    const int64_t dist = (int64_t)D[i * n + i] + D[j * n + j] - D[i * n + j] - D[j * n + i];
    const int64_t flow = (int64_t)F[a * n + a] + F[b * n + b] - F[a * n + b] - F[b * n + a];
    return force - dist * flow;
}

static void commit64(int32_t n, const int32_t *F, const int32_t *D, int32_t *perm, int32_t *pinv,
                     int64_t *I, int64_t *E, int32_t i, int32_t b, int64_t dEb) {
    const int32_t a = perm[i], j = pinv[b];
    *E -= dEb;
    perm[i] = b;
    pinv[b] = i;
    perm[j] = a;
    pinv[a] = j;
    int64_t *v = malloc(sizeof(int64_t) * 2 * (size_t)n);
    for (int c = 0; c < n; c++) {
        v[c] = (int64_t)F[(size_t)c * n + b] - F[(size_t)c * n + a];
        v[n + c] = (int64_t)F[(size_t)b * n + c] - F[(size_t)a * n + c];
    }
    for (int k = 0; k < n; k++) {
        const int64_t u1 = (int64_t)D[(size_t)k * n + j] - D[(size_t)k * n + i];
        const int64_t u2 = (int64_t)D[(size_t)j * n + k] - D[(size_t)i * n + k];
        int64_t *row = I + (size_t)k * n;
        for (int c = 0; c < n; c++) row[c] += u1 * v[c] + u2 * v[n + c];
    }
    free(v);
}

/* cutoff = (int64)(maxd - 20/beta2), saturated below (CPU_PIM.cpp:234-235) */
static int64_t cutoff64(int64_t maxd, double c20) {
    const double cd = (double)maxd - c20;
    if (cd < -9223372036854775808.0) return INT64_MIN;
    return (int64_t)cd;
}

uint32_t oracle_weight64(float beta2f, int64_t d) {
    return oracle_weight_y(beta2f * (float)d);
}

static int32_t attempt_exact64(int32_t n, const int32_t *F, const int32_t *D, int32_t *perm,
                               int32_t *pinv, int64_t *I, int64_t *E, int32_t i, float beta2f,
                               double c20, uint32_t u, int64_t *dE, uint32_t *w) {
    const int32_t a = perm[i];
    int64_t maxd = 0;
    for (int b = 0; b < n; b++) {
        dE[b] = delta64(n, F, D, perm, pinv, I, i, b);
        if (dE[b] > maxd) maxd = dE[b];
    }
    const int64_t cut = cutoff64(maxd, c20);
    uint32_t S = 0;
    for (int b = 0; b < n; b++) {
        w[b] = dE[b] > cut ? oracle_weight64(beta2f, dE[b] - maxd) : 0u;
        S += w[b];
    }
    const uint32_t T = (uint32_t)(((uint64_t)u * S) >> 32);
    uint32_t P = 0;
    int32_t sel = a;
    for (int b = 0; b < n; b++) {
        P += w[b];
        if (P > T) {
            sel = b;
            break;
        }
    }
    if (sel != a) commit64(n, F, D, perm, pinv, I, E, i, sel, dE[sel]);
    return sel;
}

static int32_t attempt_reffp64(int32_t n, const int32_t *F, const int32_t *D, int32_t *perm,
                               int32_t *pinv, int64_t *I, int64_t *E, int32_t i, double beta2,
                               double r, int64_t *dE, long double *probs) {
    const int32_t a = perm[i];
    int64_t maxd = 0;
    for (int b = 0; b < n; b++) {
        dE[b] = delta64(n, F, D, perm, pinv, I, i, b);
        if (dE[b] > maxd) maxd = dE[b];
    }
    const int64_t cut = cutoff64(maxd, 20.0 / beta2);
    double probSum = 0.0;
    for (int b = 0; b < n; b++) {
        probs[b] = 0.0L;
        if (dE[b] > cut) {
            probs[b] = exp2l((long double)beta2 * (long double)dE[b]);
            probSum += probs[b];
        }
    }
    const double scaledR = r * probSum;
    double cum = 0.0;
    for (int b = 0; b < n; b++) {
        cum += probs[b];
        if (cum >= scaledR) {
            if (b != a) commit64(n, F, D, perm, pinv, I, E, i, b, dE[b]);
            return b;
        }
    }
    return a;
}

/* ------------------------------------------------------------------------- */
/* Initial permutations (CPU_PIM.cpp:134-161).                                */
/* ------------------------------------------------------------------------- */
static void init_perm_tag(int32_t n, uint64_t seed, uint32_t stream, uint32_t tag,
                          int32_t *perm) {
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = 0; i < n; i++) {
        uint32_t x[4];
        ph_draw(seed, (uint32_t)i, stream, 0u, tag, x);
        int32_t j = (int32_t)(((uint64_t)x[0] * (uint64_t)n) >> 32);
        if (i != j) {
            int32_t t = perm[i];
            perm[i] = perm[j];
            perm[j] = t;
        }
    }
}

void oracle_init_perm(int32_t n, uint64_t seed, uint32_t stream, int32_t *perm) {
    init_perm_tag(n, seed, stream, 1u, perm);
}

/* ------------------------------------------------------------------------- */
/* Solve: setup + threadToRun for every ladder.                               */
/* ------------------------------------------------------------------------- */
typedef struct {
    int32_t *perm, *pinv, *I, *best_perm;
    int64_t *I64;           /* the wide path's force matrix (I is then unused) */
    int64_t E, bestE;
    uint64_t accepted;
    int32_t pos;
} rep_t;

typedef struct {
    const oracle_problem *prob;
    const oracle_params *par;
    int32_t mode;
    rep_t *reps;            /* R reps of this ladder */
    int32_t *order;         /* R */
    int32_t tss;            /* timeSinceSwap */
    int64_t ladderBest;
    uint64_t exchanges, kicks;
    int32_t ladder;
    const float *beta2f;    /* per position */
    const double *beta2;    /* per position */
    const double *c20;      /* per position */
    int32_t *dE;
    int64_t *dE64;          /* wide path */
    int32_t wide;
    uint32_t *w;
    int32_t *os_work;       /* 4n one-shot scratch */
    uint64_t *os_owner;     /* n */
    long double *probs;
    int64_t *xE, *xbE;      /* R scratch for the exchange step */
    int32_t *xpos;
    int64_t *traj;          /* traj_E base (may be NULL) */
    int64_t traj_cap;
    int32_t RT;
} ladder_t;

int32_t oracle_exchange_step(int32_t R, const int64_t *E, const int64_t *bestE, int32_t *order,
                             int32_t *pos, int32_t *tss, int64_t *ladder_best,
                             int32_t kick_after, int32_t *kicked) {
    int32_t nex = 0;
    (*tss)++;                                   /* CPU_PIM.cpp:295 */
    for (int l = 0; l < R; l++) {               /* :297-314 */
        if (bestE[l] < *ladder_best) {
            *ladder_best = bestE[l];
            *tss = 0;
        }
    }
    for (int k = 0; k < R - 1; k++) {           /* :324-337, hottest first, bubbling */
        int32_t r1 = order[k], r2 = order[k + 1];
        if (E[r1] < E[r2]) {
            order[k] = r2;
            order[k + 1] = r1;
            pos[r1] = k + 1;                    /* the beta swap of :332 */
            pos[r2] = k;
            nex++;
            if (r2 == R - 1) *tss = 0;          /* :333-335 tests a replica index */
        }
    }
    *kicked = 0;
    if (*tss > kick_after) {                    /* :339-344, not reset by the kick */
        int32_t r1 = order[R - 1], r2 = order[0];
        order[R - 1] = r2;
        order[0] = r1;
        pos[r1] = 0;
        pos[r2] = R - 1;
        *kicked = 1;
    }
    return nex;
}

/* Fold + exchange + kick of one ladder. */
static void ladder_exchange(ladder_t *L) {
    const int32_t R = L->par->R;
    int64_t *E = L->xE, *bE = L->xbE;
    int32_t *pos = L->xpos;
    for (int l = 0; l < R; l++) {
        E[l] = L->reps[l].E;
        bE[l] = L->reps[l].bestE;
        pos[l] = L->reps[l].pos;
    }
    int32_t kicked = 0;
    L->exchanges += (uint64_t)oracle_exchange_step(R, E, bE, L->order, pos, &L->tss,
                                                   &L->ladderBest, L->par->kick_after, &kicked);
    L->kicks += (uint64_t)kicked;
    for (int l = 0; l < R; l++) L->reps[l].pos = pos[l];
}

static void ladder_sweeps(ladder_t *L, int64_t m) {
    const oracle_problem *P = L->prob;
    const oracle_params *par = L->par;
    const int32_t n = P->n, R = par->R;
    for (int l = 0; l < R; l++) {
        rep_t *rp = &L->reps[l];
        const uint32_t stream = (uint32_t)(L->ladder * R + l);
        for (int s = 0; s < par->spe; s++) {
            const uint32_t sweep = (uint32_t)(m * par->spe + s);
            if (par->flags & ORACLE_FLAG_ONE_SHOT) {
                int32_t acc = oracle_one_shot_step(n, P->F, P->D, rp->perm, rp->pinv, rp->I,
                                                   &rp->E, L->beta2f[rp->pos], L->c20[rp->pos],
                                                   par->seed, sweep, stream, L->os_work,
                                                   L->os_owner, L->w, NULL, NULL);
                rp->accepted += (uint64_t)acc;
                if (acc > 0 && rp->E < rp->bestE) {
                    rp->bestE = rp->E;
                    memcpy(rp->best_perm, rp->perm, sizeof(int32_t) * (size_t)n);
                }
                continue;
            }
            for (int i = 0; i < n; i++) {
                uint32_t x[4];
                ph_draw(par->seed, (uint32_t)i, sweep, stream, 0u, x);
                int32_t a = rp->perm[i], sel;
                if (L->wide && L->mode == ORACLE_MODE_EXACT) {
                    sel = attempt_exact64(n, P->F, P->D, rp->perm, rp->pinv, rp->I64, &rp->E, i,
                                          L->beta2f[rp->pos], L->c20[rp->pos], x[0], L->dE64,
                                          L->w);
                } else if (L->wide) {
                    double r = (double)((((uint64_t)(x[0] >> 5)) << 26) | (uint64_t)(x[1] >> 6)) *
                               0x1p-53;
                    sel = attempt_reffp64(n, P->F, P->D, rp->perm, rp->pinv, rp->I64, &rp->E, i,
                                          L->beta2[rp->pos], r, L->dE64, L->probs);
                } else if (L->mode == ORACLE_MODE_EXACT) {
                    sel = oracle_attempt_exact(n, P->F, P->D, rp->perm, rp->pinv, rp->I, &rp->E, i,
                                               L->beta2f[rp->pos], L->c20[rp->pos], x[0], L->dE,
                                               L->w);
                } else {
                    double r = (double)((((uint64_t)(x[0] >> 5)) << 26) | (uint64_t)(x[1] >> 6)) *
                               0x1p-53;
                    sel = attempt_reffp(n, P->F, P->D, rp->perm, rp->pinv, rp->I, &rp->E, i,
                                        L->beta2[rp->pos], r, L->dE, L->probs);
                }
                if (sel != a) {
                    rp->accepted++;
                    if (rp->E < rp->bestE) { /* CPU_PIM.cpp:271-272 */
                        rp->bestE = rp->E;
                        memcpy(rp->best_perm, rp->perm, sizeof(int32_t) * (size_t)n);
                    }
                }
            }
        }
    }
    if (L->traj && m < L->traj_cap)
        for (int l = 0; l < R; l++) L->traj[m * L->RT + L->ladder * R + l] = L->reps[l].E;
}

typedef struct {
    ladder_t *ladders;
    int32_t first, count;
    int64_t megas;
} thread_job;

static void *thread_main(void *arg) {
    thread_job *job = (thread_job *)arg;
    for (int t = job->first; t < job->first + job->count; t++) {
        ladder_t *L = &job->ladders[t];
        for (int64_t m = 0; m < job->megas; m++) {
            ladder_exchange(L);
            ladder_sweeps(L, m);
        }
    }
    return NULL;
}

int oracle_solve(const oracle_problem *prob, const oracle_params *par, oracle_result *res,
                 int32_t mode, int32_t nthreads) {
    const int32_t n = prob->n, R = par->R, T = par->T;
    if (n < 1 || R < 1 || T < 1 || par->spe < 1 || par->sweeps < 0 || par->sweeps % par->spe)
        return -1;
    if (par->sweeps > (int64_t)0xffffffff) return -1;
    for (int k = 0; k < R; k++)
        if (!(par->beta_ln[k] > 0.0)) return -1;
    /* int32 force matrices when the bounds allow, else the wide (int64) path; one-shot
     * steps need small entries anyway */
    const int32_t wide = width_class(n, prob->F, prob->D);
    if (wide < 0 || (wide && (par->flags & ORACLE_FLAG_ONE_SHOT))) return -4;
    const int32_t RT = R * T;
    const int64_t megas = par->sweeps / par->spe;

    float *beta2f = malloc(sizeof(float) * R);
    double *beta2 = malloc(sizeof(double) * R), *c20 = malloc(sizeof(double) * R);
    for (int k = 0; k < R; k++) {
        beta2[k] = oracle_beta2(par->beta_ln[k]);
        beta2f[k] = (float)beta2[k];
        c20[k] = 20.0 / beta2[k];
    }
    rep_t *reps = calloc((size_t)RT, sizeof(rep_t));
    ladder_t *lad = calloc((size_t)T, sizeof(ladder_t));
    int32_t *order = malloc(sizeof(int32_t) * RT);
    int32_t *init = malloc(sizeof(int32_t) * n);
    for (int t = 0; t < T; t++) {
        if (!(par->flags & ORACLE_FLAG_INIT_PER_REPLICA)) {
            if (par->init_perm) memcpy(init, par->init_perm, sizeof(int32_t) * n);
            else init_perm_tag(n, par->seed, (uint32_t)t, 1u, init);
        }
        for (int l = 0; l < R; l++) {
            rep_t *rp = &reps[t * R + l];
            rp->perm = malloc(sizeof(int32_t) * n);
            rp->pinv = malloc(sizeof(int32_t) * n);
            rp->best_perm = malloc(sizeof(int32_t) * n);
            if (wide) rp->I64 = malloc(sizeof(int64_t) * (size_t)n * n);
            else rp->I = malloc(sizeof(int32_t) * (size_t)n * n);
            if (par->flags & ORACLE_FLAG_INIT_PER_REPLICA) {
                if (par->init_perm) memcpy(rp->perm, par->init_perm, sizeof(int32_t) * n);
                else init_perm_tag(n, par->seed, (uint32_t)(t * R + l), 2u, rp->perm);
            } else {
                memcpy(rp->perm, init, sizeof(int32_t) * n);
            }
            for (int i = 0; i < n; i++) rp->pinv[rp->perm[i]] = i;
            memcpy(rp->best_perm, rp->perm, sizeof(int32_t) * n);
            if (wide) force_matrix64(n, prob->F, prob->D, rp->perm, rp->I64);
            else oracle_force_matrix(n, prob->F, prob->D, rp->perm, rp->I);
            rp->E = oracle_energy(n, prob->F, prob->D, rp->perm);
            rp->bestE = INT64_MAX; /* CPU_PIM.cpp:123 (INT_MAX) */
            rp->pos = l;           /* CPU_PIM.cpp:115-116 */
            order[t * R + l] = l;
        }
        ladder_t *L = &lad[t];
        L->prob = prob;
        L->par = par;
        L->mode = mode;
        L->reps = &reps[t * R];
        L->order = &order[t * R];
        L->ladderBest = INT64_MAX;
        L->ladder = t;
        L->beta2f = beta2f;
        L->beta2 = beta2;
        L->c20 = c20;
        L->dE = malloc(sizeof(int32_t) * n);
        L->dE64 = malloc(sizeof(int64_t) * n);
        L->wide = wide;
        L->w = malloc(sizeof(uint32_t) * n);
        L->os_work = malloc(sizeof(int32_t) * 4 * (size_t)n);
        L->os_owner = malloc(sizeof(uint64_t) * (size_t)n);
        L->probs = malloc(sizeof(long double) * n);
        L->xE = malloc(sizeof(int64_t) * R);
        L->xbE = malloc(sizeof(int64_t) * R);
        L->xpos = malloc(sizeof(int32_t) * R);
        L->traj = res->traj_E;
        L->traj_cap = res->traj_E ? res->traj_cap : 0;
        L->RT = RT;
    }

    int64_t megas_done = 0;
    if (par->has_goal || nthreads <= 1 || T == 1) {
        /* Ladders advance in lockstep so the goal test sees every ladder's fold
         * (CPU_PIM.cpp:346-348 uses the process-wide singleBestE). */
        int64_t globalBest = INT64_MAX;
        for (int64_t m = 0; m < megas; m++) {
            for (int t = 0; t < T; t++) {
                ladder_exchange(&lad[t]);
                if (lad[t].ladderBest < globalBest) globalBest = lad[t].ladderBest;
            }
            if (par->has_goal && globalBest <= par->goal) break;
            for (int t = 0; t < T; t++) ladder_sweeps(&lad[t], m);
            megas_done = m + 1;
        }
    } else {
        int32_t nt = nthreads < T ? nthreads : T;
        pthread_t *th = malloc(sizeof(pthread_t) * nt);
        thread_job *jobs = malloc(sizeof(thread_job) * nt);
        int32_t per = T / nt, extra = T % nt, first = 0;
        for (int k = 0; k < nt; k++) {
            jobs[k].ladders = lad;
            jobs[k].first = first;
            jobs[k].count = per + (k < extra ? 1 : 0);
            jobs[k].megas = megas;
            first += jobs[k].count;
            pthread_create(&th[k], NULL, thread_main, &jobs[k]);
        }
        for (int k = 0; k < nt; k++) pthread_join(th[k], NULL);
        free(th);
        free(jobs);
        megas_done = megas;
    }

    /* Result: best over every replica's bestE; ties go to the lowest replica index.
     * With no accepted move anywhere, the best initial state is returned. */
    int64_t best = INT64_MAX;
    int32_t bi = -1;
    for (int r = 0; r < RT; r++)
        if (reps[r].bestE < best) {
            best = reps[r].bestE;
            bi = r;
        }
    if (bi < 0) {
        int64_t e0 = INT64_MAX;
        for (int r = 0; r < RT; r++) {
            int64_t e = oracle_energy(n, prob->F, prob->D, reps[r].best_perm);
            if (e < e0) {
                e0 = e;
                bi = r;
            }
        }
        best = e0;
    }
    res->best_cost = best;
    if (res->best_perm) memcpy(res->best_perm, reps[bi].best_perm, sizeof(int32_t) * n);
    res->megas_done = megas_done;
    res->attempts = (uint64_t)megas_done * (uint64_t)par->spe * (uint64_t)n * (uint64_t)RT;
    res->accepted = 0;
    res->exchanges = 0;
    res->kicks = 0;
    for (int r = 0; r < RT; r++) {
        res->accepted += reps[r].accepted;
        if (res->final_perm) memcpy(res->final_perm + (size_t)r * n, reps[r].perm, sizeof(int32_t) * n);
        if (res->best_perms)
            memcpy(res->best_perms + (size_t)r * n, reps[r].best_perm, sizeof(int32_t) * n);
        if (res->final_E) res->final_E[r] = reps[r].E;
        if (res->final_bestE) res->final_bestE[r] = reps[r].bestE;
        if (res->final_pos) res->final_pos[r] = reps[r].pos;
        if (res->final_accepted) res->final_accepted[r] = reps[r].accepted;
    }
    for (int t = 0; t < T; t++) {
        res->exchanges += lad[t].exchanges;
        res->kicks += lad[t].kicks;
        free(lad[t].dE);
        free(lad[t].dE64);
        free(lad[t].w);
        free(lad[t].os_work);
        free(lad[t].os_owner);
        free(lad[t].probs);
        free(lad[t].xE);
        free(lad[t].xbE);
        free(lad[t].xpos);
    }
    for (int r = 0; r < RT; r++) {
        free(reps[r].perm);
        free(reps[r].pinv);
        free(reps[r].best_perm);
        free(reps[r].I);
        free(reps[r].I64);
    }
    free(reps);
    free(lad);
    free(order);
    free(init);
    free(beta2f);
    free(beta2);
    free(c20);
    return 0;
}
