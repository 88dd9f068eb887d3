// qap_pt_device.h -- device-side building blocks of the PT-SA sweep for gfx950:
// Philox4x32-10, the exact integer heat-bath weight, and the descriptors the host
// fills.  Everything here is exact integer or a fixed IEEE float sequence, so the
// result of one p-dit update does not depend on launch shape or scheduling.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qap {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Counter-based: the draw for
// one p-dit update is a pure function of (seed, location i, sweep, replica), so a
// replica's trajectory is independent of where and with whom it runs.
//   sweep attempt : ctr = {i, sweep, replica_stream, 0}
//   ladder init   : ctr = {i, ladder, 0, 1}
//   replica init  : ctr = {i, replica_stream, 0, 2}
//   key           = {seed & 0xffffffff, seed >> 32}
// ---------------------------------------------------------------------------
struct u32x4 {
    uint32_t x, y, z, w;
};

__host__ __device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = u32x4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1,
                  (uint32_t)p0};
    }
    return c;
}

// ---------------------------------------------------------------------------
// Heat-bath weight w = floor(2^20 * 2^y), y = beta2f * d, d = dE - max dE <= 0.
// 2^y = 2^k * 2^f, k = rint(y), f = y - k in [-1/2, 1/2]; 2^f = e^(f ln2) from its
// Taylor series through degree 7 (Horner with fused multiply-add), coefficients
// (ln 2)^m / m! rounded to float.  Truncation error < 2e-8 (below the 2^-20
// resolution of the weights).  Every step is a correctly rounded IEEE operation
// (mul, rint, sub, fma, exact power-of-two scaling, truncating conversion), so the
// host restatement of the same sequence gives the same integer.
// The maximum candidate gets exactly 2^20; the sum over n <= 1024 candidates stays
// below 2^31.
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t heat_weight_y(float y) {
#pragma clang fp contract(off)
    if (!(y > -32.0f)) return 0u;
    const float kf = rintf(y);
    const float f = y - kf;
    float p = 0x1.ffcbfcp-17f;
    p = fmaf(p, f, 0x1.430912p-13f);
    p = fmaf(p, f, 0x1.5d87fep-10f);
    p = fmaf(p, f, 0x1.3b2ab6p-7f);
    p = fmaf(p, f, 0x1.c6b08ep-5f);
    p = fmaf(p, f, 0x1.ebfbe0p-3f);
    p = fmaf(p, f, 0x1.62e430p-1f);
    p = fmaf(p, f, 0x1p+0f);
    const int e = (int)kf + 20;
    if (e < 0) return 0u;
    return (uint32_t)(p * (float)(1u << e));
}

__host__ __device__ inline uint32_t heat_weight(float beta2f, int32_t d) {
#pragma clang fp contract(off)
    return heat_weight_y(beta2f * (float)d);
}

// The same weight for a 64-bit difference (force matrices beyond int32): (float)d is the
// correctly rounded conversion, and the float steps from y on are the same.
__host__ __device__ inline uint32_t heat_weight64(float beta2f, int64_t d) {
#pragma clang fp contract(off)
    return heat_weight_y(beta2f * (float)d);
}

// CPU_PIM.cpp:234-235: cutoff = (int)(maxDeltaE - 20.0 / beta), saturated below.
__host__ __device__ inline int32_t heat_cutoff(int32_t maxd, double c20) {
    const double cd = (double)maxd - c20;
    if (cd < -2147483648.0) return INT32_MIN;
    return (int32_t)cd;
}

// The same cutoff in 64 bits (force matrices beyond int32).
__host__ __device__ inline int64_t heat_cutoff64(int64_t maxd, double c20) {
    const double cd = (double)maxd - c20;
    if (cd < -9223372036854775808.0) return INT64_MIN;
    return (int64_t)cd;
}

// ---------------------------------------------------------------------------
// Descriptors (host writes, device reads).
// ---------------------------------------------------------------------------
struct InstDesc {
    int32_t n, ld;            // size, row stride of the force matrix in HBM (elements)
    int32_t ld_s;             // row stride of the LDS copy (elements, odd word stride)
    int32_t sparse_upd;       // compact the nonzero flow-difference columns on update
    int32_t mul24;            // |u|, |v| of the rank-2 update fit the 24-bit multiplier
    int32_t R, T;             // replicas per ladder, ladders
    int32_t spe;              // sweeps per exchange
    int32_t kick_after;
    int32_t has_goal;
    int32_t sym;              // F and D symmetric with zero diagonal
    int32_t xmirror;          // persistent launches mirror the exchange step in every
                              //   workgroup (R*T <= 64, every cost below 2^31 - 1)
    int64_t goal;
    int64_t mat_off;          // F/D/Ft/Dt element offset
    int64_t vec_off;          // Fdiag/Ddiag element offset (n)
    int64_t p_base;           // perm/pinv/best_perm element offset of replica 0
    int64_t I_base;           // force-matrix element offset of replica 0
    int64_t I64_base;         // int64 force matrices (instances beyond int32), replica 0
    int32_t rep_base;         // first global replica index (also order[] offset)
    int32_t ladder_base;      // first global ladder index
    int32_t beta_off;         // per-position tables offset
    uint32_t seed_lo, seed_hi;
};

struct InstState {
    int64_t megas_done;
    int64_t mega_limit;
    int64_t global_best;
    uint64_t exchanges, kicks;
    uint32_t sweep_base;
    int32_t active;           // sweeps of the current mega-iteration run
    int32_t halted;           // goal reached or cancelled
    int32_t pad;
};

struct PoolDev {
    const InstDesc *inst;
    InstState *state;
    const int32_t *rep_inst;  // global replica -> instance
    const int32_t *F, *D, *Ft, *Dt;
    const int16_t *F16, *D16; // F and D saturated to int16 (read by the lazy kernel, whose
                              //   instances have every entry within 16 bits)
    const int32_t *Fdiag, *Ddiag;
    const int32_t *fclass;    // per facility: low 16 bits equal <=> identical flow row and
                              //   column; high 16 bits equal <=> twins (symmetric flow rows
                              //   equal outside the two facilities' own columns)
    const int32_t *fspan;     // per facility: nonzero columns of its flow row and column,
                              //   lo | hi << 10 | count << 20 (empty: lo 1023, hi 0, 0)
    const float *beta2f;      // per ladder position (exp2 units, float)
    const double *c20;        // per ladder position: 20 / beta2 (double)
    int32_t *perm, *pinv, *best_perm, *I;
    int64_t *I64;             // force matrices of instances whose bounds exceed int32
    int64_t *E, *bestE;
    uint64_t *accepted;
    uint64_t *cells;          // force-matrix entries rewritten by rank-2 updates
    uint64_t *fcells;         // force-matrix entries written by flushes of pending updates
                              //   (lazy placement; the second half of the cells buffer)
    int32_t *pos;             // ladder position of each replica
    int32_t *order;           // per ladder: position -> replica (local index)
    int32_t *tss;             // per ladder: timeSinceSwap
    int64_t *ladder_best;     // per ladder
    uint32_t *sync;           // persistent sweep launches: [0] fault word (a wait that gave
                              //   up), then per instance [1 + 2 k] arrivals and [2 + 2 k] the
                              //   number of exchange steps run inside persistent launches
    uint64_t *xslot;          // persistent launches with mirrored exchange steps: per replica
                              //   two parities x {cost, best}, each tag << 32 | value
                              //   (null where the pool has none)
};

}  // namespace qap
