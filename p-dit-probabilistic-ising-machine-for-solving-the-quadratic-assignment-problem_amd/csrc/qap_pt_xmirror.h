// qap_pt_xmirror.h -- the exchange step inside a persistent sweep launch, mirrored by every
// workgroup of the instance (included by qap_pt_sweep.hip).
//
// The step (CPU_PIM.cpp:293-357: best fold, bubble pass, kick, goal test, interval
// bookkeeping) is a deterministic function of the ladder state and of every replica's cost
// and best.  Inside a persistent launch nothing else writes that state -- host halts and
// limits change only between launches -- so each workgroup keeps its own copy of its
// instance's ladder state in wave 0 (lane t*R + k: the replica at position k of ladder t;
// lane t: ladder t's best and time since its last improvement) and runs the step itself.
// Replicas exchange only their costs: at the end of an interval each workgroup stores its
// cost and best, each a 64-bit word holding a 32-bit tag and the 32-bit value, into its
// slot, and polls the slots of the instance's R*T replicas (a lane each) until every tag
// is the interval's.  Compared with a counter and one exchanging workgroup this takes two
// memory round trips between workgroups instead of about eight, and no fence: every word a
// workgroup needs comes in the tagged words themselves.
//
// Preconditions, checked on the host (InstDesc::xmirror): R*T <= 64 and every cost below
// 2^31 - 1, so a cost fits the low half of a word and INT32_MAX can stand for "no best
// yet" (INT64_MAX).  The tag is the instance's exchange count before the step
// (megas_done), which grows by one per step and never repeats.  Slots are double-buffered
// by its parity: a workgroup can only write tag m + 2 after every replica has published
// tag m + 1, which each does only after reading all of tag m.
#pragma once

#include "qap_pt_device.h"
#include "qap_pt_sweep_common.h"

namespace qap {

__device__ __forceinline__ int64_t readlane64(int64_t v, int lane) {
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), lane);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// minimum over the wave, every lane gets it
__device__ __forceinline__ int64_t wave_min64(int64_t v) {
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int64_t o = __shfl_xor(v, off);
        v = o < v ? o : v;
    }
    return v;
}

// Longest a persistent workgroup waits for the rest of its ladder at an exchange step:
// 30 s of the 100 MHz real-time counter.  The launch is cooperative (every workgroup
// resident), so the wait ends when the slowest replica finishes its interval; the limit
// only turns a fault into a reported error (the fault word) instead of a hang.
constexpr uint64_t kPersistWaitTicks = 3000000000ull;

// slot word: tag in the high half, the 32-bit value in the low half
__device__ __forceinline__ uint64_t xm_word(uint32_t tag, int64_t v) {
    const int32_t lo = v >= (int64_t)INT32_MAX ? INT32_MAX : (int32_t)v;
    return ((uint64_t)tag << 32) | (uint32_t)lo;
}
// the two words of replica g for the step with exchange count m
__device__ __forceinline__ uint64_t *xm_slot(const PoolDev &P, int g, int64_t m) {
    return P.xslot + 4 * (int64_t)g + 2 * (m & 1);
}

// Every wave's copy of the instance's per-position tables (lane l < R: position l), so that
// a new position after a mirrored step reads its temperature from registers
struct XBeta {
    float beta2f;
    double c20;
};
__device__ __forceinline__ XBeta xm_beta_load(const PoolDev &P, const InstDesc &d, int lane) {
    XBeta t{0.f, 0.0};
    if (lane < d.R) {
        t.beta2f = P.beta2f[d.beta_off + lane];
        t.c20 = P.c20[d.beta_off + lane];
    }
    return t;
}
__device__ __forceinline__ void xm_beta_at(const XBeta &t, int pos, float &beta2f, double &c20) {
    const int p = __builtin_amdgcn_readfirstlane(pos);  // pos is the same in every lane
    beta2f = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t.beta2f), p));
    const uint64_t cb = (uint64_t)__double_as_longlong(t.c20);
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)cb, p);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(cb >> 32), p);
    c20 = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// wave 0's copy of one instance's exchange state
struct XMirror {
    int32_t ordv;      // lane t*R + k < R*T: replica (ladder-local) at position k of ladder t
    int32_t tssv;      // lane t < T: time since the ladder's best or its coldest rung improved
    int64_t lbestv;    // lane t < T: the ladder's best
    int64_t gbest;     // the instance's best (wave-uniform from here on)
    uint64_t exchanges, kicks;
    int64_t megas_done, mega_limit;
    uint32_t sweep_base;
    int32_t halted;
};

__device__ __forceinline__ XMirror xm_load(const PoolDev &P, int inst, const InstDesc &d, int lane) {
    XMirror x;
    const int RT = d.R * d.T;
    x.ordv = lane < RT ? P.order[d.rep_base + lane] : 0;
    x.tssv = lane < d.T ? P.tss[d.ladder_base + lane] : 0;
    x.lbestv = lane < d.T ? P.ladder_best[d.ladder_base + lane] : INT64_MAX;
    const InstState st = P.state[inst];
    x.gbest = st.global_best;
    x.exchanges = st.exchanges;
    x.kicks = st.kicks;
    x.megas_done = st.megas_done;
    x.mega_limit = st.mega_limit;
    x.sweep_base = st.sweep_base;
    x.halted = st.halted;
    return x;
}

// The state the launch leaves behind, as exchange_instance would have written it (wave 0
// of the instance's replica 0 only).
__device__ __forceinline__ void xm_store(const PoolDev &P, int inst, const InstDesc &d,
                                         const XMirror &x, bool active, int lane) {
    const int RT = d.R * d.T;
    if (lane < RT) {
        const int base = lane - lane % d.R;
        P.order[d.rep_base + lane] = x.ordv;
        P.pos[d.rep_base + base + x.ordv] = lane - base;
    }
    if (lane < d.T) {
        P.tss[d.ladder_base + lane] = x.tssv;
        P.ladder_best[d.ladder_base + lane] = x.lbestv;
    }
    if (lane == 0) {
        InstState &so = P.state[inst];
        so.global_best = x.gbest;
        so.exchanges = x.exchanges;
        so.kicks = x.kicks;
        so.megas_done = x.megas_done;
        so.sweep_base = x.sweep_base;
        so.halted = x.halted;
        so.active = active ? 1 : 0;
    }
}

// Outcome of a mirrored step: kXmRun (sweeps follow, at pos / sweep_base), kXmStop (limit
// or goal: the launch ends for the instance), kXmFault (a wait gave up; fault word set).
constexpr int kXmRun = 0, kXmStop = 1, kXmFault = 2;

// One exchange step, run by wave 0 of every workgroup of the instance.  E and bestE are
// this replica's (local index `local`); on kXmRun *pos is its new ladder position.
__device__ int xm_exchange(const PoolDev &P, const InstDesc &d, XMirror &x, int g, int local,
                           int64_t E, int64_t bestE, int lane, int *pos) {
    // exchange_instance's first test, on state every workgroup holds alike
    if (x.halted || x.megas_done >= x.mega_limit) return kXmStop;
    const int R = d.R, T = d.T, RT = R * T;
    const int64_t m = x.megas_done;
    const uint32_t tag = (uint32_t)m;
    if (lane == 0) {
        uint64_t *s = xm_slot(P, g, m);
        __hip_atomic_store(s, xm_word(tag, E), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(s + 1, xm_word(tag, bestE), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // every replica's words, lane r for replica r of the instance
    const bool live = lane < RT;
    uint64_t *s = xm_slot(P, d.rep_base + (live ? lane : 0), m);
    uint64_t w0 = 0, w1 = 0;
    bool have = !live;
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    for (;;) {
        if (!have) {
            w0 = __hip_atomic_load(s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            w1 = __hip_atomic_load(s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            have = (uint32_t)(w0 >> 32) == tag && (uint32_t)(w1 >> 32) == tag;
        }
        if (__ballot(!have) == 0) break;
        if (__builtin_amdgcn_s_memrealtime() - t0 > kPersistWaitTicks) {
            if (lane == 0)
                __hip_atomic_store(&P.sync[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return kXmFault;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    // the words' low halves: costs (every one below INT32_MAX) and bests (INT32_MAX: none
    // yet), compared as int32 -- the same order as the int64 values they stand for
    const int32_t e_rep = live ? (int32_t)(uint32_t)w0 : INT32_MAX;  // by replica
    const int32_t b_rep = live ? (int32_t)(uint32_t)w1 : INT32_MAX;
    // cost at each ladder position: lane t*R + k takes replica ord[t*R + k]'s
    const int lbase = live ? lane - lane % R : 0;
    const int32_t e_pos = __shfl(e_rep, lbase + (live ? x.ordv : 0));

    int64_t gbest = INT64_MAX;
    uint32_t nex = 0, nkick = 0;
    for (int t = 0; t < T; ++t) {  // CPU_PIM.cpp:297-344, per ladder, as exchange_instance
        const int base = t * R;
        const bool inl = lane >= base && lane < base + R;
        // minimum of the ladder's bests: a DPP maximum of the negated values
        const int32_t mb32 = -lane63(wave_incl_max(inl ? -b_rep : INT32_MIN));
        const int64_t mb = mb32 == INT32_MAX ? INT64_MAX : (int64_t)mb32;
        int32_t tss = __builtin_amdgcn_readlane(x.tssv, t) + 1;
        int64_t lbest = readlane64(x.lbestv, t);
        if (mb < lbest) {  // best fold
            lbest = mb;
            tss = 0;
        }
        int newR = x.ordv;
        if (T == 1) {
            // The bubble pass in parallel, lane k = position k.  The sequential pass carries
            // an element along the ladder and hands it the next one's place whenever the next
            // costs at most as much, so the element carried into step k is the last position
            // <= k whose cost equals the prefix minimum (a "record"); step k puts the next
            // element at k when k + 1 is not a record (a swap: cE < nE), else the carry.
            const int32_t ek = lane < R ? e_pos : INT32_MAX;
            const int32_t pmin = -wave_incl_max(-ek);  // costs stay above INT32_MIN
            const bool rec = lane < R && ek == pmin;
            const int carry = wave_incl_max(rec ? lane : -1);
            const int rcar = __shfl(x.ordv, carry);
            const int rnext = __shfl(x.ordv, lane + 1);
            const bool recn = __shfl((int)rec, lane + 1) != 0;
            const bool swap = lane < R - 1 && !recn;
            newR = lane < R - 1 ? (swap ? rnext : rcar) : (lane == R - 1 ? rcar : newR);
            nex += (uint32_t)__popcll(__ballot(swap));
            if (__ballot(swap && rnext == R - 1)) tss = 0;
        } else {
            int32_t cE = __builtin_amdgcn_readlane(e_pos, base);
            int cR = __builtin_amdgcn_readlane(x.ordv, base);
            for (int k = 0; k < R - 1; ++k) {  // bubble pass, step by step
                const int32_t nE = __builtin_amdgcn_readlane(e_pos, base + k + 1);
                const int nR = __builtin_amdgcn_readlane(x.ordv, base + k + 1);
                if (cE < nE) {
                    newR = lane == base + k ? nR : newR;
                    ++nex;
                    if (nR == R - 1) tss = 0;
                } else {
                    newR = lane == base + k ? cR : newR;
                    cE = nE;
                    cR = nR;
                }
            }
            newR = lane == base + R - 1 ? cR : newR;
        }
        if (tss > d.kick_after) {  // kick: the hottest and coldest positions trade
            const int r0 = __builtin_amdgcn_readlane(newR, base);
            const int rL = __builtin_amdgcn_readlane(newR, base + R - 1);
            newR = lane == base ? rL : (lane == base + R - 1 ? r0 : newR);
            ++nkick;
        }
        x.ordv = inl ? newR : x.ordv;
        x.tssv = lane == t ? tss : x.tssv;
        x.lbestv = lane == t ? lbest : x.lbestv;
        gbest = lbest < gbest ? lbest : gbest;
    }
    x.gbest = gbest < x.gbest ? gbest : x.gbest;
    x.exchanges += nex;
    x.kicks += nkick;
    if (d.has_goal && x.gbest <= d.goal) {  // CPU_PIM.cpp:346-348
        x.halted = 1;
        return kXmStop;
    }
    x.sweep_base = (uint32_t)(x.megas_done * d.spe);
    x.megas_done += 1;
    // this replica's new position: the lane of its ladder now holding it
    const int tm = local / R, rm = local - tm * R;
    const uint64_t at = __ballot(lane >= tm * R && lane < tm * R + R && x.ordv == rm);
    *pos = __ffsll((unsigned long long)at) - 1 - tm * R;
    return kXmRun;
}

}  // namespace qap
