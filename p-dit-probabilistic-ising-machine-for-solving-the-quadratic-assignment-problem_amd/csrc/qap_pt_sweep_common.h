// qap_pt_sweep_common.h -- wave64 building blocks shared by the sweep kernels
// (qap_pt_sweep.hip, qap_pt_sweep_lazy.hip): DPP scans, LDS-only barriers, and the
// heat-bath pick of the update rule (CPU_PIM.cpp:234-251 with integer weights).
#pragma once

#include "qap_pt_device.h"

namespace qap {

constexpr int kWave = 64;
using v4i = int __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Wave64 inclusive scans with DPP.  update_dpp(old, src, ctrl, row_mask, bank_mask,
// bound_ctrl=false) returns `old` for lanes whose source lane is outside the row or
// whose row is masked off, so `old` is the identity of the operator.
// ---------------------------------------------------------------------------
template <int CTRL, int RMASK>
__device__ __forceinline__ int dpp_mov(int v, int identity) {
    return __builtin_amdgcn_update_dpp(identity, v, CTRL, RMASK, 0xf, false);
}

__device__ __forceinline__ int wave_incl_max(int v) {
    v = max(v, dpp_mov<0x111, 0xf>(v, INT32_MIN));  // row_shr:1
    v = max(v, dpp_mov<0x112, 0xf>(v, INT32_MIN));  // row_shr:2
    v = max(v, dpp_mov<0x114, 0xf>(v, INT32_MIN));  // row_shr:4
    v = max(v, dpp_mov<0x118, 0xf>(v, INT32_MIN));  // row_shr:8
    v = max(v, dpp_mov<0x142, 0xa>(v, INT32_MIN));  // row_bcast:15 -> rows 1,3
    v = max(v, dpp_mov<0x143, 0xc>(v, INT32_MIN));  // row_bcast:31 -> rows 2,3
    return v;
}

__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t u) {
    int v = (int)u;
    v += dpp_mov<0x111, 0xf>(v, 0);
    v += dpp_mov<0x112, 0xf>(v, 0);
    v += dpp_mov<0x114, 0xf>(v, 0);
    v += dpp_mov<0x118, 0xf>(v, 0);
    v += dpp_mov<0x142, 0xa>(v, 0);
    v += dpp_mov<0x143, 0xc>(v, 0);
    return (uint32_t)v;
}

__device__ __forceinline__ int lane63(int v) { return __builtin_amdgcn_readlane(v, kWave - 1); }

__host__ __device__ constexpr int round4(int x) { return (x + 3) & ~3; }
__host__ __device__ constexpr int round8(int x) { return (x + 7) & ~7; }

// Workgroup barrier that orders LDS only.  __syncthreads() also waits for every
// outstanding global load (vmcnt(0)), which would stall the row prefetches; the sweep's
// barriers only publish LDS data, so they fence the LDS address space alone.  With a
// single wave no s_barrier is needed; the fence keeps the compiler's order.
template <int NW>
__device__ __forceinline__ void lds_sync() {
    if constexpr (NW == 1) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront", "local");
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    }
}

// Heat-bath pick of one p-dit update among the candidates a workgroup owns (candidate
// b = tid + q * NT): maximum delta (at least 0, the stay), integer weights with the
// cutoff, prefix sums in candidate order, and the owner of the first candidate whose
// inclusive prefix exceeds floor(ri * S / 2^32).  red_s slots [0,16) and [16,80) carry
// the per-wave partials; every lane returns S and whether it owns the winner.
//
// With `shift_s` set, the deltas arrive without one term that every candidate shares,
// which is read from LDS (*shift_s) only after the first barrier: the pick subtracts it
// from the deltas of the real candidates (those above INT32_MIN) and from their maximum.
// A value written to *shift_s before the pick is then published by the pick's own
// barrier, so the caller needs no barrier of its own for it.
struct Pick {
    bool owner;
    int b, j;     // winning facility and its location (a and i for the stay)
    int32_t d;    // its delta
    uint32_t S;   // total weight (0: every weight truncated, the update stays)
};

template <int NT, int CPL>
__device__ __forceinline__ Pick heat_bath_pick(int32_t (&dE)[CPL], int lmax,
                                               const int32_t (&pv)[CPL], uint32_t ri,
                                               float beta2f, double c20, int a, int i,
                                               int32_t *red_s, int tid,
                                               const int32_t *shift_s = nullptr) {
    constexpr int NW = NT / kWave;
    const int lane = tid & (kWave - 1), wid = tid / kWave;
    int maxd;
    {
        const int wm = wave_incl_max(lmax);
        if constexpr (NW == 1) {
            maxd = lane63(wm);
        } else {
            if (lane == kWave - 1) red_s[wid] = wm;
            lds_sync<NW>();
            maxd = red_s[0];
#pragma unroll
            for (int w = 1; w < NW; ++w) maxd = max(maxd, red_s[w]);
        }
    }
    if (shift_s) {  // lmax was taken over the unshifted deltas (and INT32_MIN lanes)
        const int32_t sh = *shift_s;
        maxd = max(0, maxd - sh);
#pragma unroll
        for (int q = 0; q < CPL; ++q)
            if (dE[q] != INT32_MIN) dE[q] -= sh;
    }
    const int32_t cut = heat_cutoff(maxd, c20);
    uint32_t w[CPL], incl[CPL], qbase[CPL], wbase[CPL];
    uint32_t S = 0;
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
        w[q] = (dE[q] > cut) ? heat_weight(beta2f, dE[q] - maxd) : 0u;
        incl[q] = wave_incl_sum(w[q]);
    }
    if constexpr (NW == 1) {
#pragma unroll
        for (int q = 0; q < CPL; ++q) {
            qbase[q] = S;
            wbase[q] = 0;
            S += (uint32_t)lane63((int)incl[q]);
        }
    } else {
        if (lane == kWave - 1) {
#pragma unroll
            for (int q = 0; q < CPL; ++q) red_s[16 + q * 16 + wid] = (int)incl[q];
        }
        lds_sync<NW>();
#pragma unroll
        for (int q = 0; q < CPL; ++q) {
            uint32_t tq = 0, before = 0;
#pragma unroll
            for (int ww = 0; ww < NW; ++ww) {
                const uint32_t x = (uint32_t)red_s[16 + q * 16 + ww];
                before += (ww < wid) ? x : 0u;
                tq += x;
            }
            qbase[q] = S;
            wbase[q] = before;
            S += tq;
        }
    }
    const uint32_t thr = (uint32_t)(((uint64_t)ri * (uint64_t)S) >> 32);
    Pick pk{false, a, i, 0, S};
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
        const uint32_t ex = qbase[q] + wbase[q] + incl[q] - w[q];
        if (ex <= thr && thr < ex + w[q]) {
            pk.owner = true;
            pk.b = tid + q * NT;
            pk.j = pv[q];
            pk.d = dE[q];
        }
    }
    return pk;
}

}  // namespace qap
