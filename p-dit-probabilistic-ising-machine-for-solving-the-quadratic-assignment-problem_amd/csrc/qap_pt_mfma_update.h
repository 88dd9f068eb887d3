// qap_pt_mfma_update.h -- I += U V^T for an int32 force matrix in global memory, with the
// low-rank factors U, V as 16-bit vectors in LDS, on the int8 matrix cores.
//
// Used by the lazy sweep kernel (a batch of pending rank-2 updates, CPU_PIM.cpp:263-269)
// and by the one-shot step (the rank-2 updates of every applied swap,
// GPU-CUPIM/perl-gpu/src/lib/qap/qap.cu:1301-1376).  Exactness: every 16-bit entry x
// splits into two signed bytes x = lo + 256 hi (entries need |x| <= 32639 so that hi
// stays within int8); the int8 products are exact in the int32 accumulators and are
// recombined with shifts in 32-bit two's complement, which gives the exact entries
// whenever the result fits (the host checks the force bound).
#pragma once

#include "qap_pt_sweep_common.h"

namespace qap {

using v4h = int __attribute__((ext_vector_type(4)));  // 8 packed int16

constexpr int kUpdTiles = 4;  // 16-column tiles a wave updates per pass over the rows
constexpr int kUpdDepth = 4;  // row tiles in flight per wave (default)

// Signed byte planes of 8 packed int16 entries x = lo + 256 hi: lo is the low byte read
// as int8, hi = (x - lo) / 256 = high byte + bit 7 of the low byte.  Bytes keep the
// entry order (the same order for both operands is all the product needs).
__device__ __forceinline__ void split_bytes(v4h x, long &lo, long &hi) {
    uint32_t l[2], h[2];
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        l[w] = __builtin_amdgcn_perm((uint32_t)x[2 * w + 1], (uint32_t)x[2 * w], 0x06040200u);
        const uint32_t hb = __builtin_amdgcn_perm((uint32_t)x[2 * w + 1], (uint32_t)x[2 * w], 0x07050301u);
        const uint32_t carry = (l[w] >> 7) & 0x01010101u;
        h[w] = ((hb & 0x7f7f7f7fu) + carry) ^ (hb & 0x80808080u);  // bytewise add, no carries
    }
    lo = (long)(((uint64_t)l[1] << 32) | l[0]);
    hi = (long)(((uint64_t)h[1] << 32) | h[0]);
}

// I[x][c] += sum_{k < K} U[x][k] V[c][k] for x, c < n.  I: row stride ld (>= round4(n+1)
// elements, 16-byte aligned rows).  U, V: int16 rows of stride kst (a multiple of 8) for
// x, c < n; entries K .. round8(K)-1 must be zero; K <= 32 * KCH.  Every wave of the
// workgroup calls this; the caller orders the writes before later reads (__syncthreads).
//
// Each wave takes chunks of kUpdTiles 16-column tiles and computes the transposed tiles
// (V U^T) with v_mfma_i32_16x16x32_i8, whose accumulator layout (lane l: row l % 16,
// columns 4 (l / 16) + 0..3 of the I tile) makes every lane's share of a tile one 16-byte
// load and store.  A lane's operand bytes for 32-entry chunk kc are entries
// 32 kc + 8 (l / 16) + 0..7 of one row of V (column l % 16 of the tile) or of U (row
// l % 16).  Row tiles are loaded DEPTH - 1 tiles ahead of their update; the loads
// are unconditional (lanes past the matrix read its last row / 4 columns and store
// nothing), so no branch keeps them from staying in flight.
//
// Rows [x_lo, x_hi) only (x_lo a multiple of 16): the flush of one matrix split among
// several workgroups.  U and V may be in LDS or in global memory.
template <int NW, int KCH, int DEPTH = kUpdDepth>
__device__ __forceinline__ void mfma_rank_update(int32_t *Ig, int ld, int n, const int16_t *U_s,
                                                 const int16_t *V_s, int kst, int K, int lane,
                                                 int wid, int x_lo = 0, int x_hi = -1) {
    if (x_hi < 0) x_hi = n;
    const int groups = K / 8 + ((K & 7) != 0);
    const int kch = (groups + 3) / 4;
    const int li = lane & 15, kg = lane >> 4;
    const int ntile = (n + 15) / 16;
    const int nchunk = (ntile + kUpdTiles - 1) / kUpdTiles;
    const int clast = (n - 1) & ~3;
    for (int ch = wid; ch < nchunk; ch += NW) {
        const int cb = ch * kUpdTiles * 16;  // first column of the chunk
        long a0[KCH][kUpdTiles], a1[KCH][kUpdTiles];
        bool any_hi = false;
#pragma unroll
        for (int kc = 0; kc < KCH; ++kc)
#pragma unroll
            for (int t = 0; t < kUpdTiles; ++t) {
                const int c = cb + t * 16 + li, g = 4 * kc + kg;
                v4h w = v4h{0, 0, 0, 0};
                if (c < n && g < groups) w = reinterpret_cast<const v4h *>(V_s + (size_t)c * kst)[g];
                split_bytes(w, a0[kc][t], a1[kc][t]);
                any_hi |= a1[kc][t] != 0;
            }
        const bool vh = __ballot(any_hi) != 0;  // some V entry needs its high byte

        v4i cur[DEPTH][kUpdTiles];
        auto load_rows = [&](v4i(&dst)[kUpdTiles], int x0) {
            const int xr = min(x0 + li, n - 1);
#pragma unroll
            for (int t = 0; t < kUpdTiles; ++t) {
                const int cr = min(cb + t * 16 + 4 * kg, clast);
                dst[t] = *reinterpret_cast<const v4i *>(Ig + (int64_t)xr * ld + cr);
            }
        };
        auto update_rows = [&](const v4i(&src)[kUpdTiles], int x0) {
            if (x0 >= x_hi) return;
            const int x = x0 + li;
            long b0[KCH], b1[KCH];
            bool uh[KCH];
#pragma unroll
            for (int kc = 0; kc < KCH; ++kc) {
                const int g = 4 * kc + kg;
                v4h w = v4h{0, 0, 0, 0};
                if (x < x_hi && g < groups) w = reinterpret_cast<const v4h *>(U_s + (size_t)x * kst)[g];
                split_bytes(w, b0[kc], b1[kc]);
                uh[kc] = __ballot(b1[kc] != 0) != 0;  // some U entry needs its high byte
            }
            const v4i z = v4i{0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < kUpdTiles; ++t) {
                const int c = cb + t * 16 + 4 * kg;
                if (cb + t * 16 >= n) break;
                // V U^T = a0 b0 + 2^8 (a0 b1 + a1 b0) + 2^16 a1 b1, summed mod 2^32
                v4i s0 = z, s1 = z, s2 = z;
#pragma unroll
                for (int kc = 0; kc < KCH; ++kc) {
                    if (kc >= kch) break;
                    s0 = __builtin_amdgcn_mfma_i32_16x16x32_i8(a0[kc][t], b0[kc], s0, 0, 0, 0);
                    if (uh[kc]) s1 = __builtin_amdgcn_mfma_i32_16x16x32_i8(a0[kc][t], b1[kc], s1, 0, 0, 0);
                    if (vh) {
                        s1 = __builtin_amdgcn_mfma_i32_16x16x32_i8(a1[kc][t], b0[kc], s1, 0, 0, 0);
                        if (uh[kc]) s2 = __builtin_amdgcn_mfma_i32_16x16x32_i8(a1[kc][t], b1[kc], s2, 0, 0, 0);
                    }
                }
                v4i r;
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    r[v] = (int32_t)((uint32_t)src[t][v] + (uint32_t)s0[v] + ((uint32_t)s1[v] << 8) +
                                     ((uint32_t)s2[v] << 16));
                if (x < x_hi && c < n) *reinterpret_cast<v4i *>(Ig + (int64_t)x * ld + c) = r;
            }
        };
        // scheduling barriers keep the loads in row-tile order, so waiting for the
        // oldest tile does not wait for the younger ones
#pragma unroll
        for (int dd = 0; dd < DEPTH - 1; ++dd) {
            load_rows(cur[dd], x_lo + 16 * dd);
            __builtin_amdgcn_sched_barrier(0);
        }
        for (int x0 = x_lo; x0 < x_hi; x0 += 16 * DEPTH) {
#pragma unroll
            for (int dd = 0; dd < DEPTH; ++dd) {
                load_rows(cur[(dd + DEPTH - 1) % DEPTH], x0 + 16 * (dd + DEPTH - 1));
                __builtin_amdgcn_sched_barrier(0);
                update_rows(cur[dd], x0 + 16 * dd);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
}

}  // namespace qap
