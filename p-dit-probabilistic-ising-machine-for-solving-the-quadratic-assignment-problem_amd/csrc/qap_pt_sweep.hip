// qap_pt_sweep.hip -- the hot path on gfx950: p-dit sweeps and replica exchange.
//
// sweep_kernel<NT, CPL, IM, SYM>: one workgroup of NT threads runs the sweeps of one
// exchange interval for one replica (CPU_PIM.cpp:195-290; GPU_PIM.cu:304-453).
//   * Candidate b of the heat-bath update belongs to thread b % NT (slot q = b / NT,
//     CPL slots per thread), so row reads of the force matrix, flow row reads and the
//     dense rank-2 update are unit-stride across a wavefront.
//   * Max, prefix sums and the winner search are wave64 DPP scans (row_shr 1/2/4/8,
//     row_bcast 15/31) plus one LDS hop per wave when NT > 64.  Prefix sums run in
//     candidate order b = 0..n-1 (per slot block, then across blocks), so the winner
//     is exactly "the first b whose inclusive weight sum exceeds the threshold".
//   * The force matrix I lives in LDS for the whole launch when it fits: int32 with an
//     odd word stride (column gathers I[pinv[b]][a] hit 64 distinct banks), or, when
//     F, D >= 0 and the permutation-independent bound is <= 65535, as uint16 holding
//     -I (n = 256 fits in 137 KiB).  Otherwise it is read and updated in place in HBM.
//   * Distance row i and flow row p(i) of the next update are prefetched into LDS /
//     registers one update ahead; the winning lane commits the permutation swap before
//     the broadcast barrier, so a "stay" or a no-op swap costs three barriers.
//   * Facilities with identical flow rows and columns share a class id; a swap inside
//     one class changes neither the cost nor I, so the rank-2 update is skipped.
//   * The rank-2 update I[k][c] += u1[k] v1[c] + u2[k] v2[c] (CPU_PIM.cpp:263-269) runs
//     over owned columns with flow differences held in registers, four rows in flight,
//     or over the compacted nonzero columns when F is sparse.
// exchange_kernel: best fold, deterministic exchange pass and hot/cold kick of every
// ladder of an instance (CPU_PIM.cpp:293-357), staged through LDS.
#include "qap_pt_device.h"
#include "qap_pt_launch.h"

namespace qap {

constexpr int kWave = 64;

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

// red_s (ints): [0,16) max partials, [16,80) sum partials [q*16 + wave],
// [96] winner b, [97] winner delta, [98] winner's partner location j.
constexpr int kRedWords = 128;
constexpr int kSmallArrays = 15;  // p, pinv, cls, u1, u2, col, v1, v2, rnd, drow x2, dcol x2, ddiag, spare

__host__ __device__ inline int sweep_lds_words(int n, int ld_s, int im) {
    const int small = kSmallArrays * round4(n) + kRedWords;
    if (im == IM_LDS32) return small + n * ld_s;
    if (im == IM_LDS16) return small + round4((n * ld_s + 1) / 2);
    return small;
}

template <int IM>
struct IElem {
    using T = int32_t;
};
template <>
struct IElem<IM_LDS16> {
    using T = uint16_t;  // holds -I
};

template <int NT, int CPL, int IM, bool SYM>
__global__ __launch_bounds__(NT) void sweep_kernel(PoolDev P, const int32_t *__restrict__ rep_list) {
    constexpr int NW = NT / kWave;
    using IT = typename IElem<IM>::T;
    extern __shared__ __align__(16) int32_t smem[];

    const int g = rep_list[blockIdx.x];
    const int inst = P.rep_inst[g];
    if (!P.state[inst].active) return;
    const InstDesc d = P.inst[inst];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int n = d.n, ld = d.ld, ld_s = d.ld_s, n4 = round4(n);
    const int local = g - d.rep_base;

    int32_t *p_s = smem;
    int32_t *pinv_s = p_s + n4;
    int32_t *cls_s = pinv_s + n4;
    int32_t *u1_s = cls_s + n4;
    int32_t *u2_s = u1_s + n4;
    int32_t *col_s = u2_s + n4;
    int32_t *v1_s = col_s + n4;
    int32_t *v2_s = v1_s + n4;
    uint32_t *rnd_s = reinterpret_cast<uint32_t *>(v2_s + n4);
    int32_t *drow_s = reinterpret_cast<int32_t *>(rnd_s + n4);  // 2 buffers
    int32_t *dcol_s = drow_s + 2 * n4;                           // 2 buffers
    int32_t *ddiag_s = dcol_s + 2 * n4;
    int32_t *red_s = ddiag_s + 2 * n4;  // one spare array keeps the 16-byte alignment
    IT *I_s = reinterpret_cast<IT *>(red_s + kRedWords);

    const int32_t *__restrict__ F = P.F + d.mat_off;
    const int32_t *__restrict__ D = P.D + d.mat_off;
    const int32_t *__restrict__ Ft = P.Ft + d.mat_off;
    const int32_t *__restrict__ Dt = P.Dt + d.mat_off;
    const int32_t *__restrict__ Fdg = P.Fdiag + d.vec_off;
    const int32_t *__restrict__ Ddg = P.Ddiag + d.vec_off;
    int32_t *pg = P.perm + d.p_base + (int64_t)local * n;
    int32_t *pinvg = P.pinv + d.p_base + (int64_t)local * n;
    int32_t *bestpg = P.best_perm + d.p_base + (int64_t)local * n;
    int32_t *Ig = P.I + d.I_base + (int64_t)local * n * ld;

    auto Iget = [&](int r, int c) -> int32_t {
        if constexpr (IM == IM_LDS32) return I_s[r * ld_s + c];
        else if constexpr (IM == IM_LDS16) return -(int32_t)I_s[r * ld_s + c];
        else return Ig[(int64_t)r * ld + c];
    };
    auto Iaddr = [&](int r, int c) -> IT * {
        if constexpr (IM == IM_HBM) return Ig + (int64_t)r * ld + c;
        else return I_s + r * ld_s + c;
    };

    // ---- stage the replica into LDS -------------------------------------------
    for (int x = tid; x < n; x += NT) {
        p_s[x] = pg[x];
        pinv_s[x] = pinvg[x];
        cls_s[x] = P.fclass[d.vec_off + x];
        if constexpr (!SYM) ddiag_s[x] = Ddg[x];
    }
    int cw_st = 1;
    while (cw_st < n && cw_st < NT) cw_st <<= 1;
    const int sh_st = __builtin_ctz(cw_st);
    if constexpr (IM != IM_HBM) {
        const int rw = NT >> sh_st;
        for (int k = tid >> sh_st; k < n; k += rw)
            for (int c = tid & (cw_st - 1); c < n; c += cw_st) {
                const int32_t v = Ig[(int64_t)k * ld + c];
                if constexpr (IM == IM_LDS16) I_s[k * ld_s + c] = (uint16_t)(-v);
                else I_s[k * ld_s + c] = v;
            }
    }
    // distance row 0 (and column 0) into buffer 0
    for (int x = tid; x < n; x += NT) {
        drow_s[x] = D[x];
        if constexpr (!SYM) dcol_s[x] = Dt[x];
    }
    __syncthreads();

    int64_t E = P.E[g];
    int64_t bestE = P.bestE[g];
    uint64_t acc = P.accepted[g];
    uint64_t cells = P.cells[g];
    const int pos = P.pos[g];
    const float beta2f = P.beta2f[d.beta_off + pos];
    const double c20 = P.c20[d.beta_off + pos];
    const uint32_t sweep0 = P.state[inst].sweep_base;
    const bool sparse_upd = d.sparse_upd != 0;

    // Per-thread copies of the flow row of the current facility a at owned columns.
    int32_t fr[CPL], ftr[CPL], fdg[CPL];
    int a_cur = p_s[0];
    int32_t faa = SYM ? 0 : Fdg[a_cur];
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
        const int b = tid + q * NT;
        fr[q] = (b < n) ? F[(int64_t)a_cur * n + b] : 0;
        ftr[q] = (!SYM && b < n) ? Ft[(int64_t)a_cur * n + b] : 0;
        fdg[q] = (!SYM && b < n) ? Fdg[b] : 0;
    }
    int buf = 0;

    for (int s = 0; s < d.spe; ++s) {
        const uint32_t sweep = sweep0 + (uint32_t)s;
        __syncthreads();
        for (int x = tid; x < n; x += NT)
            rnd_s[x] = philox4x32_10(u32x4{(uint32_t)x, sweep, (uint32_t)local, 0u}, d.seed_lo,
                                     d.seed_hi)
                           .x;
        __syncthreads();

        for (int i = 0; i < n; ++i) {
            const int a = p_s[i];
            if (a != a_cur) {  // the previous swap moved the facility we prefetched
                a_cur = a;
                if constexpr (!SYM) faa = Fdg[a];
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const int b = tid + q * NT;
                    fr[q] = (b < n) ? F[(int64_t)a * n + b] : 0;
                    if constexpr (!SYM) ftr[q] = (b < n) ? Ft[(int64_t)a * n + b] : 0;
                }
            }
            // prefetch for the next update: distance row i+1, flow row p(i+1)
            const int inext = (i + 1 == n) ? 0 : i + 1;
            const int a_nx = p_s[inext];
            int32_t dn[CPL], dtn[CPL], frn[CPL], ftrn[CPL];
            int32_t faa_n = 0;
            if constexpr (!SYM) faa_n = Fdg[a_nx];
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const int b = tid + q * NT;
                const bool ok = b < n;
                dn[q] = ok ? D[(int64_t)inext * n + b] : 0;
                frn[q] = ok ? F[(int64_t)a_nx * n + b] : 0;
                dtn[q] = (!SYM && ok) ? Dt[(int64_t)inext * n + b] : 0;
                ftrn[q] = (!SYM && ok) ? Ft[(int64_t)a_nx * n + b] : 0;
            }

            const int32_t *drow = drow_s + buf * n4;
            const int32_t *dcol = dcol_s + buf * n4;
            const int32_t Iia = Iget(i, a);
            const int32_t Dii = SYM ? 0 : ddiag_s[i];

            // (1) O(1) deltas of every candidate (CPU_PIM.cpp:220-232)
            int32_t dE[CPL];
            int lmax = 0;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const int b = tid + q * NT;
                dE[q] = INT32_MIN;
                if (b < n) {
                    const int j = pinv_s[b];
                    const int32_t force = Iget(i, b) + Iget(j, a) - Iia - Iget(j, b);
                    int32_t v;
                    if constexpr (SYM) {
                        v = force - 4 * drow[j] * fr[q];
                    } else {
                        const int32_t dist = Dii + ddiag_s[j] - drow[j] - dcol[j];
                        const int32_t flow = faa + fdg[q] - fr[q] - ftr[q];
                        v = force - dist * flow;
                    }
                    dE[q] = v;
                    lmax = max(lmax, v);
                }
            }
            int maxd;
            {
                const int wm = wave_incl_max(lmax);
                if constexpr (NW == 1) {
                    maxd = lane63(wm);
                } else {
                    if (lane == kWave - 1) red_s[wid] = wm;
                    __syncthreads();
                    maxd = red_s[0];
#pragma unroll
                    for (int w = 1; w < NW; ++w) maxd = max(maxd, red_s[w]);
                }
            }

            // (2) integer heat-bath weights and prefix sums in candidate order
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
                __syncthreads();
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

            // (3) winner: first b whose inclusive prefix exceeds floor(u * S / 2^32);
            //     it commits the permutation swap before the broadcast barrier.
            const uint32_t thr = (uint32_t)(((uint64_t)rnd_s[i] * (uint64_t)S) >> 32);
            if (S == 0 && tid == 0) {  // every weight truncated to zero: stay
                red_s[96] = a;
                red_s[97] = 0;
                red_s[98] = i;
            }
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const uint32_t ex = qbase[q] + wbase[q] + incl[q] - w[q];
                if (ex <= thr && thr < ex + w[q]) {
                    const int b = tid + q * NT;
                    const int j = pinv_s[b];
                    red_s[96] = b;
                    red_s[97] = dE[q];
                    red_s[98] = j;
                    if (b != a) {
                        p_s[i] = b;
                        p_s[j] = a;
                        pinv_s[b] = i;
                        pinv_s[a] = j;
                    }
                }
            }
            // next distance row into the other buffer (nobody reads it this update)
            {
                int32_t *dnb = drow_s + (buf ^ 1) * n4;
                int32_t *dtb = dcol_s + (buf ^ 1) * n4;
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const int b = tid + q * NT;
                    if (b < n) {
                        dnb[b] = dn[q];
                        if constexpr (!SYM) dtb[b] = dtn[q];
                    }
                }
            }
            __syncthreads();
            const int b = red_s[96];
            const int32_t dsel = red_s[97];
            const int j = red_s[98];

            if (b != a) {
                // (4) commit (CPU_PIM.cpp:252-283)
                E -= dsel;
                ++acc;
                if (cls_s[a] != cls_s[b]) {
                    int32_t v1[CPL], v2[CPL];
                    uint32_t nz = 0;
#pragma unroll
                    for (int q = 0; q < CPL; ++q) {
                        const int c = tid + q * NT;
                        v1[q] = 0;
                        v2[q] = 0;
                        if (c < n) {
                            const int32_t fb = F[(int64_t)b * n + c];
                            const int32_t dj = D[(int64_t)j * n + c];
                            if constexpr (SYM) {
                                v1[q] = 2 * (fb - fr[q]);           // 2 (F[b][c] - F[a][c])
                                u1_s[c] = dj - drow[c];              // D[j][k] - D[i][k]
                            } else {
                                const int32_t ftb = Ft[(int64_t)b * n + c];
                                const int32_t dtj = Dt[(int64_t)j * n + c];
                                v1[q] = ftb - ftr[q];                // F[c][b] - F[c][a]
                                v2[q] = fb - fr[q];                  // F[b][c] - F[a][c]
                                u1_s[c] = dtj - dcol[c];             // D[k][j] - D[k][i]
                                u2_s[c] = dj - drow[c];              // D[j][k] - D[i][k]
                            }
                            nz += (v1[q] | v2[q]) != 0;
                        }
                    }
                    uint32_t nnz = (uint32_t)n;
                    if (sparse_upd) {
                        const uint32_t inc = wave_incl_sum(nz);
                        uint32_t base;
                        if constexpr (NW == 1) {
                            base = inc - nz;
                            nnz = (uint32_t)lane63((int)inc);
                        } else {
                            if (lane == kWave - 1) red_s[16 + wid] = (int)inc;
                            __syncthreads();
                            uint32_t before = 0, tot = 0;
#pragma unroll
                            for (int ww = 0; ww < NW; ++ww) {
                                const uint32_t x = (uint32_t)red_s[16 + ww];
                                before += (ww < wid) ? x : 0u;
                                tot += x;
                            }
                            base = before + inc - nz;
                            nnz = tot;
                        }
#pragma unroll
                        for (int q = 0; q < CPL; ++q) {
                            if ((v1[q] | v2[q]) != 0) {
                                col_s[base] = tid + q * NT;
                                v1_s[base] = v1[q];
                                v2_s[base] = v2[q];
                                ++base;
                            }
                        }
                    }
                    __syncthreads();
                    cells += (uint64_t)n * nnz;
                    if (!sparse_upd) {
                        // owned columns, four rows in flight
                        for (int k0 = 0; k0 < n; k0 += 4) {
                            int32_t x[4][CPL], uk1[4], uk2[4];
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int k = k0 + r;
                                uk1[r] = (k < n) ? u1_s[k] : 0;
                                uk2[r] = (!SYM && k < n) ? u2_s[k] : 0;
#pragma unroll
                                for (int q = 0; q < CPL; ++q) {
                                    const int c = tid + q * NT;
                                    x[r][q] = (k < n && c < n) ? *Iaddr(k, c) : 0;
                                }
                            }
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int k = k0 + r;
#pragma unroll
                                for (int q = 0; q < CPL; ++q) {
                                    const int c = tid + q * NT;
                                    if (k < n && c < n) {
                                        const int32_t dv = SYM ? uk1[r] * v1[q]
                                                               : uk1[r] * v1[q] + uk2[r] * v2[q];
                                        if constexpr (IM == IM_LDS16)
                                            *Iaddr(k, c) = (uint16_t)(x[r][q] - dv);
                                        else
                                            *Iaddr(k, c) = x[r][q] + dv;
                                    }
                                }
                            }
                        }
                    } else if (nnz > 0) {
                        int cw = 1;
                        while (cw < (int)nnz && cw < NT) cw <<= 1;
                        const int shw = __builtin_ctz(cw);
                        const int rw = NT >> shw;
                        const int cx = tid & (cw - 1);
                        const int ry = tid >> shw;
                        for (int ci = cx; ci < (int)nnz; ci += cw) {
                            const int c = col_s[ci];
                            const int32_t a1 = v1_s[ci];
                            const int32_t a2 = v2_s[ci];
                            for (int k0 = ry; k0 < n; k0 += 4 * rw) {
                                int32_t x[4];
#pragma unroll
                                for (int r = 0; r < 4; ++r) {
                                    const int k = k0 + r * rw;
                                    x[r] = (k < n) ? *Iaddr(k, c) : 0;
                                }
#pragma unroll
                                for (int r = 0; r < 4; ++r) {
                                    const int k = k0 + r * rw;
                                    if (k < n) {
                                        const int32_t dv = SYM ? u1_s[k] * a1
                                                               : u1_s[k] * a1 + u2_s[k] * a2;
                                        if constexpr (IM == IM_LDS16)
                                            *Iaddr(k, c) = (uint16_t)(x[r] - dv);
                                        else
                                            *Iaddr(k, c) = x[r] + dv;
                                    }
                                }
                            }
                        }
                    }
                    __syncthreads();
                }
                if (E < bestE) {  // CPU_PIM.cpp:271-272, plus the permutation
                    bestE = E;
                    for (int x = tid; x < n; x += NT) bestpg[x] = p_s[x];
                }
            }
            // rotate the prefetched rows in
            buf ^= 1;
            a_cur = a_nx;
            if constexpr (!SYM) faa = faa_n;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                fr[q] = frn[q];
                if constexpr (!SYM) ftr[q] = ftrn[q];
            }
        }
    }

    __syncthreads();
    for (int x = tid; x < n; x += NT) {
        pg[x] = p_s[x];
        pinvg[x] = pinv_s[x];
    }
    if constexpr (IM != IM_HBM) {
        const int rw = NT >> sh_st;
        for (int k = tid >> sh_st; k < n; k += rw)
            for (int c = tid & (cw_st - 1); c < n; c += cw_st) {
                if constexpr (IM == IM_LDS16) Ig[(int64_t)k * ld + c] = -(int32_t)I_s[k * ld_s + c];
                else Ig[(int64_t)k * ld + c] = I_s[k * ld_s + c];
            }
    }
    if (tid == 0) {
        P.E[g] = E;
        P.bestE[g] = bestE;
        P.accepted[g] = acc;
        P.cells[g] = cells;
    }
}

// ---------------------------------------------------------------------------
// Exchange: one workgroup per instance.  Energies, bests, ladder order and positions
// are staged into LDS with coalesced loads; one thread per ladder runs the sequential
// fold / bubble pass / kick (CPU_PIM.cpp:295-344) on LDS; the goal test
// (:346-348) uses the best over all ladders of the instance.
// ---------------------------------------------------------------------------
constexpr int kExMaxRT = 2048;

__global__ __launch_bounds__(64) void exchange_kernel(PoolDev P) {
    __shared__ int64_t E_s[kExMaxRT];
    __shared__ int64_t bE_s[kExMaxRT];
    __shared__ int32_t ord_s[kExMaxRT];
    __shared__ int32_t pos_s[kExMaxRT];
    __shared__ int64_t gbest_s[64];
    __shared__ uint32_t nex_s[64], nkick_s[64];
    __shared__ int go_s;
    const int inst = blockIdx.x;
    const int tid = threadIdx.x;
    InstState &st = P.state[inst];
    const InstDesc d = P.inst[inst];
    if (tid == 0) {
        go_s = !(st.halted || st.megas_done >= st.mega_limit);
        if (!go_s) st.active = 0;
    }
    __syncthreads();
    if (!go_s) return;
    const int R = d.R, T = d.T, RT = R * T;
    const bool staged = RT <= kExMaxRT;
    int64_t *E = staged ? E_s : P.E + d.rep_base;
    int64_t *bE = staged ? bE_s : P.bestE + d.rep_base;
    int32_t *ord = staged ? ord_s : P.order + d.rep_base;
    int32_t *pos = staged ? pos_s : P.pos + d.rep_base;
    if (staged) {
        for (int x = tid; x < RT; x += 64) {
            E_s[x] = P.E[d.rep_base + x];
            bE_s[x] = P.bestE[d.rep_base + x];
            ord_s[x] = P.order[d.rep_base + x];
            pos_s[x] = P.pos[d.rep_base + x];
        }
        __syncthreads();
    }
    int64_t gbest = INT64_MAX;
    uint32_t nex = 0, nkick = 0;
    for (int t = tid; t < T; t += 64) {
        const int L = d.ladder_base + t;
        const int base = t * R;
        int32_t tss = P.tss[L] + 1;
        int64_t lbest = P.ladder_best[L];
        for (int l = 0; l < R; ++l) {
            const int64_t b = bE[base + l];
            if (b < lbest) {
                lbest = b;
                tss = 0;
            }
        }
        for (int k = 0; k < R - 1; ++k) {
            const int r1 = ord[base + k], r2 = ord[base + k + 1];
            if (E[base + r1] < E[base + r2]) {
                ord[base + k] = r2;
                ord[base + k + 1] = r1;
                pos[base + r1] = k + 1;
                pos[base + r2] = k;
                ++nex;
                if (r2 == R - 1) tss = 0;
            }
        }
        if (tss > d.kick_after) {
            const int r1 = ord[base + R - 1], r2 = ord[base];
            ord[base + R - 1] = r2;
            ord[base] = r1;
            pos[base + r1] = 0;
            pos[base + r2] = R - 1;
            ++nkick;
        }
        P.tss[L] = tss;
        P.ladder_best[L] = lbest;
        gbest = lbest < gbest ? lbest : gbest;
    }
    gbest_s[tid] = gbest;
    nex_s[tid] = nex;
    nkick_s[tid] = nkick;
    __syncthreads();
    if (staged) {
        for (int x = tid; x < RT; x += 64) {
            P.order[d.rep_base + x] = ord_s[x];
            P.pos[d.rep_base + x] = pos_s[x];
        }
    }
    if (tid == 0) {
        int64_t gb = st.global_best;
        uint64_t ne = 0, nk = 0;
        for (int x = 0; x < 64; ++x) {
            gb = gbest_s[x] < gb ? gbest_s[x] : gb;
            ne += nex_s[x];
            nk += nkick_s[x];
        }
        st.global_best = gb;
        st.exchanges += ne;
        st.kicks += nk;
        if (d.has_goal && gb <= d.goal) {
            st.halted = 1;
            st.active = 0;
        } else {
            st.sweep_base = (uint32_t)(st.megas_done * d.spe);
            st.megas_done += 1;
            st.active = 1;
        }
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int NT, int CPL, int IM, bool SYM>
static hipError_t launch_sweep_t(const PoolDev &P, const int32_t *list, int nblocks, int lds_bytes,
                                 hipStream_t s) {
    auto k = sweep_kernel<NT, CPL, IM, SYM>;
    static bool attr_done = false;
    if (!attr_done) {
        (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  kMaxLdsBytes);
        attr_done = true;
    }
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(NT), lds_bytes, s, P, list);
    return hipGetLastError();
}

template <int NT, int CPL>
static hipError_t launch_sweep_c(const PoolDev &P, const SweepShape &sh, const int32_t *list,
                                 int nblocks, int lds_bytes, hipStream_t s) {
    switch (sh.im * 2 + (sh.sym ? 1 : 0)) {
    case IM_LDS32 * 2 + 1: return launch_sweep_t<NT, CPL, IM_LDS32, true>(P, list, nblocks, lds_bytes, s);
    case IM_LDS32 * 2 + 0: return launch_sweep_t<NT, CPL, IM_LDS32, false>(P, list, nblocks, lds_bytes, s);
    case IM_LDS16 * 2 + 1: return launch_sweep_t<NT, CPL, IM_LDS16, true>(P, list, nblocks, lds_bytes, s);
    case IM_LDS16 * 2 + 0: return launch_sweep_t<NT, CPL, IM_LDS16, false>(P, list, nblocks, lds_bytes, s);
    case IM_HBM * 2 + 1: return launch_sweep_t<NT, CPL, IM_HBM, true>(P, list, nblocks, lds_bytes, s);
    case IM_HBM * 2 + 0: return launch_sweep_t<NT, CPL, IM_HBM, false>(P, list, nblocks, lds_bytes, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_sweep(const PoolDev &P, const SweepShape &sh, const int32_t *list, int nblocks,
                        int lds_bytes, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    switch (sh.nt * 8 + sh.cpl) {
    case 64 * 8 + 1: return launch_sweep_c<64, 1>(P, sh, list, nblocks, lds_bytes, s);
    case 64 * 8 + 2: return launch_sweep_c<64, 2>(P, sh, list, nblocks, lds_bytes, s);
    case 64 * 8 + 4: return launch_sweep_c<64, 4>(P, sh, list, nblocks, lds_bytes, s);
    case 128 * 8 + 1: return launch_sweep_c<128, 1>(P, sh, list, nblocks, lds_bytes, s);
    case 128 * 8 + 2: return launch_sweep_c<128, 2>(P, sh, list, nblocks, lds_bytes, s);
    case 256 * 8 + 1: return launch_sweep_c<256, 1>(P, sh, list, nblocks, lds_bytes, s);
    case 256 * 8 + 2: return launch_sweep_c<256, 2>(P, sh, list, nblocks, lds_bytes, s);
    case 256 * 8 + 4: return launch_sweep_c<256, 4>(P, sh, list, nblocks, lds_bytes, s);
    default: return hipErrorInvalidValue;
    }
}

bool sweep_shape_supported(int nt, int cpl) {
    switch (nt * 8 + cpl) {
    case 64 * 8 + 1: case 64 * 8 + 2: case 64 * 8 + 4: case 128 * 8 + 1: case 128 * 8 + 2:
    case 256 * 8 + 1: case 256 * 8 + 2: case 256 * 8 + 4: return true;
    default: return false;
    }
}

int sweep_lds_bytes(int n, int ld_s, int im) { return 4 * sweep_lds_words(n, ld_s, im); }

hipError_t launch_exchange(const PoolDev &P, int count, hipStream_t s) {
    hipLaunchKernelGGL(exchange_kernel, dim3(count), dim3(64), 0, s, P);
    return hipGetLastError();
}

}  // namespace qap
