// qap_pt_sweep.hip -- the hot path on gfx950: p-dit sweeps and replica exchange.
//
// sweep_kernel<NT, CPL, IM, SYM, SP>: one workgroup of NT threads runs the sweeps of one
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
//     one class changes neither the cost nor I, so the rank-2 update is skipped.  On
//     symmetric instances, twins (flow rows equal outside their own two columns, e.g.
//     two facilities of one dense block) change two columns of I: with I as uint16 in
//     LDS or in HBM every thread updates its rows of those columns straight away, with
//     no scan for nonzero columns.
//   * The rank-2 update I[k][c] += u1[k] v1[c] + u2[k] v2[c] (CPU_PIM.cpp:263-269) runs
//     over the static range of nonzero flow columns with 128-bit accesses (LDS
//     placements), over owned columns with the flow differences in registers (HBM), or,
//     when F is sparse, over the exact nonzero columns: per-wave ballot masks published
//     with one barrier, one or two columns walked straight from the masks, more
//     compacted into a list.
// exchange_kernel: best fold, deterministic exchange pass and hot/cold kick of every
// ladder of an instance (CPU_PIM.cpp:293-357), staged through LDS.
#include "qap_pt_device.h"
#include "qap_pt_launch.h"
#include "qap_pt_sweep_common.h"
#include "qap_pt_xmirror.h"

namespace qap {

// red_s (ints): [0,16) max partials, [16,80) sum partials [q*16 + wave] (in the commit of
// a sparse update, [16, 16 + 2 NW CPL) holds the uint64 nonzero-column masks instead),
// [96] winner b, [97] winner delta, [98] winner's partner location j, [100, 103) the
// exchange step's outcome in a persistent launch, [140, 140 + 2 NW) per-wave rewritten-cell counts (uint64) at the end of a launch.
constexpr int kRedWords = 160;
// sparse flow updates with at most this many nonzero columns walk the columns directly
// (one LDS round trip per column, in sequence); longer sets batch rows over a compacted
// list, which is faster from three columns on (n = 100: 12.3 vs 10.5 M attempts/s at 8)
constexpr int kWalkCols = 2;
// p, pinv, cls, u1, u2, col, v1, v2, rnd, ddiag, drow x2, dcol x2, frow x2, ftrow x2,
// vd1, vd2, fspan -- each round8(n) words; the uint16 placement adds the launch's best
// permutation
constexpr int kSmallArrays = 21;

__host__ __device__ inline int sweep_lds_words(int n, int ld_s, int im) {
    const int small = (kSmallArrays + (im == IM_LDS16 ? 1 : 0)) * round8(n) + kRedWords;
    if (im == IM_LDS32) return small + n * ld_s;
    if (im == IM_LDS16) return small + round4((n * ld_s + 1) / 2);
    return small;
}

// Product of a distance difference and a flow difference in the rank-2 update.  With
// |u| < 2^23 and |v| < 2^23 (checked on the host) and |u v| < 2^31 (the delta bound),
// the 24-bit multiplier gives the exact product at full rate.
template <bool M24>
__device__ __forceinline__ int32_t umul(int32_t u, int32_t v) {
    if constexpr (M24) return __mul24(u, v);
    else return u * v;
}

// Dense rank-2 update I[k][c] += u1[k] v1[c] (+ u2[k] v2[c]) over the columns a thread
// owns (colp[q] = column base, lanes without a column point at the padding column n,
// where v = 0).  Rows go in groups of G with the next group's force entries and row
// factors (four rows per 16-byte LDS read) in flight while the current group is written.
template <int CPL, int IM, bool SYM, bool M24, typename IT>
__device__ __forceinline__ void dense_update(IT *const colp[CPL], int32_t rs, int n,
                                             const int32_t *u1_s, const int32_t *u2_s,
                                             const int32_t v1[CPL], const int32_t v2[CPL]) {
    constexpr int G = (CPL == 1) ? 16 : (CPL == 2 ? 8 : 4);
    int32_t xc[G][CPL], xn[G][CPL], uc1[G], un1[G], uc2[G], un2[G];
    auto load_group = [&](int k0) {
#pragma unroll
        for (int r4 = 0; r4 < G; r4 += 4) {
            const int4 a = *reinterpret_cast<const int4 *>(u1_s + k0 + r4);
            un1[r4] = a.x; un1[r4 + 1] = a.y; un1[r4 + 2] = a.z; un1[r4 + 3] = a.w;
            if constexpr (!SYM) {
                const int4 b = *reinterpret_cast<const int4 *>(u2_s + k0 + r4);
                un2[r4] = b.x; un2[r4 + 1] = b.y; un2[r4 + 2] = b.z; un2[r4 + 3] = b.w;
            }
        }
#pragma unroll
        for (int r = 0; r < G; ++r) {
            const int k = min(k0 + r, n - 1);
#pragma unroll
            for (int q = 0; q < CPL; ++q) xn[r][q] = (int32_t)colp[q][(int64_t)k * rs];
        }
    };
    load_group(0);
    for (int k0 = 0; k0 < n; k0 += G) {
#pragma unroll
        for (int r = 0; r < G; ++r) {
            uc1[r] = un1[r];
            if constexpr (!SYM) uc2[r] = un2[r];
#pragma unroll
            for (int q = 0; q < CPL; ++q) xc[r][q] = xn[r][q];
        }
        if (k0 + G < n) load_group(k0 + G);
#pragma unroll
        for (int r = 0; r < G; ++r) {
            const int k = k0 + r;
            if (k < n) {  // uniform
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const int32_t dv = SYM ? umul<M24>(uc1[r], v1[q])
                                           : umul<M24>(uc1[r], v1[q]) + umul<M24>(uc2[r], v2[q]);
                    if constexpr (IM == IM_LDS16) colp[q][(int64_t)k * rs] = (uint16_t)(xc[r][q] - dv);
                    else colp[q][(int64_t)k * rs] = xc[r][q] + dv;
                }
            }
        }
    }
}

// Rank-2 update over the column range of vectors [V0, V1) of an LDS force matrix with
// 16-byte rows: W = 4 int32 or W = 8 uint16 (holding -I) columns per 128-bit access.
// Each lane owns one vector column and a fixed residue of rows.
// vd1/vd2 hold v for every column (zero outside the nonzero set).
template <int NT, int IM, bool SYM, bool M24>
__device__ __forceinline__ void vec_update(int4 *I4, int ldv, int n, int V0, int V1,
                                           const int32_t *u1_s, const int32_t *u2_s,
                                           const int32_t *vd1, const int32_t *vd2, int tid) {
    constexpr int W = (IM == IM_LDS16) ? 8 : 4;
    const int Vn = V1 - V0;
    if (Vn <= 0) return;
    const int rpp = NT / Vn;  // >= 1: Vn <= ceil(n / W) <= NT
    const int cx = tid % Vn, ry = tid / Vn;
    if (ry >= rpp) return;
    const int vi = V0 + cx;
    int32_t a1[W], a2[W];
#pragma unroll
    for (int t4 = 0; t4 < W; t4 += 4) {
        const int4 x = *reinterpret_cast<const int4 *>(vd1 + vi * W + t4);
        a1[t4] = x.x; a1[t4 + 1] = x.y; a1[t4 + 2] = x.z; a1[t4 + 3] = x.w;
        if constexpr (!SYM) {
            const int4 y = *reinterpret_cast<const int4 *>(vd2 + vi * W + t4);
            a2[t4] = y.x; a2[t4 + 1] = y.y; a2[t4 + 2] = y.z; a2[t4 + 3] = y.w;
        } else {
            a2[t4] = a2[t4 + 1] = a2[t4 + 2] = a2[t4 + 3] = 0;
        }
    }
    // Rows go in groups of G loads in flight, with no bounds test inside a full group;
    // the last, partial group clamps its loads and predicates its stores.
    constexpr int G = 8;
    const int step = rpp * ldv;  // int4 elements between consecutive rows of this lane
    // uint16 placement: the word of columns (2h, 2h+1) gets -u v for both halves at once
    // with packed 16-bit integer multiply-adds (modulo 2^16, exact: the entries fit)
    typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
    u16x2 pk1[4], pk2[4];
    if constexpr (IM == IM_LDS16) {
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            pk1[h] = u16x2{(unsigned short)a1[2 * h], (unsigned short)a1[2 * h + 1]};
            pk2[h] = u16x2{(unsigned short)a2[2 * h], (unsigned short)a2[2 * h + 1]};
        }
    }
    auto update_row = [&](int4 x, int32_t w1, int32_t w2) -> int4 {
        int32_t e[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            if constexpr (IM == IM_LDS16) {
                // two uint16 (-I) per word: subtract the increments, wrap to 16 bits
                const unsigned short n1 = (unsigned short)(-w1), n2 = (unsigned short)(-w2);
                u16x2 r = __builtin_bit_cast(u16x2, e[h]);
                r = u16x2{n1, n1} * pk1[h] + r;
                if constexpr (!SYM) r = u16x2{n2, n2} * pk2[h] + r;
                e[h] = __builtin_bit_cast(int32_t, r);
            } else {
                int32_t dv = umul<M24>(w1, a1[h]);
                if constexpr (!SYM) dv += umul<M24>(w2, a2[h]);
                e[h] += dv;
            }
        }
        return int4{e[0], e[1], e[2], e[3]};
    };
    int4 *rowp = I4 + ry * ldv + vi;
    int k0 = ry;
    for (; k0 + (G - 1) * rpp < n; k0 += G * rpp, rowp += G * step) {
        int4 x[G];
        int32_t w1[G], w2[G];
#pragma unroll
        for (int r = 0; r < G; ++r) {
            x[r] = rowp[r * step];
            w1[r] = u1_s[k0 + r * rpp];
            w2[r] = SYM ? 0 : u2_s[k0 + r * rpp];
        }
#pragma unroll
        for (int r = 0; r < G; ++r) rowp[r * step] = update_row(x[r], w1[r], w2[r]);
    }
    if (k0 < n) {  // fewer than G rows left for this lane
        int4 x[G];
        int32_t w1[G], w2[G];
#pragma unroll
        for (int r = 0; r < G; ++r) {
            const int k = min(k0 + r * rpp, n - 1);
            x[r] = I4[k * ldv + vi];
            w1[r] = u1_s[k];
            w2[r] = SYM ? 0 : u2_s[k];
        }
#pragma unroll
        for (int r = 0; r < G; ++r)
            if (k0 + r * rpp < n) rowp[r * step] = update_row(x[r], w1[r], w2[r]);
    }
}

template <int IM>
struct IElem {
    using T = int32_t;
};
template <>
struct IElem<IM_LDS16> {
    using T = uint16_t;  // holds -I
};

__device__ void exchange_instance(const PoolDev &P, int inst, int cap, unsigned char *ex_smem,
                                  int tid);

// The exchange step between two intervals of a persistent launch, called by every thread of
// replica g's workgroup (local index `local`, cost E, best bestE).  Returns kXmRun with the
// replica's new ladder position and first sweep in *pos and *sweep0, kXmStop when the
// instance ends at this step (limit or goal), kXmFault when a wait gave up (the fault word
// is set).  With `mirror` every workgroup runs the step itself on its copy xm (wave 0);
// otherwise the last replica to arrive at the instance's counter runs exchange_instance and
// advances the generation word to `target`, for which the others wait.
__device__ int persistent_exchange(const PoolDev &P, const InstDesc &d, int inst, int g, int local,
                                   int64_t E, int64_t bestE, uint32_t target, bool mirror,
                                   XMirror &xm, int32_t *red_s, int tid, int *pos,
                                   uint32_t *sweep0) {
    const int lane = tid & (kWave - 1), wid = tid / kWave;
    if (mirror) {
        if (wid == 0) {
            int np = 0;
            const int r = xm_exchange(P, d, xm, g, local, E, bestE, lane, &np);
            if (lane == 0) {
                red_s[100] = r;
                red_s[101] = np;
                red_s[102] = (int)xm.sweep_base;
            }
        }
        __syncthreads();
        const int r = red_s[100];
        if (r == kXmRun) {
            *pos = red_s[101];
            *sweep0 = (uint32_t)red_s[102];
        }
        return r;
    }
    if (tid == 0) {  // this replica's cost and best, for the exchange step
        P.E[g] = E;
        P.bestE[g] = bestE;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t prev = __hip_atomic_fetch_add(&P.sync[1 + 2 * inst], 1u, __ATOMIC_ACQ_REL,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        red_s[100] = prev == (uint32_t)(d.R * d.T - 1);  // the instance's last arrival
        red_s[101] = 0;
    }
    __syncthreads();
    if (red_s[100]) {
        if (wid == 0) {  // every other replica has arrived: run the step, release them
            if (lane == 0)
                __hip_atomic_store(&P.sync[1 + 2 * inst], 0u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            exchange_instance(P, inst, 0, nullptr, lane);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            if (lane == 0)
                __hip_atomic_store(&P.sync[2 + 2 * inst], target, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
    } else if (tid == 0) {
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        // relaxed polls (an acquire per poll would invalidate the caches other
        // workgroups read from); the acquire fence after the barrier below follows
        while (__hip_atomic_load(&P.sync[2 + 2 * inst], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT) != target) {
            __builtin_amdgcn_s_sleep(2);
            if (__builtin_amdgcn_s_memrealtime() - t0 > kPersistWaitTicks) {
                __hip_atomic_store(&P.sync[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                red_s[101] = 1;
                break;
            }
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (red_s[101]) return kXmFault;
    const InstState stn = P.state[inst];
    if (!stn.active) return kXmStop;  // limit or goal reached at this step
    *pos = P.pos[g];
    *sweep0 = stn.sweep_base;
    return kXmRun;
}

// max_iv > 1 (persistent launch, LDS placements): the workgroup runs up to max_iv exchange
// intervals with its replica resident in LDS.  Between intervals the instance's exchange
// step (the one exchange_kernel runs between launches) runs inside the launch: mirrored by
// every workgroup from the replicas' tagged costs (persistent_exchange; instances with
// InstDesc::xmirror), or run by the last replica to arrive at the instance's counter, which
// releases the others by advancing the instance's generation word.
template <int NT, int CPL, int IM, bool SYM, bool SP>
__global__ __launch_bounds__(NT) void sweep_kernel(PoolDev P, const int32_t *__restrict__ rep_list,
                                                   int max_iv) {
    constexpr int NW = NT / kWave;
    using IT = typename IElem<IM>::T;
    extern __shared__ __align__(16) int32_t smem[];

    const int g = rep_list[blockIdx.x];
    const int inst = P.rep_inst[g];
    if (!P.state[inst].active) return;
    const InstDesc d = P.inst[inst];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int n = d.n, ld = d.ld, ld_s = d.ld_s, n4 = round8(n);  // array stride
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
    int32_t *frow_s = ddiag_s + n4;                              // 2 buffers
    int32_t *ftrow_s = frow_s + 2 * n4;                          // 2 buffers
    int32_t *vd1_s = ftrow_s + 2 * n4;                           // v of every column
    int32_t *vd2_s = vd1_s + n4;
    int32_t *fsp_s = vd2_s + n4;                                 // nonzero flow columns
    int32_t *bestp_s = fsp_s + n4;  // uint16 placement: this launch's best permutation
    int32_t *red_s = bestp_s + (IM == IM_LDS16 ? n4 : 0);
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
        fsp_s[x] = P.fspan[d.vec_off + x];
        if constexpr (!SYM) ddiag_s[x] = Ddg[x];
    }
    // HBM rows are 16-byte aligned (ld % 4 == 0): the matrix is n * L4 quads, walked flat
    // by the whole workgroup (quad g = tid + m NT at row g / L4, column quad g % L4) with
    // kStageU loads in flight per thread.
    const int L4 = ld >> 2, Q = n * L4;
    const int dq = NT / L4, dr = NT - dq * L4;  // advance of (row, quad) per NT quads
    const int row_t = tid / L4, cx_t = tid - row_t * L4;
    auto advance = [&](int &row, int &cx) {
        cx += dr;
        row += dq;
        if (cx >= L4) {
            cx -= L4;
            ++row;
        }
    };
    constexpr int kStageU = 16;
    if constexpr (IM == IM_LDS32 || IM == IM_LDS16) {
        // (native vector types: HIP's int4 struct copies are not promoted to registers)
        const v4i *Ig4 = reinterpret_cast<const v4i *>(Ig);
        const int ldq = ld_s >> 2;  // LDS row stride in 16-byte (int32) / 8-byte (uint16) units
        int row = row_t, cx = cx_t;
        for (int g0 = tid; g0 < Q; g0 += kStageU * NT) {
            v4i v[kStageU];
            int ro[kStageU], co[kStageU];
#pragma unroll
            for (int u = 0; u < kStageU; ++u) {
                ro[u] = row;
                co[u] = cx;
                v[u] = Ig4[(int64_t)min(row, n - 1) * L4 + cx];
                advance(row, cx);
            }
#pragma unroll
            for (int u = 0; u < kStageU; ++u) {
                if (ro[u] >= n) continue;
                if constexpr (IM == IM_LDS32) {
                    reinterpret_cast<v4i *>(I_s)[ro[u] * ldq + co[u]] = v[u];
                } else {  // 4 int32 -> 4 uint16 holding -I, one 8-byte store
                    const uint32_t w0 = ((uint32_t)(-v[u].x) & 0xffffu) | ((uint32_t)(-v[u].y) << 16);
                    const uint32_t w1 = ((uint32_t)(-v[u].z) & 0xffffu) | ((uint32_t)(-v[u].w) << 16);
                    reinterpret_cast<uint2 *>(I_s)[ro[u] * ldq + co[u]] = uint2{w0, w1};
                }
            }
        }
    }
    // distance row 0 (and column 0) and the flow row of p(0) into buffer 0
    lds_sync<NW>();
    int a_cur = p_s[0];
    for (int x = tid; x < n; x += NT) {
        drow_s[x] = D[x];
        frow_s[x] = F[(int64_t)a_cur * n + x];
        if constexpr (!SYM) {
            dcol_s[x] = Dt[x];
            ftrow_s[x] = Ft[(int64_t)a_cur * n + x];
        }
    }
    lds_sync<NW>();

    int64_t E = P.E[g];
    int64_t bestE = P.bestE[g];
    bool improved = false;  // bestp_s holds this launch's best permutation
    uint64_t acc = P.accepted[g];
    uint64_t cells = 0;  // this wave's share: n x the nonzero flow differences it owns
    int pos = P.pos[g];
    float beta2f = P.beta2f[d.beta_off + pos];
    double c20 = P.c20[d.beta_off + pos];
    uint32_t sweep0 = P.state[inst].sweep_base;
    // persistent launch: the instance's exchange state mirrored in wave 0, or the exchange
    // steps of this instance done inside persistent launches so far (read before this
    // workgroup first arrives, so no step of this launch can have advanced it)
    const bool mirror = max_iv > 1 && d.xmirror && P.xslot;
    XMirror xm{};
    int xstat = kXmRun;
    if (mirror && wid == 0) xm = xm_load(P, inst, d, lane);
    uint32_t gen0 = 0;
    if (max_iv > 1 && !mirror)
        gen0 = __hip_atomic_load(&P.sync[2 + 2 * inst], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // SP: the launch holds sparse-flow instances only, and the dense update path is not
    // built (grey n = 256: 2.2 % faster; the sparse path is exact for any flow)
    const bool sparse_upd = SP || d.sparse_upd != 0;

    // F[b][b] of the owned candidates (constant for the launch) and their locations
    // pinv[b] (kept in registers, updated on every swap)
    int32_t fdg[CPL], pv[CPL];
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
        const int b = tid + q * NT;
        fdg[q] = (!SYM && b < n) ? Fdg[b] : 0;
        pv[q] = (b < n) ? pinv_s[b] : 0;
    }
    int buf = 0;
    // facility at the current site, carried in a register from update to update (the
    // swap at site i moves it only when the partner location is the next site)
    int a_reg = a_cur;
    // Rows are loaded two updates ahead of their use: during the update of site i the
    // registers below receive distance row i+2 and the flow row of p(i+2); they are stored
    // into the other LDS buffer during the update of site i+1, so an L2 round trip has a
    // whole update to land instead of the part after the loads' issue.  a_rec is the
    // facility whose flow row is in the registers (p(i+2) when issued: a swap in between
    // that moves it is caught by the a != a_cur test at the top of the update).
    int32_t dn[CPL], dtn[CPL], frn[CPL], ftrn[CPL];
    int a_rec;
    {
        const int i1 = n > 1 ? 1 : 0;
        a_rec = p_s[i1];
#pragma unroll
        for (int q = 0; q < CPL; ++q) {
            const int bc = min(tid + q * NT, n - 1);
            dn[q] = D[(int64_t)i1 * n + bc];
            frn[q] = F[(int64_t)a_rec * n + bc];
            if constexpr (!SYM) {
                dtn[q] = Dt[(int64_t)i1 * n + bc];
                ftrn[q] = Ft[(int64_t)a_rec * n + bc];
            } else {
                dtn[q] = 0;
                ftrn[q] = 0;
            }
        }
    }

    for (int iv = 0; iv < max_iv; ++iv) {
    if (iv > 0) {  // ---- exchange step between intervals (persistent launch) --------
        xstat = persistent_exchange(P, d, inst, g, local, E, bestE, gen0 + (uint32_t)iv, mirror,
                                    xm, red_s, tid, &pos, &sweep0);
        if (xstat != kXmRun) break;  // the instance ends here, or a wait gave up
        // (the temperature tables are not kept in registers here as in sweep_spec_kernel:
        // the three extra registers cost grey n = 256 4 %)
        beta2f = P.beta2f[d.beta_off + pos];
        c20 = P.c20[d.beta_off + pos];
    }
    for (int s = 0; s < d.spe; ++s) {
        const uint32_t sweep = sweep0 + (uint32_t)s;
        lds_sync<NW>();
        for (int x = tid; x < n; x += NT)
            rnd_s[x] = philox4x32_10(u32x4{(uint32_t)x, sweep, (uint32_t)local, 0u}, d.seed_lo,
                                     d.seed_hi)
                           .x;
        lds_sync<NW>();

        for (int i = 0; i < n; ++i) {
            const int a = a_reg;
            const uint32_t ri = rnd_s[i];
            if (a != a_cur) {  // the previous swap moved the facility we prefetched
                a_cur = a;
                for (int x = tid; x < n; x += NT) {
                    frow_s[buf * n4 + x] = F[(int64_t)a * n + x];
                    if constexpr (!SYM) ftrow_s[buf * n4 + x] = Ft[(int64_t)a * n + x];
                }
                lds_sync<NW>();
            }
            const int32_t *drow = drow_s + buf * n4;
            const int32_t *dcol = dcol_s + buf * n4;
            const int32_t *frow = frow_s + buf * n4;    // F[a][.]
            const int32_t *ftrow = ftrow_s + buf * n4;  // F[.][a]
            const int32_t Iia = Iget(i, a);
            const int32_t Dii = SYM ? 0 : ddiag_s[i];
            const int32_t faa = SYM ? 0 : frow[a];

            // (1) O(1) deltas of every candidate (CPU_PIM.cpp:220-232)
            int32_t dE[CPL];
            int lmax = 0;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const int b = tid + q * NT;
                dE[q] = INT32_MIN;
                if (b < n) {
                    const int j = pv[q];
                    const int32_t force = Iget(i, b) + Iget(j, a) - Iia - Iget(j, b);
                    int32_t v;
                    if constexpr (SYM) {
                        v = force - 4 * drow[j] * frow[b];
                    } else {
                        const int32_t dist = Dii + ddiag_s[j] - drow[j] - dcol[j];
                        const int32_t flow = faa + fdg[q] - frow[b] - ftrow[b];
                        v = force - dist * flow;
                    }
                    dE[q] = v;
                    lmax = max(lmax, v);
                }
            }
            // rows of the next site (loaded during the previous update) into the other
            // buffers, which nobody reads this update; then the loads for the site after
            const int inext = (i + 1 == n) ? 0 : i + 1;
            const int inext2 = (inext + 1 == n) ? 0 : inext + 1;
            const int a_nx = p_s[inext];
            const int a_nx2 = p_s[inext2];
            {
                int32_t *dnb = drow_s + (buf ^ 1) * n4;
                int32_t *dtb = dcol_s + (buf ^ 1) * n4;
                int32_t *fnb = frow_s + (buf ^ 1) * n4;
                int32_t *ftb = ftrow_s + (buf ^ 1) * n4;
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const int bb = tid + q * NT;
                    if (bb < n) {
                        dnb[bb] = dn[q];
                        fnb[bb] = frn[q];
                        if constexpr (!SYM) {
                            dtb[bb] = dtn[q];
                            ftb[bb] = ftrn[q];
                        }
                    }
                }
            }
            const int a_nbuf = a_rec;  // the facility whose flow row is now in buf ^ 1
            a_rec = a_nx2;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const int bc = min(tid + q * NT, n - 1);
                dn[q] = D[(int64_t)inext2 * n + bc];
                frn[q] = F[(int64_t)a_nx2 * n + bc];
                if constexpr (!SYM) {
                    dtn[q] = Dt[(int64_t)inext2 * n + bc];
                    ftrn[q] = Ft[(int64_t)a_nx2 * n + bc];
                }
            }
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

            // (3) winner: first b whose inclusive prefix exceeds floor(u * S / 2^32);
            //     it commits the permutation swap before the broadcast barrier.
            const uint32_t thr = (uint32_t)(((uint64_t)ri * (uint64_t)S) >> 32);
            int my_b = a, my_j = i;
            int32_t my_d = 0;
            bool owner = false;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const uint32_t ex = qbase[q] + wbase[q] + incl[q] - w[q];
                if (ex <= thr && thr < ex + w[q]) {
                    owner = true;
                    my_b = tid + q * NT;
                    my_j = pv[q];
                    my_d = dE[q];
                }
            }
            if (owner && my_b != a) {  // the winner commits the permutation swap
                p_s[i] = my_b;
                p_s[my_j] = a;
            }
            int b, j;
            int32_t dsel;
            if constexpr (NW == 1) {
                // one wave: the winner's values travel by readlane, no LDS round trip
                const uint64_t bal = __ballot(owner);
                if (bal) {
                    const int src = __ffsll((unsigned long long)bal) - 1;
                    b = __builtin_amdgcn_readlane(my_b, src);
                    j = __builtin_amdgcn_readlane(my_j, src);
                    dsel = __builtin_amdgcn_readlane(my_d, src);
                } else {  // every weight truncated to zero: stay
                    b = a;
                    j = i;
                    dsel = 0;
                }
                lds_sync<NW>();
            } else {
                if (owner) {
                    red_s[96] = my_b;
                    red_s[97] = my_d;
                    red_s[98] = my_j;
                }
                if (S == 0 && tid == 0) {  // every weight truncated to zero: stay
                    red_s[96] = a;
                    red_s[97] = 0;
                    red_s[98] = i;
                }
                lds_sync<NW>();
                b = red_s[96];
                dsel = red_s[97];
                j = red_s[98];
            }
            if (b != a) {  // owned inverse-permutation entries follow the swap
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const int bb = tid + q * NT;
                    if (bb == b) pv[q] = i;
                    else if (bb == a) pv[q] = j;
                }
            }

            if (b != a) {
                // (4) commit (CPU_PIM.cpp:252-283)
                E -= dsel;
                ++acc;
                const int cla = cls_s[a], clb = cls_s[b];
                // (not in the int32 LDS kernel: the extra code there cost 2.5-3.5 % on
                // instances without twins, n = 64 dense and n = 100 sparse flow)
                if (SYM && IM != IM_LDS32 && ((cla ^ clb) & 0xffff) != 0 && ((cla ^ clb) >> 16) == 0) {
                    // twin facilities (symmetric flow rows equal outside columns a and b):
                    // v = 2 (F[b] - F[a]) is 2 F[a][b] at column a, its negative at column
                    // b and zero elsewhere, so the update is two columns of I += u v^T with
                    // u = D[j] - D[i]; every thread updates its own rows, no column scan
                    const int32_t f2 = 2 * frow[b];
                    for (int k = tid; k < n; k += NT) {
                        const int32_t dv = (D[(int64_t)j * n + k] - drow[k]) * f2;
                        IT *ra = Iaddr(k, a), *rb = Iaddr(k, b);
                        const int32_t xa = (int32_t)*ra, xb = (int32_t)*rb;
                        if constexpr (IM == IM_LDS16) {  // -I stored
                            *ra = (uint16_t)(xa - dv);
                            *rb = (uint16_t)(xb + dv);
                        } else {
                            *ra = xa + dv;
                            *rb = xb - dv;
                        }
                    }
                    if (wid == 0) cells += 2 * (uint64_t)n;
                    if constexpr (IM == IM_HBM) __syncthreads();
                    else lds_sync<NW>();
                } else if ((cla ^ clb) & 0xffff) {
                    // The flow differences are nonzero only inside the union of the two
                    // facilities' nonzero flow columns (static, from the host).  A dense
                    // instance updates that range with 128-bit accesses; a sparse one counts
                    // the nonzeros first and compacts them when the range is much longer.
                    const uint32_t sa = (uint32_t)fsp_s[a], sb = (uint32_t)fsp_s[b];
                    const int s_lo = (int)min(sa & 1023u, sb & 1023u);
                    const int s_hi = (int)max((sa >> 10) & 1023u, (sb >> 10) & 1023u);
                    constexpr int VW = (IM == IM_LDS16) ? 8 : 4;
                    int32_t v1[CPL], v2[CPL];
                    uint32_t nz = 0;
#pragma unroll
                    for (int q = 0; q < CPL; ++q) {
                        const int c = tid + q * NT;
                        v1[q] = 0;
                        v2[q] = 0;
                        const int cc = min(c, n - 1);
                        const int32_t fb = F[(int64_t)b * n + cc];
                        const int32_t dj = D[(int64_t)j * n + cc];
                        int32_t ftb = 0, dtj = 0;
                        if constexpr (!SYM) {
                            ftb = Ft[(int64_t)b * n + cc];
                            dtj = Dt[(int64_t)j * n + cc];
                        }
                        if (c < n) {
                            if constexpr (SYM) {
                                v1[q] = 2 * (fb - frow[c]);          // 2 (F[b][c] - F[a][c])
                                u1_s[c] = dj - drow[c];              // D[j][k] - D[i][k]
                            } else {
                                v1[q] = ftb - ftrow[c];              // F[c][b] - F[c][a]
                                v2[q] = fb - frow[c];                // F[b][c] - F[a][c]
                                u1_s[c] = dtj - dcol[c];             // D[k][j] - D[k][i]
                                u2_s[c] = dj - drow[c];              // D[j][k] - D[i][k]
                            }
                            nz += (v1[q] | v2[q]) != 0;
                        }
                    }
                    uint32_t nnz = (uint32_t)n;
                    const int c_lo = s_lo, c_hi = s_hi;
                    bool vec_path = IM != IM_HBM;  // dense flow: the range, 128-bit accesses
                    // per-wave ballot masks of the nonzero columns, slot q of wave w at
                    // word w * CPL + q (the pick's partial sums there are dead by now)
                    uint64_t *msk_s = reinterpret_cast<uint64_t *>(red_s + 16);
                    uint64_t mk[NW * CPL];  // the masks, read once (dependent LDS reads are slow)
                    if (sparse_upd) {
                        // v of every column and one mask per wave and slot, then one barrier:
                        // every wave knows the exact nonzero count and the column set
#pragma unroll
                        for (int q = 0; q < CPL; ++q) {
                            const int c = tid + q * NT;
                            if (c < n4) {
                                vd1_s[c] = v1[q];
                                if constexpr (!SYM) vd2_s[c] = v2[q];
                            }
                            const uint64_t bm = __ballot((v1[q] | v2[q]) != 0);
                            if (lane == 0) msk_s[wid * CPL + q] = bm;
                        }
                        lds_sync<NW>();
                        nnz = 0;
#pragma unroll
                        for (int m = 0; m < NW * CPL; ++m) {
                            mk[m] = msk_s[m];
                            nnz += (uint32_t)__popcll(mk[m]);
                        }
                        vec_path = IM != IM_HBM && c_hi - c_lo + 1 <= 4 * (int)nnz + 4 * VW;
                    } else {
                        if (vec_path) {  // v of every column for the vector path
#pragma unroll
                            for (int q = 0; q < CPL; ++q) {
                                const int c = tid + q * NT;
                                if (c < n4) {
                                    vd1_s[c] = v1[q];
                                    if constexpr (!SYM) vd2_s[c] = v2[q];
                                }
                            }
                        }
                        lds_sync<NW>();
                    }
                    {
                        uint32_t cnt = 0;
#pragma unroll
                        for (int q = 0; q < CPL; ++q)
                            cnt += (uint32_t)__popcll(__ballot((v1[q] | v2[q]) != 0));
                        cells += (uint64_t)n * cnt;
                    }
                    if (vec_path) {
                        int4 *I4 = reinterpret_cast<int4 *>(I_s);
                        const int V0 = c_lo / VW, V1 = c_hi / VW + 1;
                        if (d.mul24)
                            vec_update<NT, IM, SYM, true>(I4, ld_s / VW, n, V0, V1, u1_s, u2_s,
                                                          vd1_s, vd2_s, tid);
                        else
                            vec_update<NT, IM, SYM, false>(I4, ld_s / VW, n, V0, V1, u1_s, u2_s,
                                                           vd1_s, vd2_s, tid);
                    } else if (!sparse_upd) {
                        IT *colp[CPL];
#pragma unroll
                        for (int q = 0; q < CPL; ++q) colp[q] = Iaddr(0, min(tid + q * NT, n));
                        const int32_t rs = (IM == IM_HBM) ? ld : ld_s;
                        if (d.mul24)
                            dense_update<CPL, IM, SYM, true>(colp, rs, n, u1_s, u2_s, v1, v2);
                        else
                            dense_update<CPL, IM, SYM, false>(colp, rs, n, u1_s, u2_s, v1, v2);
                    } else if (nnz <= (uint32_t)kWalkCols) {
                        // few columns (a swap inside a block of equal flow rows changes two):
                        // walk the set bits of the masks, all threads over the rows of each
                        // The row factors are read once per row; a column costs one LDS round
                        // trip (its v and the entries are independent reads).
                        const int rs = (IM == IM_HBM) ? ld : ld_s;
                        const bool m24 = d.mul24 != 0;
                        for (int k0 = tid; k0 < n; k0 += NT) {
                            const int32_t w1 = u1_s[k0];
                            const int32_t w2 = SYM ? 0 : u2_s[k0];
                            const int off = __mul24(k0, rs);
#pragma unroll
                            for (int m = 0; m < NW * CPL; ++m) {
                                uint64_t word = mk[m];
                                const int cbase = (m / CPL) * kWave + (m % CPL) * NT;
                                while (word) {
                                    const int c = cbase + __builtin_ctzll(word);
                                    word &= word - 1;
                                    IT *cp = Iaddr(0, c);
                                    const int32_t x = (int32_t)cp[off];
                                    const int32_t a1 = vd1_s[c];
                                    int32_t dv = m24 ? __mul24(w1, a1) : w1 * a1;
                                    if constexpr (!SYM) {
                                        const int32_t a2 = vd2_s[c];
                                        dv += m24 ? __mul24(w2, a2) : w2 * a2;
                                    }
                                    if constexpr (IM == IM_LDS16) cp[off] = (uint16_t)(x - dv);
                                    else cp[off] = x + dv;
                                }
                            }
                        }
                    } else {
                        // compacted column list, positions from the masks (no scan)
                        uint32_t before = 0;
#pragma unroll
                        for (int m = 0; m < NW * CPL; ++m)
                            before += (m < wid * CPL) ? (uint32_t)__popcll(mk[m]) : 0u;
                        const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
                        for (int q = 0; q < CPL; ++q) {
                            uint64_t mine = 0;
#pragma unroll
                            for (int m = 0; m < NW * CPL; ++m)
                                if (m == wid * CPL + q) mine = mk[m];
                            if ((v1[q] | v2[q]) != 0) {
                                const uint32_t at = before + (uint32_t)__popcll(mine & lt);
                                col_s[at] = tid + q * NT;
                                v1_s[at] = v1[q];
                                v2_s[at] = v2[q];
                            }
                            before += (uint32_t)__popcll(mine);
                        }
                        lds_sync<NW>();
                        int cw = 1;
                        while (cw < (int)nnz && cw < NT) cw <<= 1;
                        const int shw = __builtin_ctz(cw);
                        const int rw = NT >> shw;
                        const int cx = tid & (cw - 1);
                        const int ry = tid >> shw;
                        const int64_t rs = (IM == IM_HBM) ? (int64_t)ld : (int64_t)ld_s;
                        constexpr int G = 8;
                        // Rows go G at a time per lane with no bounds test inside a group;
                        // lanes past the compacted list work on the padding column.  (A
                        // per-element choice of the 24-bit multiplier here turned into a
                        // branch per row: n = 100 sparse flow went from 12.3 to 10.6 M/s.)
                        for (int ci0 = 0; ci0 < (int)nnz; ci0 += cw) {
                            const int ci = ci0 + cx;
                            const bool live = ci < (int)nnz;
                            const int c = live ? col_s[ci] : n;
                            const int32_t a1 = live ? v1_s[ci] : 0;
                            const int32_t a2 = live ? v2_s[ci] : 0;
                            IT *cp = Iaddr(0, c);
                            for (int k0 = ry; k0 < n; k0 += G * rw) {
                                int32_t x[G], w1[G], w2[G];
#pragma unroll
                                for (int r = 0; r < G; ++r) {
                                    const int k = min(k0 + r * rw, n - 1);
                                    x[r] = (int32_t)cp[k * rs];
                                    w1[r] = u1_s[k];
                                    if constexpr (!SYM) w2[r] = u2_s[k];
                                }
#pragma unroll
                                for (int r = 0; r < G; ++r) {
                                    const int k = k0 + r * rw;
                                    const int32_t dv = SYM ? w1[r] * a1 : w1[r] * a1 + w2[r] * a2;
                                    if (k < n) {
                                        if constexpr (IM == IM_LDS16)
                                            cp[k * rs] = (uint16_t)(x[r] - dv);
                                        else
                                            cp[k * rs] = x[r] + dv;
                                    }
                                }
                            }
                        }
                    }
                    if constexpr (IM == IM_HBM) __syncthreads();
                    else lds_sync<NW>();
                }
                if (E < bestE) {  // CPU_PIM.cpp:271-272, plus the permutation
                    bestE = E;
                    improved = true;
                    // uint16 placement: copied within LDS and written out at the end of
                    // the launch (grey n = 256: 3.6 % faster than global stores in the
                    // loop; in the int32 LDS kernel the LDS copy measured 0.5 % slower)
                    if constexpr (IM == IM_LDS16)
                        for (int x = tid; x < n; x += NT) bestp_s[x] = p_s[x];
                    else
                        for (int x = tid; x < n; x += NT) bestpg[x] = p_s[x];
                }
            }
            // the prefetched rows become current
            buf ^= 1;
            a_cur = a_nbuf;
            a_reg = (b != a && j == inext) ? a : a_nx;
        }
    }
    }  // intervals

    if (lane == 0) reinterpret_cast<uint64_t *>(red_s + 140)[wid] = cells;
    lds_sync<NW>();
    for (int x = tid; x < n; x += NT) {
        pg[x] = p_s[x];
        if (IM == IM_LDS16 && improved) bestpg[x] = bestp_s[x];
    }
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
        const int bb = tid + q * NT;
        if (bb < n) pinvg[bb] = pv[q];
    }
    if constexpr (IM == IM_LDS32 || IM == IM_LDS16) {  // write the matrix back, flat as staged
        v4i *Ig4 = reinterpret_cast<v4i *>(Ig);
        const int ldq = ld_s >> 2;
        int row = row_t, cx = cx_t;
        for (int gq = tid; gq < Q; gq += NT) {
            v4i o;
            if constexpr (IM == IM_LDS32) {
                o = reinterpret_cast<const v4i *>(I_s)[row * ldq + cx];
            } else {
                const uint2 w = reinterpret_cast<const uint2 *>(I_s)[row * ldq + cx];
                const int c = 4 * cx;
                o.x = (c + 0 < n) ? -(int32_t)(w.x & 0xffffu) : 0;
                o.y = (c + 1 < n) ? -(int32_t)(w.x >> 16) : 0;
                o.z = (c + 2 < n) ? -(int32_t)(w.y & 0xffffu) : 0;
                o.w = (c + 3 < n) ? -(int32_t)(w.y >> 16) : 0;
            }
            Ig4[(int64_t)row * L4 + cx] = o;
            advance(row, cx);
        }
    }
    if (tid == 0) {
        P.E[g] = E;
        P.bestE[g] = bestE;
        P.accepted[g] = acc;
        uint64_t tot = cells;
        if constexpr (NW > 1) {
            const uint64_t *wc = reinterpret_cast<const uint64_t *>(red_s + 140);
#pragma unroll
            for (int w = 1; w < NW; ++w) tot += wc[w];
        }
        P.cells[g] += tot;
    }
    // the mirrored exchange state, left as exchange_instance would have left it
    if (mirror && local == 0 && wid == 0 && xstat != kXmFault)
        xm_store(P, inst, d, xm, xstat == kXmRun, lane);
}

}  // namespace qap

// the speculative sweep kernel shares this file's exchange step and persistent protocol
#include "qap_pt_sweep_spec.h"

namespace qap {

// ---------------------------------------------------------------------------
// Exchange: one workgroup per instance.  Energies, bests, ladder order and positions
// are staged into LDS with coalesced loads; one thread per ladder runs the sequential
// fold / bubble pass / kick (CPU_PIM.cpp:295-344) on LDS; the goal test
// (:346-348) uses the best over all ladders of the instance.
// ---------------------------------------------------------------------------
constexpr int kExMaxRT = 2048;
constexpr int kExWaveLadders = 8;  // at most this many ladders of <= 64 rungs: wave path

// LDS of an exchange workgroup staging up to cap replicas: sized by the pool, so a launch
// does not ask for room for kExMaxRT replicas beside the sweep workgroups of other pools
// (with 50 KB a serving bank's exchange waited for CUs to free: 43 us per launch).
__host__ __device__ inline int exchange_lds_bytes(int cap) { return 24 * cap; }

// sum over the wave, every lane gets it
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) v += (uint32_t)__shfl_xor((int)v, off);
    return v;
}

// One instance's exchange step (CPU_PIM.cpp:295-357), run by one wave (lane 0..63):
// fold of the replica bests into the ladder bests, the bubble pass, the kick, the goal test
// and the instance's interval bookkeeping.  With cap > 0 the energies, bests, ladder order
// and positions of up to cap replicas are staged through LDS at ex_smem (24 bytes per
// replica, coalesced loads issued in one round); with cap == 0 they are read and written in
// global memory directly (the persistent sweep kernel, whose LDS holds its replica).
__device__ void exchange_instance(const PoolDev &P, int inst, int cap, unsigned char *ex_smem,
                                  int tid) {
    int64_t *E_s = reinterpret_cast<int64_t *>(ex_smem);
    int64_t *bE_s = E_s + cap;
    int32_t *ord_s = reinterpret_cast<int32_t *>(bE_s + cap);
    int32_t *pos_s = ord_s + cap;
    // Every load of the step is issued first, in one round (all of it is valid memory
    // whether or not the instance still runs); the decision to run comes after.
    const InstDesc d = P.inst[inst];
    const InstState st = P.state[inst];
    const int R = d.R, T = d.T, RT = R * T;
    const bool staged = RT <= cap;
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
    }
    int32_t tss_pf = 0;
    int64_t lbest_pf = INT64_MAX;
    if (tid < T) {
        tss_pf = P.tss[d.ladder_base + tid];
        lbest_pf = P.ladder_best[d.ladder_base + tid];
    }
    lds_sync<1>();
    if (st.halted || st.megas_done >= st.mega_limit) {
        if (tid == 0) P.state[inst].active = 0;
        return;
    }
    int64_t gbest = INT64_MAX;
    uint32_t nex = 0, nkick = 0;
    if (R <= kWave && T <= kExWaveLadders) {
        // Few short ladders: the wave walks each ladder together, lane l holding ladder
        // position l, so the sequential bubble pass runs on lane reads of registers
        // instead of dependent LDS round trips (the same pass, step for step).
        for (int t = 0; t < T; ++t) {
            const int L = d.ladder_base + t;
            const int base = t * R;
            const bool live = tid < R;
            const int r_l = live ? ord[base + tid] : 0;
            const int64_t e_l = live ? E[base + r_l] : INT64_MAX;
            const int64_t mb = wave_min64(live ? bE[base + tid] : INT64_MAX);
            int32_t tss = __builtin_amdgcn_readlane(tss_pf, t) + 1;  // lane t prefetched ladder t
            int64_t lbest = readlane64(lbest_pf, t);
            if (mb < lbest) {  // best fold (CPU_PIM.cpp:297-321)
                lbest = mb;
                tss = 0;
            }
            // bubble pass (CPU_PIM.cpp:323-344): the element carried from position k
            // meets the one at k + 1; the lower cost moves on, the other stays at k
            int64_t cE = readlane64(e_l, 0);
            int cR = __builtin_amdgcn_readlane(r_l, 0);
            int newR = r_l;
            for (int k = 0; k < R - 1; ++k) {
                const int64_t nE = readlane64(e_l, k + 1);
                const int nR = __builtin_amdgcn_readlane(r_l, k + 1);
                if (cE < nE) {
                    newR = tid == k ? nR : newR;
                    ++nex;
                    if (nR == R - 1) tss = 0;
                } else {
                    newR = tid == k ? cR : newR;
                    cE = nE;
                    cR = nR;
                }
            }
            newR = tid == R - 1 ? cR : newR;
            if (tss > d.kick_after) {  // kick: the hottest and coldest positions trade
                const int r0 = __builtin_amdgcn_readlane(newR, 0);
                const int rL = __builtin_amdgcn_readlane(newR, R - 1);
                newR = tid == 0 ? rL : (tid == R - 1 ? r0 : newR);
                ++nkick;
            }
            if (live) {
                ord[base + tid] = newR;
                pos[base + newR] = tid;
            }
            if (tid == 0) {
                P.tss[L] = tss;
                P.ladder_best[L] = lbest;
            }
            gbest = lbest < gbest ? lbest : gbest;
        }
        // the counts and best are wave-uniform: lane 0 reports them
        gbest = readlane64(gbest, 0);
        nex = (uint32_t)__builtin_amdgcn_readlane((int)nex, 0);
        nkick = (uint32_t)__builtin_amdgcn_readlane((int)nkick, 0);
    } else {
        for (int t = tid; t < T; t += 64) {
            const int L = d.ladder_base + t;
            const int base = t * R;
            int32_t tss = (t == tid ? tss_pf : P.tss[L]) + 1;
            int64_t lbest = t == tid ? lbest_pf : P.ladder_best[L];
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
        gbest = wave_min64(gbest);
        nex = wave_sum_u32(nex);
        nkick = wave_sum_u32(nkick);
    }
    lds_sync<1>();
    if (staged) {
        for (int x = tid; x < RT; x += 64) {
            P.order[d.rep_base + x] = ord_s[x];
            P.pos[d.rep_base + x] = pos_s[x];
        }
    }
    if (tid == 0) {
        const int64_t gb = gbest < st.global_best ? gbest : st.global_best;
        InstState &so = P.state[inst];
        so.global_best = gb;
        so.exchanges = st.exchanges + nex;
        so.kicks = st.kicks + nkick;
        if (d.has_goal && gb <= d.goal) {
            so.halted = 1;
            so.active = 0;
        } else {
            so.sweep_base = (uint32_t)(st.megas_done * d.spe);
            so.megas_done = st.megas_done + 1;
            so.active = 1;
        }
    }
}

__global__ __launch_bounds__(64) void exchange_kernel(PoolDev P, int cap) {
    extern __shared__ __align__(16) unsigned char ex_smem[];
    exchange_instance(P, blockIdx.x, cap, ex_smem, threadIdx.x);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
// max_iv > 1: a persistent launch of max_iv exchange intervals, cooperative so that every
// workgroup is resident (hipErrorCooperativeLaunchTooLarge when the device cannot hold them)
// max_iv == 0: nothing is launched; the return says whether nblocks workgroups of the
// kernel fit on the device at once (hipSuccess) or not (hipErrorCooperativeLaunchTooLarge).
template <int NT, int CPL, int IM, bool SYM>
static hipError_t launch_sweep_t(const PoolDev &P, bool sparse, const int32_t *list, int nblocks,
                                 int lds_bytes, int max_iv, hipStream_t s) {
    auto k = sparse ? sweep_kernel<NT, CPL, IM, SYM, true> : sweep_kernel<NT, CPL, IM, SYM, false>;
    const hipError_t e = sparse ? allow_full_lds<sweep_kernel<NT, CPL, IM, SYM, true>>()
                                : allow_full_lds<sweep_kernel<NT, CPL, IM, SYM, false>>();
    if (e != hipSuccess) return e;
    if (max_iv == 0) {
        int per_cu = 0;
        const hipError_t q = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(k), NT, (size_t)lds_bytes);
        if (q != hipSuccess) return q;
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return hipErrorInvalidDevice;
        return (int64_t)per_cu * cus >= nblocks ? hipSuccess : hipErrorCooperativeLaunchTooLarge;
    }
    if (max_iv <= 1) {
        hipLaunchKernelGGL(k, dim3(nblocks), dim3(NT), lds_bytes, s, P, list, 1);
        return hipGetLastError();
    }
    PoolDev pd = P;
    const int32_t *lp = list;
    int iv = max_iv;
    void *args[] = {&pd, &lp, &iv};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k), dim3(nblocks), dim3(NT),
                                      args, (unsigned)lds_bytes, s);
}

template <int NT, int CPL>
static hipError_t launch_sweep_c(const PoolDev &P, const SweepShape &sh, const int32_t *list,
                                 int nblocks, int lds_bytes, int max_iv, hipStream_t s) {
    switch (sh.im * 2 + (sh.sym ? 1 : 0)) {
    case IM_LDS32 * 2 + 1: return launch_sweep_t<NT, CPL, IM_LDS32, true>(P, sh.sparse, list, nblocks, lds_bytes, max_iv, s);
    case IM_LDS32 * 2 + 0: return launch_sweep_t<NT, CPL, IM_LDS32, false>(P, sh.sparse, list, nblocks, lds_bytes, max_iv, s);
    case IM_LDS16 * 2 + 1: return launch_sweep_t<NT, CPL, IM_LDS16, true>(P, sh.sparse, list, nblocks, lds_bytes, max_iv, s);
    case IM_LDS16 * 2 + 0: return launch_sweep_t<NT, CPL, IM_LDS16, false>(P, sh.sparse, list, nblocks, lds_bytes, max_iv, s);
    case IM_HBM * 2 + 1: return launch_sweep_t<NT, CPL, IM_HBM, true>(P, sh.sparse, list, nblocks, lds_bytes, 1, s);
    case IM_HBM * 2 + 0: return launch_sweep_t<NT, CPL, IM_HBM, false>(P, sh.sparse, list, nblocks, lds_bytes, 1, s);
    default: return hipErrorInvalidValue;
    }
}

// The speculative kernel, launched like launch_sweep_t (max_iv 0: occupancy check only).
template <int S>
static hipError_t launch_spec_t(const PoolDev &P, const int32_t *list, int nblocks, int lds_bytes,
                                int max_iv, hipStream_t s) {
    auto k = sweep_spec_kernel<S>;
    const hipError_t e = allow_full_lds<sweep_spec_kernel<S>>();
    if (e != hipSuccess) return e;
    if (max_iv == 0) {
        int per_cu = 0, dev = 0, cus = 0;
        hipError_t q = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(k), S * kWave, (size_t)lds_bytes);
        if (q == hipSuccess) q = hipGetDevice(&dev);
        if (q == hipSuccess) q = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (q != hipSuccess) return q;
        return (int64_t)per_cu * cus >= nblocks ? hipSuccess : hipErrorCooperativeLaunchTooLarge;
    }
    if (max_iv <= 1) {
        hipLaunchKernelGGL(k, dim3(nblocks), dim3(S * kWave), lds_bytes, s, P, list, 1);
        return hipGetLastError();
    }
    PoolDev pd = P;
    const int32_t *lp = list;
    int iv = max_iv;
    void *args[] = {&pd, &lp, &iv};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k), dim3(nblocks),
                                      dim3(S * kWave), args, (unsigned)lds_bytes, s);
}

static hipError_t launch_sweep_iv(const PoolDev &P, const SweepShape &sh, const int32_t *list,
                                  int nblocks, int lds_bytes, int max_iv, hipStream_t s) {
    if (sh.spec > 0) {
        if (sh.im != IM_LDS32 || !sh.sym) return hipErrorInvalidValue;
        switch (sh.spec) {
        case 2: return launch_spec_t<2>(P, list, nblocks, lds_bytes, max_iv, s);
        case 4: return launch_spec_t<4>(P, list, nblocks, lds_bytes, max_iv, s);
        case 8: return launch_spec_t<8>(P, list, nblocks, lds_bytes, max_iv, s);
        case 16: return launch_spec_t<16>(P, list, nblocks, lds_bytes, max_iv, s);
        default: return hipErrorInvalidValue;
        }
    }
    switch (sh.nt * 8 + sh.cpl) {
    case 64 * 8 + 1: return launch_sweep_c<64, 1>(P, sh, list, nblocks, lds_bytes, max_iv, s);
    case 64 * 8 + 2: return launch_sweep_c<64, 2>(P, sh, list, nblocks, lds_bytes, max_iv, s);
    case 64 * 8 + 4: return launch_sweep_c<64, 4>(P, sh, list, nblocks, lds_bytes, max_iv, s);
    case 128 * 8 + 1: return launch_sweep_c<128, 1>(P, sh, list, nblocks, lds_bytes, max_iv, s);
    case 128 * 8 + 2: return launch_sweep_c<128, 2>(P, sh, list, nblocks, lds_bytes, max_iv, s);
    case 256 * 8 + 1: return launch_sweep_c<256, 1>(P, sh, list, nblocks, lds_bytes, max_iv, s);
    case 256 * 8 + 2: return launch_sweep_c<256, 2>(P, sh, list, nblocks, lds_bytes, max_iv, s);
    case 256 * 8 + 4: return launch_sweep_c<256, 4>(P, sh, list, nblocks, lds_bytes, max_iv, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_sweep(const PoolDev &P, const SweepShape &sh, const int32_t *list, int nblocks,
                        int lds_bytes, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (sh.im == IM_LAZY) return launch_sweep_lazy(P, sh, list, nblocks, lds_bytes, s);
    if (sh.im == IM_ONESHOT) return launch_oneshot(P, sh, list, nblocks, lds_bytes, s);
    if (sh.im == IM_HBM64) return launch_sweep64(P, sh, list, nblocks, lds_bytes, s);
    return launch_sweep_iv(P, sh, list, nblocks, lds_bytes, 1, s);
}

hipError_t launch_sweep_persistent(const PoolDev &P, const SweepShape &sh, const int32_t *list,
                                   int nblocks, int lds_bytes, int intervals, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (sh.im != IM_LDS32 && sh.im != IM_LDS16) return hipErrorInvalidValue;
    return launch_sweep_iv(P, sh, list, nblocks, lds_bytes, intervals < 1 ? 1 : intervals, s);
}

bool sweep_persistent_fits(const SweepShape &sh, int nblocks, int lds_bytes) {
    if (sh.im != IM_LDS32 && sh.im != IM_LDS16) return false;
    return launch_sweep_iv(PoolDev{}, sh, nullptr, nblocks, lds_bytes, 0, nullptr) == hipSuccess;
}

bool sweep_shape_supported(int nt, int cpl, int im) {
    if (im == IM_ONESHOT) return (nt == 256 || nt == 512) && (cpl == 1 || cpl == 2 || cpl == 4);
    if (im == IM_HBM64)
        return (cpl == 1 && (nt == 64 || nt == 128 || nt == 256)) || (nt == 256 && (cpl == 2 || cpl == 4));
    if (im == IM_LAZY && nt == 512 && (cpl == 1 || cpl == 2)) return true;
    switch (nt * 8 + cpl) {
    case 64 * 8 + 1: case 64 * 8 + 2: case 64 * 8 + 4: case 128 * 8 + 1: case 128 * 8 + 2:
    case 256 * 8 + 1: case 256 * 8 + 2: case 256 * 8 + 4: return true;
    default: return false;
    }
}

int shape_lds_bytes(int n, int ld_s, const SweepShape &sh) {
    if (sh.spec > 0) return 4 * spec_lds_words(n, ld_s, sh.spec);
    return sweep_lds_bytes(n, ld_s, sh.im);
}

int sweep_lds_bytes(int n, int ld_s, int im) {
    if (im == IM_ONESHOT) return oneshot_lds_bytes(n);
    if (im == IM_HBM64) return sweep64_lds_bytes(n);
    return im == IM_LAZY ? lazy_lds_bytes(n) : 4 * sweep_lds_words(n, ld_s, im);
}

hipError_t launch_exchange(const PoolDev &P, int count, int max_rt, hipStream_t s) {
    const int cap = max_rt < 1 ? 1 : (max_rt < kExMaxRT ? max_rt : kExMaxRT);
    hipLaunchKernelGGL(exchange_kernel, dim3(count), dim3(64), exchange_lds_bytes(cap), s, P, cap);
    return hipGetLastError();
}

}  // namespace qap
