// qap_pt_oneshot.hip -- one-shot parallel-proposal steps (GPU-CUPIM/perl-gpu/src/lib/qap/
// qap.cu:1037-1609), one workgroup per replica, force matrix in HBM.
//
// One step of a replica:
//   proposals  every location pi draws a partner pj (swap the facilities at pi and pj;
//              pj == pi is the stay) from the step's snapshot with the exact heat-bath
//              rule over pj = 0..n-1 in location order: integer weights with the
//              reference cutoff, 32-bit fixed-point threshold from the Philox word of
//              ctr {pi, step, replica, 3} (qap.cu:1037-1140 draws the same kind of
//              proposal with float weights).  One wave per location, candidate pj in
//              lane pj % 64, slot pj / 64.
//   claims     each moving proposal takes both of its locations with the 64-bit key
//              (~(delta ^ 2^31) << 32) | pi in an LDS atomic minimum: largest delta
//              first, then smallest location (qap.cu:1157-1199).
//   winners    a proposal whose key holds both locations is applied (qap.cu:1203-1221).
//              The winners are disjoint, so the force matrix moves by the sum of their
//              rank-2 updates taken from the snapshot (qap.cu:1301-1376): the update
//              vectors become 16-bit columns of U and V in LDS and the sum is one integer
//              matrix product on the matrix cores.
//   cost       recomputed exactly as -1/2 sum_i I[i][p(i)] (qap.cu:1418-1457), with
//              the new diagonal I[i][p(i)] kept in LDS for the next step's deltas.
// Exchange, counters and the best tracking are those of the sweep kernels; one step
// counts as one sweep of n attempts.
#include "qap_pt_device.h"
#include "qap_pt_launch.h"
#include "qap_pt_mfma_update.h"
#include "qap_pt_sweep_common.h"

namespace qap {

namespace {

constexpr int kOsRed = 64;     // reduction slots
constexpr int kOsMaxKCH = 4;   // update vectors per matrix pass: at most 4 x 32
constexpr int kPanel = 64;     // facilities (force-matrix columns) staged in LDS at a time
constexpr int kPanStride = kPanel + 1;  // LDS row stride of a panel: column reads conflict-free
// p, pinv, dg, rnd, prop, del, list -- round8(n) words each -- and owner (2 words each)
constexpr int kOsArrays = 9;

__host__ __device__ inline int os_fixed_words(int n) { return kOsArrays * round8(n) + kOsRed; }

// Update vectors one matrix pass holds (kcap: a multiple of 32, at most 128) and the U / V
// row stride (kst int16: an odd number of 16-byte groups, so that the 128-bit operand
// reads of consecutive rows by consecutive lanes fall in distinct LDS banks).
struct OsLayout {
    int kcap, kst;
};

__host__ __device__ inline OsLayout os_layout(int n, int lds_bytes) {
    const int avail = lds_bytes - 4 * os_fixed_words(n);
    const int g = avail > 0 ? avail / (4 * round8(n) * 8) : 0;  // 8-entry groups that fit
    int kg = g > 1 ? ((g - 1) / 4) * 4 : 0;                     // multiple of 4, kg + 1 <= g
    if (kg > 4 * kOsMaxKCH) kg = 4 * kOsMaxKCH;
    return OsLayout{8 * kg, kg > 0 ? 8 * (kg + 1) : 0};
}

__device__ __forceinline__ int64_t wave_sum64(int64_t v) {
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) v += __shfl_xor(v, off);
    return v;
}

}  // namespace

int oneshot_lds_bytes(int n) {
    const int want = 4 * os_fixed_words(n) + 4 * round8(n) * 8 * (4 * kOsMaxKCH + 1);
    return want < kMaxLdsBytes ? want : kMaxLdsBytes;
}

bool oneshot_fits(int n) {
    const OsLayout l = os_layout(n, oneshot_lds_bytes(n));
    return l.kcap >= 32 && l.kst >= kPanStride;  // a column panel fits the U / V region
}

template <int NT, int CPL>
__global__ __launch_bounds__(NT) void oneshot_kernel(PoolDev P, const int32_t *__restrict__ rep_list,
                                                     int lds_bytes) {
    constexpr int NW = NT / kWave;  // candidates pj = lane + 64 q, q < CPL: n <= 64 CPL
    // 32-entry chunks of update vectors per matrix pass: eight-wave groups have half the
    // registers per lane, so they take half as many per pass
    constexpr int KCH = NT >= 512 ? kOsMaxKCH / 2 : kOsMaxKCH;
    // locations whose loads a wave issues together (the same 16 per workgroup either way)
    constexpr int kOsBatch = NT >= 512 ? 2 : 4;
    extern __shared__ __align__(16) int32_t smem[];

    const int g = rep_list[blockIdx.x];
    const int inst = P.rep_inst[g];
    if (!P.state[inst].active) return;
    const InstDesc d = P.inst[inst];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int n = d.n, ld = d.ld, n4 = round8(n);
    const int local = g - d.rep_base;
    const OsLayout lay = os_layout(n, lds_bytes);
    // the host launches only shapes that fit: a location per thread, a candidate slot per
    // lane and location, and a panel inside the update-vector region
    if (lay.kcap < 32 || n > NT || n > kWave * CPL || lay.kst < kPanStride) return;

    int32_t *p_s = smem;
    int32_t *pinv_s = p_s + n4;
    int32_t *dg_s = pinv_s + n4;  // I[x][p(x)]
    uint32_t *rnd_s = reinterpret_cast<uint32_t *>(dg_s + n4);
    int32_t *prop_s = reinterpret_cast<int32_t *>(rnd_s + n4);
    int32_t *del_s = prop_s + n4;
    int32_t *list_s = del_s + n4;
    unsigned long long *own_s = reinterpret_cast<unsigned long long *>(list_s + n4);
    int32_t *red_s = reinterpret_cast<int32_t *>(own_s + n4);
    int16_t *U_s = reinterpret_cast<int16_t *>(red_s + kOsRed);  // [n4][kst]
    int16_t *V_s = U_s + (size_t)n4 * lay.kst;                    // [n4][kst]
    int32_t *pan_s = reinterpret_cast<int32_t *>(U_s);  // [n][kPanStride] during proposals

    const int32_t *__restrict__ F = P.F + d.mat_off;
    const int32_t *__restrict__ D = P.D + d.mat_off;
    const int32_t *__restrict__ Ft = P.Ft + d.mat_off;
    const int32_t *__restrict__ Dt = P.Dt + d.mat_off;
    const int32_t *__restrict__ Fd = P.Fdiag + d.vec_off;
    const int32_t *__restrict__ Dd = P.Ddiag + d.vec_off;
    int32_t *pg = P.perm + d.p_base + (int64_t)local * n;
    int32_t *pinvg = P.pinv + d.p_base + (int64_t)local * n;
    int32_t *bestpg = P.best_perm + d.p_base + (int64_t)local * n;
    int32_t *Ig = P.I + d.I_base + (int64_t)local * n * ld;

    for (int x = tid; x < n; x += NT) {
        p_s[x] = pg[x];
        pinv_s[x] = pinvg[x];
    }
    lds_sync<NW>();
    for (int x = tid; x < n; x += NT) dg_s[x] = Ig[(int64_t)x * ld + p_s[x]];
    lds_sync<NW>();

    int64_t E = P.E[g];
    int64_t bestE = P.bestE[g];
    uint64_t acc = P.accepted[g];
    uint64_t cells = 0;
    const int pos = P.pos[g];
    const float beta2f = P.beta2f[d.beta_off + pos];
    const double c20 = P.c20[d.beta_off + pos];
    const uint32_t step0 = P.state[inst].sweep_base;
    const bool sym = d.sym != 0;

    for (int s = 0; s < d.spe; ++s) {
        const uint32_t step = step0 + (uint32_t)s;
        for (int x = tid; x < n; x += NT) {
            rnd_s[x] = philox4x32_10(u32x4{(uint32_t)x, step, (uint32_t)local, 3u}, d.seed_lo,
                                     d.seed_hi)
                           .x;
            own_s[x] = ~0ull;
        }
        lds_sync<NW>();

        // ---- proposals, one panel of kPanel facilities at a time: the panel's columns of I
        // are staged in LDS with coalesced 16-byte loads, then the locations holding
        // those facilities (pinv of the panel) take their column I[.][a] from LDS.  Wave
        // wid takes the panel's locations wid * kOsBatch, ...; the loads of kOsBatch
        // locations are issued before any of their picks.
        for (int f0 = 0; f0 < n; f0 += kPanel) {
            const int fw = min(kPanel, n - f0);
            const int ncg = (fw + 3) / 4;  // 4-column groups of the panel
            for (int e = tid; e < n * (kPanel / 4); e += NT) {
                const int x = e / (kPanel / 4), cg = e % (kPanel / 4);
                if (cg < ncg) {  // columns f0 + 4 cg .. + 3 < round4(n) <= ld: inside row x
                    const v4i v = *reinterpret_cast<const v4i *>(Ig + (int64_t)x * ld + f0 + 4 * cg);
                    int32_t *dst = pan_s + x * kPanStride + 4 * cg;
                    dst[0] = v.x;
                    dst[1] = v.y;
                    dst[2] = v.z;
                    dst[3] = v.w;
                }
            }
            lds_sync<NW>();
        for (int pb = wid * kOsBatch; pb < fw; pb += NW * kOsBatch) {
            int32_t dE[kOsBatch][CPL];
#pragma unroll
            for (int t = 0; t < kOsBatch; ++t) {
                const int pi = pinv_s[f0 + min(pb + t, fw - 1)];
                const int a = p_s[pi];
                const int32_t *rowI = Ig + (int64_t)pi * ld;
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const int pj = min(lane + q * kWave, n - 1);
                    const int b = p_s[pj];
                    const int32_t r = rowI[b];                       // I[pi][b]
                    const int32_t c = pan_s[pj * kPanStride + (a - f0)];  // I[pj][a]
                    int32_t df;
                    if (sym) {
                        df = 4 * D[(int64_t)pi * n + pj] * F[(int64_t)a * n + b];
                    } else {
                        const int32_t dt = Dd[pi] + Dd[pj] - D[(int64_t)pi * n + pj] -
                                           Dt[(int64_t)pi * n + pj];
                        const int32_t ft = Fd[a] + Fd[b] - F[(int64_t)a * n + b] -
                                           Ft[(int64_t)a * n + b];
                        df = dt * ft;
                    }
                    dE[t][q] = r + c - dg_s[pi] - dg_s[pj] - df;
                }
            }
            // the kOsBatch picks run phase by phase, so their wave scans (dependent DPP
            // chains) interleave; a slot past the last location takes no part
            int32_t maxd[kOsBatch];
#pragma unroll
            for (int t = 0; t < kOsBatch; ++t) {
                int lmax = INT32_MIN;
#pragma unroll
                for (int q = 0; q < CPL; ++q)
                    if (lane + q * kWave < n) lmax = max(lmax, dE[t][q]);
                maxd[t] = wave_incl_max(lmax);
            }
            uint32_t w[kOsBatch][CPL], incl[kOsBatch][CPL];
#pragma unroll
            for (int t = 0; t < kOsBatch; ++t) {
                maxd[t] = lane63(maxd[t]);  // >= 0: the stay is a candidate
                const int32_t cut = heat_cutoff(maxd[t], c20);
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const bool live = lane + q * kWave < n && dE[t][q] > cut;
                    w[t][q] = live ? heat_weight(beta2f, dE[t][q] - maxd[t]) : 0u;
                }
            }
#pragma unroll
            for (int t = 0; t < kOsBatch; ++t)
#pragma unroll
                for (int q = 0; q < CPL; ++q) incl[t][q] = wave_incl_sum(w[t][q]);
#pragma unroll
            for (int t = 0; t < kOsBatch; ++t) {
                if (pb + t >= fw) break;
                const int pi = pinv_s[f0 + pb + t];
                uint32_t S = 0, ex[CPL];
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    ex[q] = S + incl[t][q] - w[t][q];
                    S += (uint32_t)lane63((int)incl[t][q]);
                }
                const uint32_t thr = (uint32_t)(((uint64_t)rnd_s[pi] * (uint64_t)S) >> 32);
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    if (w[t][q] > 0 && ex[q] <= thr && thr < ex[q] + w[t][q]) {
                        const int pj = lane + q * kWave;
                        prop_s[pi] = pj;
                        del_s[pi] = pj == pi ? 0 : dE[t][q];
                    }
                }
                if (S == 0 && lane == 0) {  // every weight truncated to zero: stay
                    prop_s[pi] = pi;
                    del_s[pi] = 0;
                }
            }
        }
            lds_sync<NW>();  // the panel is read; proposals of its locations are in LDS
        }

        // ---- claims on both locations of every moving proposal
        for (int pi = tid; pi < n; pi += NT) {
            const int pj = prop_s[pi];
            if (pj != pi) {
                const unsigned long long key =
                    ((unsigned long long)(~((uint32_t)del_s[pi] ^ 0x80000000u)) << 32) | (uint32_t)pi;
                atomicMin(&own_s[pi], key);
                atomicMin(&own_s[pj], key);
            }
        }
        lds_sync<NW>();

        // ---- winners, compacted in location order (location pi = thread pi)
        bool won = false;
        if (tid < n) {
            const int pj = prop_s[tid];
            won = pj != tid && (uint32_t)own_s[tid] == (uint32_t)tid &&
                  (uint32_t)own_s[pj] == (uint32_t)tid;
        }
        const uint32_t incl = wave_incl_sum(won ? 1u : 0u);
        if (lane == kWave - 1) red_s[wid] = (int)incl;
        lds_sync<NW>();
        int m = 0, before = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            before += (w < wid) ? red_s[w] : 0;
            m += red_s[w];
        }
        if (won) list_s[before + (int)incl - 1] = tid;
        lds_sync<NW>();
        if (m == 0) continue;

        // ---- sum of the winners' rank-2 updates (from the snapshot), in passes of at
        // most kcap update vectors
        const int per = sym ? 1 : 2;  // update vectors per swap
        const int spp = min(lay.kcap, 32 * KCH) / per;
        for (int s0 = 0; s0 < m; s0 += spp) {
            const int ms = min(spp, m - s0), K = ms * per, K8 = round8(K);
            for (int e = tid; e < ms * n; e += NT) {
                const int sw = e / n, x = e - sw * n;
                const int pi = list_s[s0 + sw], pj = prop_s[pi];
                const int a = p_s[pi], b = p_s[pj];
                int32_t u1, v1, u2 = 0, v2 = 0;
                if (sym) {  // 2 (D[j][x] - D[i][x]) (F[b][x] - F[a][x]), the 2 on v
                    u1 = D[(int64_t)pj * n + x] - D[(int64_t)pi * n + x];
                    v1 = 2 * (F[(int64_t)b * n + x] - F[(int64_t)a * n + x]);
                } else {    // (D[x][j] - D[x][i]) (F[x][b] - F[x][a]) + (D[j][x] - D[i][x]) (F[b][x] - F[a][x])
                    u1 = Dt[(int64_t)pj * n + x] - Dt[(int64_t)pi * n + x];
                    v1 = Ft[(int64_t)b * n + x] - Ft[(int64_t)a * n + x];
                    u2 = D[(int64_t)pj * n + x] - D[(int64_t)pi * n + x];
                    v2 = F[(int64_t)b * n + x] - F[(int64_t)a * n + x];
                }
                const int k = sw * per;
                U_s[(size_t)x * lay.kst + k] = (int16_t)u1;
                V_s[(size_t)x * lay.kst + k] = (int16_t)v1;
                if (!sym) {
                    U_s[(size_t)x * lay.kst + k + 1] = (int16_t)u2;
                    V_s[(size_t)x * lay.kst + k + 1] = (int16_t)v2;
                }
                cells += (uint64_t)((v1 | v2) != 0) * (uint64_t)n;
            }
            for (int e = tid; e < (K8 - K) * n; e += NT) {  // zero tail up to a group of 8
                const int k = K + e / n, x = e % n;
                U_s[(size_t)x * lay.kst + k] = 0;
                V_s[(size_t)x * lay.kst + k] = 0;
            }
            lds_sync<NW>();
            mfma_rank_update<NW, KCH, NT >= 512 ? 2 : kUpdDepth>(Ig, ld, n, U_s, V_s, lay.kst, K,
                                                                lane, wid);
            __syncthreads();  // matrix writes visible, U / V free for the next pass
        }

        // ---- the winners' swaps (disjoint), then the new diagonal and the exact cost
        for (int k = tid; k < m; k += NT) {
            const int pi = list_s[k], pj = prop_s[pi];
            const int a = p_s[pi], b = p_s[pj];
            p_s[pi] = b;
            p_s[pj] = a;
            pinv_s[b] = pi;
            pinv_s[a] = pj;
        }
        lds_sync<NW>();
        int64_t part = 0;
        for (int x = tid; x < n; x += NT) {
            const int32_t v = Ig[(int64_t)x * ld + p_s[x]];
            dg_s[x] = v;
            part += v;
        }
        part = wave_sum64(part);
        if (lane == 0) reinterpret_cast<int64_t *>(red_s + 16)[wid] = part;
        lds_sync<NW>();
        int64_t sum = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) sum += reinterpret_cast<const int64_t *>(red_s + 16)[w];
        E = -sum / 2;
        acc += (uint64_t)m;
        if (E < bestE) {  // CPU_PIM.cpp:271-272 after the step, plus the permutation
            bestE = E;
            for (int x = tid; x < n; x += NT) bestpg[x] = p_s[x];
        }
        lds_sync<NW>();
    }

    if (lane == 0) reinterpret_cast<uint64_t *>(red_s + 32)[wid] = cells;
    lds_sync<NW>();
    for (int x = tid; x < n; x += NT) {
        pg[x] = p_s[x];
        pinvg[x] = pinv_s[x];
    }
    if (tid == 0) {
        uint64_t tot = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) tot += reinterpret_cast<const uint64_t *>(red_s + 32)[w];
        P.E[g] = E;
        P.bestE[g] = bestE;
        P.accepted[g] = acc;
        P.cells[g] += tot;
    }
}

template <int NT, int CPL>
hipError_t launch_oneshot_t(const PoolDev &P, const int32_t *list, int nblocks, int lds_bytes,
                            hipStream_t s) {
    auto k = oneshot_kernel<NT, CPL>;
    const hipError_t e = allow_full_lds<oneshot_kernel<NT, CPL>>();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(NT), lds_bytes, s, P, list, lds_bytes);
    return hipGetLastError();
}

hipError_t launch_oneshot(const PoolDev &P, const SweepShape &sh, const int32_t *list, int nblocks,
                          int lds_bytes, hipStream_t s) {
    switch (sh.nt * 8 + sh.cpl) {
    case 256 * 8 + 1: return launch_oneshot_t<256, 1>(P, list, nblocks, lds_bytes, s);
    case 256 * 8 + 2: return launch_oneshot_t<256, 2>(P, list, nblocks, lds_bytes, s);
    case 256 * 8 + 4: return launch_oneshot_t<256, 4>(P, list, nblocks, lds_bytes, s);
    case 512 * 8 + 1: return launch_oneshot_t<512, 1>(P, list, nblocks, lds_bytes, s);
    case 512 * 8 + 2: return launch_oneshot_t<512, 2>(P, list, nblocks, lds_bytes, s);
    case 512 * 8 + 4: return launch_oneshot_t<512, 4>(P, list, nblocks, lds_bytes, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace qap
