// qap_pt_kernels.hip -- gfx950 kernels of the PT-SA annealer.
//
//   sweep_kernel     one workgroup = one replica; runs the sweeps of one exchange
//                    interval (CPU_PIM.cpp:195-290, GPU_PIM.cu:304-453).  Candidate b of
//                    the heat-bath update is owned by thread b / CPL; max, prefix sum
//                    and winner selection are wave64 shuffles + ballots with one LDS
//                    hop between waves.  The force matrix lives in LDS for the whole
//                    launch when it fits (staged in and out with 16-byte accesses) and
//                    is read/updated in place in HBM (L2-resident) otherwise.
//   exchange_kernel  best fold, deterministic exchange pass and hot/cold kick of every
//                    ladder (CPU_PIM.cpp:293-357), one thread per instance.
//   force_kernel     I = -(D G^T + D^T G') for a permutation, 16x16 tiles through LDS
//                    (CPU_PIM.cpp:164-181); used at set-up and by the verifier.
//   energy_kernel    E = sum_ij D[i][j] F[p(i)][p(j)] (CPU_PIM.cpp:184-192).
#include "qap_pt_device.h"
#include "qap_pt_launch.h"

namespace qap {

constexpr int kWave = 64;

__device__ inline int wave_max(int v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m, kWave));
    return v;
}

__device__ inline uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, kWave);
        if (lane >= d) v += t;
    }
    return v;
}

// Scratch slots inside red_s (ints): [0,16) max partials, [16,32) scan partials,
// [32,34) winner broadcast.
template <int NW>
__device__ inline int block_max(int v, int *red_s, int wid, int lane) {
    v = wave_max(v);
    if constexpr (NW == 1) {
        return v;
    } else {
        if (lane == 0) red_s[wid] = v;
        __syncthreads();
        int r = red_s[0];
#pragma unroll
        for (int w = 1; w < NW; ++w) r = max(r, red_s[w]);
        return r;
    }
}

template <int NW>
__device__ inline void block_excl_scan(uint32_t v, int *red_s, int wid, int lane, uint32_t &excl,
                                       uint32_t &total) {
    const uint32_t inc = wave_incl_scan(v, lane);
    if constexpr (NW == 1) {
        excl = inc - v;
        total = __shfl(inc, kWave - 1, kWave);
    } else {
        uint32_t *part = reinterpret_cast<uint32_t *>(red_s + 16);
        if (lane == kWave - 1) part[wid] = inc;
        __syncthreads();
        uint32_t before = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t x = part[w];
            before += (w < wid) ? x : 0u;
            tot += x;
        }
        excl = before + inc - v;
        total = tot;
    }
}

__host__ __device__ constexpr int round4(int x) { return (x + 3) & ~3; }

// LDS layout (int32 words): p, pinv, u1, u2, v1, v2, col, rnd (each round4(n)),
// 64 scratch words, then the force matrix n*ld (when staged).
__host__ __device__ inline int sweep_lds_words(int n, int ld, bool lds_i) {
    return 8 * round4(n) + 64 + (lds_i ? n * ld : 0);
}

template <int NT, int CPL, bool LDS_I, bool SYM>
__global__ __launch_bounds__(NT) void sweep_kernel(PoolDev P, const int32_t *__restrict__ rep_list) {
    constexpr int NW = NT / kWave;
    extern __shared__ __align__(16) int32_t smem[];

    const int g = rep_list[blockIdx.x];
    const int inst = P.rep_inst[g];
    if (!P.state[inst].active) return;
    const InstDesc d = P.inst[inst];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int n = d.n, ld = d.ld, n4 = round4(n);
    const int local = g - d.rep_base;

    int32_t *p_s = smem;
    int32_t *pinv_s = p_s + n4;
    int32_t *u1_s = pinv_s + n4;
    int32_t *u2_s = u1_s + n4;
    int32_t *v1_s = u2_s + n4;
    int32_t *v2_s = v1_s + n4;
    int32_t *col_s = v2_s + n4;
    uint32_t *rnd_s = reinterpret_cast<uint32_t *>(col_s + n4);
    int32_t *red_s = reinterpret_cast<int32_t *>(rnd_s + n4);
    int32_t *I_s = red_s + 64;

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

    for (int x = tid; x < n; x += NT) {
        p_s[x] = pg[x];
        pinv_s[x] = pinvg[x];
    }
    if constexpr (LDS_I) {
        const int words4 = (n * ld) >> 2;  // ld % 4 == 0
        const int4 *src = reinterpret_cast<const int4 *>(Ig);
        int4 *dst = reinterpret_cast<int4 *>(I_s);
        for (int x = tid; x < words4; x += NT) dst[x] = src[x];
    }
    int32_t *Ib;
    if constexpr (LDS_I) Ib = I_s;
    else Ib = Ig;

    int64_t E = P.E[g];
    int64_t bestE = P.bestE[g];
    uint64_t acc = P.accepted[g];
    uint64_t cells = P.cells[g];
    const int pos = P.pos[g];
    const float beta2f = P.beta2f[d.beta_off + pos];
    const double c20 = P.c20[d.beta_off + pos];
    const uint32_t sweep0 = P.state[inst].sweep_base;

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
            const int32_t *Irow = Ib + (int64_t)i * ld;
            const int32_t Iia = Irow[a];
            const int32_t *Drow_i = D + (int64_t)i * n;
            const int32_t *Frow_a = F + (int64_t)a * n;
            const int32_t *Dcol_i = Dt + (int64_t)i * n;
            const int32_t *Fcol_a = Ft + (int64_t)a * n;
            const int32_t Dii = SYM ? 0 : Ddg[i];
            const int32_t Faa = SYM ? 0 : Fdg[a];

            // (1) O(1) deltas of every candidate (CPU_PIM.cpp:220-232).
            int32_t dE[CPL];
            int lmax = 0;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const int b = tid * CPL + q;
                dE[q] = INT32_MIN;
                if (b < n) {
                    const int j = pinv_s[b];
                    const int32_t *Irow_j = Ib + (int64_t)j * ld;
                    const int32_t force = Irow[b] + Irow_j[a] - Iia - Irow_j[b];
                    int32_t v;
                    if constexpr (SYM) {
                        v = force - 4 * Drow_i[j] * Frow_a[b];
                    } else {
                        const int32_t dist = Dii + Ddg[j] - Drow_i[j] - Dcol_i[j];
                        const int32_t flow = Faa + Fdg[b] - Frow_a[b] - Fcol_a[b];
                        v = force - dist * flow;
                    }
                    dE[q] = v;
                    lmax = max(lmax, v);
                }
            }
            const int maxd = block_max<NW>(lmax, red_s, wid, lane);

            // (2) integer heat-bath weights and their prefix sum (CPU_PIM.cpp:234-251).
            const int32_t cut = heat_cutoff(maxd, c20);
            uint32_t w[CPL];
            uint32_t lsum = 0;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                w[q] = (dE[q] > cut) ? heat_weight(beta2f, dE[q] - maxd) : 0u;
                lsum += w[q];
            }
            uint32_t excl, S;
            block_excl_scan<NW>(lsum, red_s, wid, lane, excl, S);

            // (3) first candidate whose inclusive prefix exceeds floor(u * S / 2^32).
            const uint32_t thr = (uint32_t)(((uint64_t)rnd_s[i] * (uint64_t)S) >> 32);
            const bool owner = (excl <= thr) && (thr < excl + lsum);
            int sel = 0;
            int32_t dsel = 0;
            if (owner) {
                uint32_t run = excl;
                bool found = false;
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    run += w[q];
                    if (!found && run > thr) {
                        found = true;
                        sel = tid * CPL + q;
                        dsel = dE[q];
                    }
                }
            }
            if constexpr (NW == 1) {
                const uint64_t bal = __ballot(owner);
                const int src = __ffsll((unsigned long long)bal) - 1;
                sel = __shfl(sel, src, kWave);
                dsel = __shfl(dsel, src, kWave);
            } else {
                if (owner) {
                    red_s[32] = sel;
                    red_s[33] = dsel;
                }
                __syncthreads();
                sel = red_s[32];
                dsel = red_s[33];
            }
            if (sel == a) continue;  // "stay": no swap

            // (4) commit: cost, permutation, rank-2 force update (CPU_PIM.cpp:252-283).
            const int j = pinv_s[sel];
            const int b = sel;
            E -= dsel;
            ++acc;
            const int32_t *Drow_j = D + (int64_t)j * n;
            for (int k = tid; k < n; k += NT) {
                if constexpr (SYM) {
                    u1_s[k] = Drow_j[k] - Drow_i[k];
                } else {
                    u1_s[k] = Dt[(int64_t)j * n + k] - Dcol_i[k];  // D[k][j] - D[k][i]
                    u2_s[k] = Drow_j[k] - Drow_i[k];              // D[j][k] - D[i][k]
                }
            }
            // columns with a nonzero flow difference, compacted in order
            uint32_t nz = 0;
            int32_t cv1[CPL], cv2[CPL];
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const int c = tid * CPL + q;
                cv1[q] = 0;
                cv2[q] = 0;
                if (c < n) {
                    if constexpr (SYM) {
                        cv1[q] = 2 * (F[(int64_t)b * n + c] - Frow_a[c]);
                    } else {
                        cv1[q] = Ft[(int64_t)b * n + c] - Fcol_a[c];  // F[c][b] - F[c][a]
                        cv2[q] = F[(int64_t)b * n + c] - Frow_a[c];   // F[b][c] - F[a][c]
                    }
                    nz += (cv1[q] | cv2[q]) != 0;
                }
            }
            uint32_t base, nnz;
            block_excl_scan<NW>(nz, red_s, wid, lane, base, nnz);
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                if ((cv1[q] | cv2[q]) != 0) {
                    col_s[base] = tid * CPL + q;
                    v1_s[base] = cv1[q];
                    v2_s[base] = cv2[q];
                    ++base;
                }
            }
            __syncthreads();
            cells += (uint64_t)n * nnz;
            if (nnz > 0) {
                // 2-D split: CW threads across the compacted columns, NT/CW rows at once.
                int cw = 1;
                while (cw < (int)nnz && cw < NT) cw <<= 1;
                const int rw = NT / cw;
                const int cx = tid & (cw - 1);
                const int ry = tid / cw;
                for (int k = ry; k < n; k += rw) {
                    int32_t *row = Ib + (int64_t)k * ld;
                    const int32_t uk1 = u1_s[k];
                    if constexpr (SYM) {
                        for (int ci = cx; ci < (int)nnz; ci += cw) row[col_s[ci]] += uk1 * v1_s[ci];
                    } else {
                        const int32_t uk2 = u2_s[k];
                        for (int ci = cx; ci < (int)nnz; ci += cw)
                            row[col_s[ci]] += uk1 * v1_s[ci] + uk2 * v2_s[ci];
                    }
                }
            }
            if (tid == 0) {
                p_s[i] = b;
                p_s[j] = a;
                pinv_s[b] = i;
                pinv_s[a] = j;
            }
            __syncthreads();
            if (E < bestE) {  // CPU_PIM.cpp:271-272, plus the permutation
                bestE = E;
                for (int x = tid; x < n; x += NT) bestpg[x] = p_s[x];
            }
        }
    }

    __syncthreads();
    for (int x = tid; x < n; x += NT) {
        pg[x] = p_s[x];
        pinvg[x] = pinv_s[x];
    }
    if constexpr (LDS_I) {
        const int words4 = (n * ld) >> 2;
        const int4 *src = reinterpret_cast<const int4 *>(I_s);
        int4 *dst = reinterpret_cast<int4 *>(Ig);
        for (int x = tid; x < words4; x += NT) dst[x] = src[x];
    }
    if (tid == 0) {
        P.E[g] = E;
        P.bestE[g] = bestE;
        P.accepted[g] = acc;
        P.cells[g] = cells;
    }
}

// One thread per instance: fold, exchange pass, kick, goal (CPU_PIM.cpp:293-357).
__global__ void exchange_kernel(PoolDev P, int count) {
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= count) return;
    InstState &st = P.state[inst];
    const InstDesc d = P.inst[inst];
    if (st.halted || st.megas_done >= st.mega_limit) {
        st.active = 0;
        return;
    }
    int64_t gbest = st.global_best;
    uint64_t nex = 0, nkick = 0;
    const int R = d.R;
    for (int t = 0; t < d.T; ++t) {
        const int L = d.ladder_base + t;
        const int base = d.rep_base + t * R;
        int32_t *order = P.order + base;
        int32_t tss = P.tss[L] + 1;
        int64_t lbest = P.ladder_best[L];
        for (int l = 0; l < R; ++l) {
            const int64_t b = P.bestE[base + l];
            if (b < lbest) {
                lbest = b;
                tss = 0;
            }
        }
        for (int k = 0; k < R - 1; ++k) {
            const int r1 = order[k], r2 = order[k + 1];
            if (P.E[base + r1] < P.E[base + r2]) {
                order[k] = r2;
                order[k + 1] = r1;
                P.pos[base + r1] = k + 1;
                P.pos[base + r2] = k;
                ++nex;
                if (r2 == R - 1) tss = 0;
            }
        }
        if (tss > d.kick_after) {
            const int r1 = order[R - 1], r2 = order[0];
            order[R - 1] = r2;
            order[0] = r1;
            P.pos[base + r1] = 0;
            P.pos[base + r2] = R - 1;
            ++nkick;
        }
        P.tss[L] = tss;
        P.ladder_best[L] = lbest;
        gbest = lbest < gbest ? lbest : gbest;
    }
    st.global_best = gbest;
    st.exchanges += nex;
    st.kicks += nkick;
    if (d.has_goal && gbest <= d.goal) {
        st.halted = 1;
        st.active = 0;
        return;
    }
    st.sweep_base = (uint32_t)(st.megas_done * d.spe);
    st.megas_done += 1;
    st.active = 1;
}

__global__ void set_limits_kernel(InstState *state, const int64_t *limits, int count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) state[k].mega_limit = limits[k];
}

__global__ void halt_kernel(InstState *state, int count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) {
        state[k].halted = 1;
        state[k].active = 0;
    }
}

// I (row stride ld) of nrep replicas of one instance, 16x16 output tiles.
// blockIdx.z = replica (local index), perms at perm + z*n.
__global__ __launch_bounds__(256) void force_kernel(const int32_t *__restrict__ F,
                                                    const int32_t *__restrict__ D,
                                                    const int32_t *__restrict__ perm, int n,
                                                    int ld, int32_t *__restrict__ Iout) {
    __shared__ int32_t sD[16][17], sDt[16][17], sG[16][17], sG2[16][17];
    const int z = blockIdx.z;
    const int32_t *p = perm + (int64_t)z * n;
    int32_t *I = Iout + (int64_t)z * n * ld;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int i0 = blockIdx.y * 16, a0 = blockIdx.x * 16;
    int32_t acc = 0;
    for (int j0 = 0; j0 < n; j0 += 16) {
        const int j = j0 + tx;
        const int ii = i0 + ty, aa = a0 + ty;
        const int pj = j < n ? p[j] : 0;
        sD[ty][tx] = (ii < n && j < n) ? D[(int64_t)ii * n + j] : 0;        // D[i][j]
        sDt[ty][tx] = (ii < n && j < n) ? D[(int64_t)j * n + ii] : 0;       // D[j][i]
        sG[ty][tx] = (aa < n && j < n) ? F[(int64_t)aa * n + pj] : 0;       // F[a][p(j)]
        sG2[ty][tx] = (aa < n && j < n) ? F[(int64_t)pj * n + aa] : 0;      // F[p(j)][a]
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) acc += sD[ty][k] * sG[tx][k] + sDt[ty][k] * sG2[tx][k];
        __syncthreads();
    }
    const int i = i0 + ty, a = a0 + tx;
    if (i < n && a < n) I[(int64_t)i * ld + a] = -acc;
}

// E = sum_ij D[i][j] F[p(i)][p(j)] for nrep replicas (blockIdx.x = replica).
__global__ __launch_bounds__(256) void energy_kernel(const int32_t *__restrict__ F,
                                                     const int32_t *__restrict__ D,
                                                     const int32_t *__restrict__ perm, int n,
                                                     int64_t *__restrict__ Eout) {
    __shared__ int64_t part[256];
    const int z = blockIdx.x;
    const int32_t *p = perm + (int64_t)z * n;
    int64_t acc = 0;
    for (int64_t x = threadIdx.x; x < (int64_t)n * n; x += 256) {
        const int i = (int)(x / n), j = (int)(x % n);
        acc += (int64_t)D[x] * F[(int64_t)p[i] * n + p[j]];
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) Eout[z] = part[0];
}

// Counts entries where a and b (nrep matrices n x ld, first n columns) differ.
__global__ void count_mismatch_kernel(const int32_t *__restrict__ a, const int32_t *__restrict__ b,
                                      int n, int ld, int64_t total_rows,
                                      unsigned long long *__restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.y + threadIdx.y;
    if (r >= total_rows) return;
    unsigned long long c = 0;
    for (int x = threadIdx.x; x < n; x += blockDim.x) c += a[r * ld + x] != b[r * ld + x];
    if (c) atomicAdd(out, c);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int NT, int CPL, bool LDS_I, bool SYM>
static hipError_t launch_sweep_t(const PoolDev &P, const int32_t *list, int nblocks, int lds_bytes,
                                 hipStream_t s) {
    auto k = sweep_kernel<NT, CPL, LDS_I, SYM>;
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
    if (sh.lds_i) {
        return sh.sym ? launch_sweep_t<NT, CPL, true, true>(P, list, nblocks, lds_bytes, s)
                      : launch_sweep_t<NT, CPL, true, false>(P, list, nblocks, lds_bytes, s);
    }
    return sh.sym ? launch_sweep_t<NT, CPL, false, true>(P, list, nblocks, lds_bytes, s)
                  : launch_sweep_t<NT, CPL, false, false>(P, list, nblocks, lds_bytes, s);
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

int sweep_lds_bytes(int n, int ld, bool lds_i) { return 4 * sweep_lds_words(n, ld, lds_i); }

hipError_t launch_exchange(const PoolDev &P, int count, hipStream_t s) {
    hipLaunchKernelGGL(exchange_kernel, dim3((count + 63) / 64), dim3(64), 0, s, P, count);
    return hipGetLastError();
}

hipError_t launch_set_limits(InstState *state, const int64_t *limits, int count, hipStream_t s) {
    hipLaunchKernelGGL(set_limits_kernel, dim3((count + 63) / 64), dim3(64), 0, s, state, limits,
                       count);
    return hipGetLastError();
}

hipError_t launch_halt(InstState *state, int count, hipStream_t s) {
    hipLaunchKernelGGL(halt_kernel, dim3((count + 63) / 64), dim3(64), 0, s, state, count);
    return hipGetLastError();
}

hipError_t launch_force(const int32_t *F, const int32_t *D, const int32_t *perm, int n, int ld,
                        int nrep, int32_t *Iout, hipStream_t s) {
    const int tiles = (n + 15) / 16;
    hipLaunchKernelGGL(force_kernel, dim3(tiles, tiles, nrep), dim3(256), 0, s, F, D, perm, n, ld,
                       Iout);
    return hipGetLastError();
}

hipError_t launch_energy(const int32_t *F, const int32_t *D, const int32_t *perm, int n, int nrep,
                         int64_t *Eout, hipStream_t s) {
    hipLaunchKernelGGL(energy_kernel, dim3(nrep), dim3(256), 0, s, F, D, perm, n, Eout);
    return hipGetLastError();
}

hipError_t launch_count_mismatch(const int32_t *a, const int32_t *b, int n, int ld,
                                 int64_t total_rows, unsigned long long *out, hipStream_t s) {
    const int rows_per_block = 4;
    const int64_t blocks = (total_rows + rows_per_block - 1) / rows_per_block;
    hipLaunchKernelGGL(count_mismatch_kernel, dim3((unsigned)blocks), dim3(64, rows_per_block), 0, s,
                       a, b, n, ld, total_rows, out);
    return hipGetLastError();
}

}  // namespace qap
