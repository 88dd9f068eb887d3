// qap_pt_sweep64.hip -- p-dit sweeps of instances whose entries are too large for int32
// force matrices: the force matrix is int64 in HBM and updated in place, deltas are
// int64 (GPU-CUPIM's ReplicaPool<F, Z> with Z = int64_t, qap.cu:369-411, chosen by
// --large-energies, qap_gpu.cu:789-793).
//
// sweep64_kernel<NT, CPL>: one workgroup per replica, candidate b owned by thread b % NT,
// the general update rule (CPU_PIM.cpp:206-287) with asymmetric matrices and diagonals:
//   dE_b = I[i][b] + I[j][a] - I[i][a] - I[j][b] - (D_ii + D_jj - D_ij - D_ji)(F_aa + F_bb
//          - F_ab - F_ba),  j = pinv[b], a = p[i];
// heat-bath pick over the 2^20-scaled integer weights of the 64-bit differences
// (heat_weight64, heat_cutoff64: the float steps of the int32 kernels from y on), and on a
// move the rank-2 update I[k][c] += (D[k][j] - D[k][i])(F[c][b] - F[c][a]) +
// (D[j][k] - D[i][k])(F[b][c] - F[a][c]) (CPU_PIM.cpp:263-269) in int64.  The host admits an
// instance here only when 4 B + 16 maxD maxF < 2^62 (B the force bound) and the cost
// n^2 maxD maxF < 2^62, so every sum is exact.  Philox counters and exchange are those of
// the other kernels; the trajectory is the oracle's wide path bit for bit.
#include "qap_pt_device.h"
#include "qap_pt_launch.h"
#include "qap_pt_sweep_common.h"

namespace qap {

namespace {

// p, pinv, rnd, Ddiag, Fdiag (int32) and u1, u2, v1, v2 (int64), round8(n) entries each,
// then reduction slots
constexpr int kRed64 = 32;   // int64 slots: [0, NW) max partials, [24] winner delta
constexpr int kRed32 = 160;  // uint32 slots: [q * 16 + wave] sum partials, [96] b, [97] j

__host__ __device__ inline int sweep64_lds_words(int n) {
    return 5 * round8(n) + 2 * 4 * round8(n) + 2 * kRed64 + kRed32;
}

__device__ __forceinline__ int64_t wave_max64(int64_t v) {
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int64_t o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    return v;
}

}  // namespace

template <int NT, int CPL>
__global__ __launch_bounds__(NT) void sweep64_kernel(PoolDev P, const int32_t *__restrict__ rep_list) {
    constexpr int NW = NT / kWave;
    extern __shared__ __align__(16) int32_t smem[];

    const int g = rep_list[blockIdx.x];
    const int inst = P.rep_inst[g];
    if (!P.state[inst].active) return;
    const InstDesc d = P.inst[inst];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int n = d.n, ld = d.ld, n4 = round8(n);
    const int local = g - d.rep_base;

    int32_t *p_s = smem;
    int32_t *pinv_s = p_s + n4;
    uint32_t *rnd_s = reinterpret_cast<uint32_t *>(pinv_s + n4);
    int32_t *ddg_s = reinterpret_cast<int32_t *>(rnd_s + n4);
    int32_t *fdg_s = ddg_s + n4;
    int64_t *u1_s = reinterpret_cast<int64_t *>(fdg_s + n4);  // n4 is a multiple of 8
    int64_t *u2_s = u1_s + n4;
    int64_t *v1_s = u2_s + n4;
    int64_t *v2_s = v1_s + n4;
    int64_t *red64 = v2_s + n4;
    uint32_t *red32 = reinterpret_cast<uint32_t *>(red64 + kRed64);

    const int32_t *__restrict__ F = P.F + d.mat_off;
    const int32_t *__restrict__ D = P.D + d.mat_off;
    int32_t *pg = P.perm + d.p_base + (int64_t)local * n;
    int32_t *pinvg = P.pinv + d.p_base + (int64_t)local * n;
    int32_t *bestpg = P.best_perm + d.p_base + (int64_t)local * n;
    int64_t *Ig = P.I64 + d.I64_base + (int64_t)local * n * ld;

    for (int x = tid; x < n; x += NT) {
        p_s[x] = pg[x];
        pinv_s[x] = pinvg[x];
        ddg_s[x] = P.Ddiag[d.vec_off + x];
        fdg_s[x] = P.Fdiag[d.vec_off + x];
    }
    int64_t E = P.E[g];
    int64_t bestE = P.bestE[g];
    uint64_t acc = P.accepted[g];
    uint64_t cells = 0;  // this thread's share: n x the nonzero flow-difference columns it updates
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
            const uint32_t ri = rnd_s[i];
            const int64_t Iia = Ig[(int64_t)i * ld + a];
            const int64_t Dii = ddg_s[i], Faa = fdg_s[a];

            // (1) deltas of every candidate (CPU_PIM.cpp:220-232)
            int64_t dE[CPL];
            int64_t lmax = 0;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const int b = tid + q * NT;
                dE[q] = INT64_MIN;
                if (b < n) {
                    const int j = pinv_s[b];
                    const int64_t force = Ig[(int64_t)i * ld + b] + Ig[(int64_t)j * ld + a] - Iia -
                                          Ig[(int64_t)j * ld + b];
                    const int64_t dist = Dii + ddg_s[j] - D[(int64_t)i * n + j] - D[(int64_t)j * n + i];
                    const int64_t flow = Faa + fdg_s[b] - F[(int64_t)a * n + b] - F[(int64_t)b * n + a];
                    dE[q] = force - dist * flow;
                    lmax = dE[q] > lmax ? dE[q] : lmax;
                }
            }
            int64_t maxd = wave_max64(lmax);
            if constexpr (NW > 1) {
                if (lane == 0) red64[wid] = maxd;
                __syncthreads();
                maxd = red64[0];
#pragma unroll
                for (int w = 1; w < NW; ++w) maxd = red64[w] > maxd ? red64[w] : maxd;
            }

            // (2) integer heat-bath weights and prefix sums in candidate order
            const int64_t cut = heat_cutoff64(maxd, c20);
            uint32_t w[CPL], incl[CPL], qbase[CPL], wbase[CPL];
            uint32_t S = 0;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                w[q] = (dE[q] > cut) ? heat_weight64(beta2f, dE[q] - maxd) : 0u;
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
                    for (int q = 0; q < CPL; ++q) red32[q * 16 + wid] = incl[q];
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    uint32_t tq = 0, before = 0;
#pragma unroll
                    for (int ww = 0; ww < NW; ++ww) {
                        const uint32_t x = red32[q * 16 + ww];
                        before += (ww < wid) ? x : 0u;
                        tq += x;
                    }
                    qbase[q] = S;
                    wbase[q] = before;
                    S += tq;
                }
            }

            // (3) winner: the first b whose inclusive prefix exceeds floor(u S / 2^32)
            const uint32_t thr = (uint32_t)(((uint64_t)ri * (uint64_t)S) >> 32);
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const uint32_t ex = qbase[q] + wbase[q] + incl[q] - w[q];
                if (ex <= thr && thr < ex + w[q]) {
                    const int b = tid + q * NT;
                    red32[96] = (uint32_t)b;
                    red32[97] = (uint32_t)pinv_s[b];
                    red64[24] = dE[q];
                }
            }
            if (S == 0 && tid == 0) {  // every weight truncated to zero: stay
                red32[96] = (uint32_t)a;
                red32[97] = (uint32_t)i;
                red64[24] = 0;
            }
            __syncthreads();
            const int b = (int)red32[96], j = (int)red32[97];
            const int64_t dsel = red64[24];
            __syncthreads();  // every thread has read the winner before the slots are reused
            if (b == a) continue;

            // (4) commit (CPU_PIM.cpp:252-283)
            E -= dsel;
            ++acc;
            if (tid == 0) {
                p_s[i] = b;
                p_s[j] = a;
                pinv_s[b] = i;
                pinv_s[a] = j;
            }
            for (int k = tid; k < n; k += NT) {
                u1_s[k] = (int64_t)D[(int64_t)k * n + j] - D[(int64_t)k * n + i];
                u2_s[k] = (int64_t)D[(int64_t)j * n + k] - D[(int64_t)i * n + k];
                v1_s[k] = (int64_t)F[(int64_t)k * n + b] - F[(int64_t)k * n + a];
                v2_s[k] = (int64_t)F[(int64_t)b * n + k] - F[(int64_t)a * n + k];
            }
            __syncthreads();
            // I[k][c] += u1[k] v1[c] + u2[k] v2[c]: a thread per column, down the rows
            for (int c = tid; c < n; c += NT) {
                const int64_t a1 = v1_s[c], a2 = v2_s[c];
                if ((a1 | a2) == 0) continue;
                cells += (uint64_t)n;
                int64_t *col = Ig + c;
                for (int k = 0; k < n; ++k) col[(int64_t)k * ld] += u1_s[k] * a1 + u2_s[k] * a2;
            }
            __threadfence_block();
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
    if (cells) atomicAdd(reinterpret_cast<unsigned long long *>(&P.cells[g]),
                         (unsigned long long)cells);
    if (tid == 0) {
        P.E[g] = E;
        P.bestE[g] = bestE;
        P.accepted[g] = acc;
    }
}

// I (int64, row stride ld) = -(D G^T + D^T G') of nrep replicas, 16x16 output tiles;
// products and sums in int64 (CPU_PIM.cpp:164-181).  blockIdx.z = replica.
__global__ __launch_bounds__(256) void force64_kernel(const int32_t *__restrict__ F,
                                                      const int32_t *__restrict__ D,
                                                      const int32_t *__restrict__ perm, int n,
                                                      int ld, int64_t *__restrict__ Iout) {
    __shared__ int32_t sD[16][17], sDt[16][17], sG[16][17], sG2[16][17];
    const int z = blockIdx.z;
    const int32_t *p = perm + (int64_t)z * n;
    int64_t *I = Iout + (int64_t)z * n * ld;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int i0 = blockIdx.y * 16, a0 = blockIdx.x * 16;
    int64_t acc = 0;
    for (int j0 = 0; j0 < n; j0 += 16) {
        const int j = j0 + tx;
        const int ii = i0 + ty, aa = a0 + ty;
        const int pj = j < n ? p[j] : 0;
        sD[ty][tx] = (ii < n && j < n) ? D[(int64_t)ii * n + j] : 0;   // D[i][j]
        sDt[ty][tx] = (ii < n && j < n) ? D[(int64_t)j * n + ii] : 0;  // D[j][i]
        sG[ty][tx] = (aa < n && j < n) ? F[(int64_t)aa * n + pj] : 0;  // F[a][p(j)]
        sG2[ty][tx] = (aa < n && j < n) ? F[(int64_t)pj * n + aa] : 0; // F[p(j)][a]
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k)
            acc += (int64_t)sD[ty][k] * sG[tx][k] + (int64_t)sDt[ty][k] * sG2[tx][k];
        __syncthreads();
    }
    const int i = i0 + ty, a = a0 + tx;
    if (i < n && a < n) I[(int64_t)i * ld + a] = -acc;
}

// Counts entries where a and b (total_rows int64 rows of stride ld, first n columns) differ.
__global__ void count_mismatch64_kernel(const int64_t *__restrict__ a, const int64_t *__restrict__ b,
                                        int n, int ld, int64_t total_rows,
                                        unsigned long long *__restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.y + threadIdx.y;
    if (r >= total_rows) return;
    unsigned long long c = 0;
    for (int x = threadIdx.x; x < n; x += blockDim.x) c += a[r * ld + x] != b[r * ld + x];
    if (c) atomicAdd(out, c);
}

template <int NT, int CPL>
static hipError_t launch64_t(const PoolDev &P, const int32_t *list, int nblocks, int lds_bytes,
                             hipStream_t s) {
    const hipError_t e = allow_full_lds<sweep64_kernel<NT, CPL>>();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((sweep64_kernel<NT, CPL>), dim3(nblocks), dim3(NT), lds_bytes, s, P, list);
    return hipGetLastError();
}

hipError_t launch_sweep64(const PoolDev &P, const SweepShape &sh, const int32_t *list, int nblocks,
                          int lds_bytes, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    switch (sh.nt * 8 + sh.cpl) {
    case 64 * 8 + 1: return launch64_t<64, 1>(P, list, nblocks, lds_bytes, s);
    case 128 * 8 + 1: return launch64_t<128, 1>(P, list, nblocks, lds_bytes, s);
    case 256 * 8 + 1: return launch64_t<256, 1>(P, list, nblocks, lds_bytes, s);
    case 256 * 8 + 2: return launch64_t<256, 2>(P, list, nblocks, lds_bytes, s);
    case 256 * 8 + 4: return launch64_t<256, 4>(P, list, nblocks, lds_bytes, s);
    default: return hipErrorInvalidValue;
    }
}

int sweep64_lds_bytes(int n) { return 4 * sweep64_lds_words(n); }

hipError_t launch_force64(const int32_t *F, const int32_t *D, const int32_t *perm, int n, int ld,
                          int nrep, int64_t *Iout, hipStream_t s) {
    const int tiles = (n + 15) / 16;
    hipLaunchKernelGGL(force64_kernel, dim3(tiles, tiles, nrep), dim3(256), 0, s, F, D, perm, n,
                       ld, Iout);
    return hipGetLastError();
}

hipError_t launch_count_mismatch64(const int64_t *a, const int64_t *b, int n, int ld,
                                   int64_t total_rows, unsigned long long *out, hipStream_t s) {
    const int rows_per_block = 4;
    const int64_t blocks = (total_rows + rows_per_block - 1) / rows_per_block;
    hipLaunchKernelGGL(count_mismatch64_kernel, dim3((unsigned)blocks), dim3(64, rows_per_block), 0,
                       s, a, b, n, ld, total_rows, out);
    return hipGetLastError();
}

}  // namespace qap
