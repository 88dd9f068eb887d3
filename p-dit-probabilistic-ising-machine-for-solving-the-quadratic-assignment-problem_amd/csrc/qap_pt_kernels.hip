// qap_pt_kernels.hip -- set-up and verification kernels (the sweep and exchange
// kernels are in qap_pt_sweep.hip).
//
//   force_mfma_kernel  the set-up force matrices on the int8 matrix cores: one product
//                      per distinct starting permutation, written to every replica that
//                      starts from it (CPU_PIM.cpp:164-181, GPU-CUPIM's
//                      recompute_force_mats, qap.cu:369-411).
//   force_kernel       I = -(D G^T + D^T G') for a permutation, 16x16 tiles through LDS
//                      on the vector ALUs; the verifier's independent rebuild.
//   energy_kernel      E = sum_ij D[i][j] F[p(i)][p(j)] (CPU_PIM.cpp:184-192).
#include "qap_pt_device.h"
#include "qap_pt_launch.h"
#include "qap_pt_sweep_common.h"

namespace qap {

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

// ---------------------------------------------------------------------------
// Force matrices on the matrix cores.
//
//   -I[i][a] = sum_j D[i][j] F[a][p(j)] + D[j][i] F[p(j)][a]
//            = sum_{k < 2n} A[i][k] B[a][k],   A = [D | D^T],  B = [F P^T | F^T P^T]
// (row a of B: row a of F, then row a of F^T, both gathered by p).  Entries split into
// signed base-256 digits x = sum_s x_s 256^s, x_s in [-128, 127] (PA digits for D, PB
// for F, from the instance's largest magnitudes); every digit pair is one int8 product
// v_mfma_i32_16x16x32_i8, exact in its int32 accumulator (|partial| <= 2n 2^14 < 2^26),
// and the products recombine as sum_L 2^(8L) acc_L in 32-bit two's complement, where
// acc_L sums the pairs with s + t = L (pairs with s + t >= 4 vanish mod 2^32).  The
// result is exact because the entries fit int32 (the host's force bound).
//
// A workgroup (4 waves) computes a 64 (a) x 64 (i) tile: per 64-wide k chunk it stages
// both operands' digit planes in LDS (rows of 64 bytes at a 72-byte stride, so the 16
// rows one operand read touches fall in distinct banks), then each wave multiplies its
// 32 x 32 quarter as 2 x 2 tiles of 16 x 16.  The accumulator layout (lane l: rows
// a = 4 (l / 16) + 0..3, column i = l % 16 of the tile) makes each lane's share of a
// tile one 16-byte store in row i of I; the tile is stored to the `group` replicas that
// share the permutation (blockIdx.z = group index, permutation of its first replica).
// ---------------------------------------------------------------------------
constexpr int kFmTile = 64;     // output tile (a and i)
constexpr int kFmK = 64;        // k per staged chunk
constexpr int kFmRow = 72;      // LDS bytes per staged row
constexpr int kFmMaxPlanes = 4;

__global__ __launch_bounds__(256) void force_mfma_kernel(
    const int32_t *__restrict__ F, const int32_t *__restrict__ D, const int32_t *__restrict__ Ft,
    const int32_t *__restrict__ Dt, const int32_t *__restrict__ perm, int n, int ld, int group,
    int pa, int pb, int32_t *__restrict__ Iout) {
    __shared__ __align__(16) unsigned char sA[kFmMaxPlanes][kFmTile * kFmRow];  // D digits, rows i
    __shared__ __align__(16) unsigned char sB[kFmMaxPlanes][kFmTile * kFmRow];  // F digits, rows a
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int a0 = blockIdx.x * kFmTile, i0 = blockIdx.y * kFmTile;
    const int rep0 = blockIdx.z * group;
    const int32_t *p = perm + (int64_t)rep0 * n;
    const int K = 2 * n;

    // staging: thread t takes row r = t / 16 + 16 h (h = 0..3) and columns 4 (t % 16) + 0..3
    auto digits = [](int32_t x, int np, uint32_t (&w)[kFmMaxPlanes], int byte) {
        int64_t r = x;
#pragma unroll
        for (int s = 0; s < kFmMaxPlanes; ++s) {
            if (s >= np) break;
            const int32_t d = (int32_t)(int8_t)(uint8_t)(r & 0xff);
            w[s] |= ((uint32_t)(uint8_t)d) << (8 * byte);
            r = (r - d) >> 8;
        }
    };
    const int sr = tid >> 4, sc = 4 * (tid & 15);

    const int li = lane & 15, kg = lane >> 4;
    const int wa = (wid >> 1) * 32, wi = (wid & 1) * 32;  // this wave's quarter
    v4i acc[4][2][2];
#pragma unroll
    for (int L = 0; L < 4; ++L)
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) acc[L][x][y] = v4i{0, 0, 0, 0};

    for (int k0 = 0; k0 < K; k0 += kFmK) {
        __syncthreads();  // the previous chunk's operands are consumed
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int r = sr + 16 * h;
            uint32_t wA[kFmMaxPlanes] = {0, 0, 0, 0}, wB[kFmMaxPlanes] = {0, 0, 0, 0};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int k = k0 + sc + e;
                const int i = i0 + r, a = a0 + r;
                int32_t xa = 0, xb = 0;
                if (k < K) {
                    const bool lo = k < n;
                    const int kk = lo ? k : k - n;
                    if (i < n) xa = lo ? D[(int64_t)i * n + kk] : Dt[(int64_t)i * n + kk];
                    if (a < n) {
                        const int pk = p[kk];
                        xb = lo ? F[(int64_t)a * n + pk] : Ft[(int64_t)a * n + pk];
                    }
                }
                digits(xa, pa, wA, e);
                digits(xb, pb, wB, e);
            }
#pragma unroll
            for (int s = 0; s < kFmMaxPlanes; ++s) {
                if (s < pa) *reinterpret_cast<uint32_t *>(&sA[s][r * kFmRow + sc]) = wA[s];
                if (s < pb) *reinterpret_cast<uint32_t *>(&sB[s][r * kFmRow + sc]) = wB[s];
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kFmK / 32; ++ks) {
            const int kb = 32 * ks + 8 * kg;
            long opA[kFmMaxPlanes][2], opB[kFmMaxPlanes][2];  // A: F digits (rows a), B: D digits (cols i)
#pragma unroll
            for (int s = 0; s < kFmMaxPlanes; ++s)
#pragma unroll
                for (int x = 0; x < 2; ++x) {
                    opA[s][x] = s < pb ? *reinterpret_cast<const long *>(&sB[s][(wa + 16 * x + li) * kFmRow + kb]) : 0;
                    opB[s][x] = s < pa ? *reinterpret_cast<const long *>(&sA[s][(wi + 16 * x + li) * kFmRow + kb]) : 0;
                }
#pragma unroll
            for (int t = 0; t < kFmMaxPlanes; ++t) {      // F digit
                if (t >= pb) break;
#pragma unroll
                for (int s = 0; s < kFmMaxPlanes; ++s) {  // D digit
                    if (s >= pa || s + t >= 4) break;
#pragma unroll
                    for (int x = 0; x < 2; ++x)
#pragma unroll
                        for (int y = 0; y < 2; ++y)
                            acc[s + t][x][y] = __builtin_amdgcn_mfma_i32_16x16x32_i8(
                                opA[t][x], opB[s][y], acc[s + t][x][y], 0, 0, 0);
                }
            }
        }
    }

    // -I = sum_L 2^(8 L) acc_L (mod 2^32); lane: column i, rows a .. a + 3 of each tile
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int a = a0 + wa + 16 * x + 4 * kg, i = i0 + wi + 16 * y + li;
            if (i >= n || a >= n) continue;
            v4i o;
#pragma unroll
            for (int v = 0; v < 4; ++v)
                o[v] = (int32_t)(0u - ((uint32_t)acc[0][x][y][v] + ((uint32_t)acc[1][x][y][v] << 8) +
                                       ((uint32_t)acc[2][x][y][v] << 16) +
                                       ((uint32_t)acc[3][x][y][v] << 24)));
            for (int r = 0; r < group; ++r)
                *reinterpret_cast<v4i *>(Iout + (int64_t)(rep0 + r) * n * ld + (int64_t)i * ld + a) = o;
        }
}

// Set-up costs from the force matrices: sum_i I[i][p(i)] = -2 E exactly (the two terms of
// I[i][p(i)] each sum to E over i), an O(n) pass per permutation.  One wave per group of
// replicas sharing a permutation (blockIdx.x = group); the verifier keeps the independent
// O(n^2) energy_kernel.
__global__ __launch_bounds__(64) void energy_from_force_kernel(const int32_t *__restrict__ I,
                                                               const int32_t *__restrict__ perm,
                                                               int n, int ld, int group,
                                                               int64_t *__restrict__ Eout) {
    const int rep0 = blockIdx.x * group;
    const int32_t *p = perm + (int64_t)rep0 * n;
    const int32_t *M = I + (int64_t)rep0 * n * ld;
    int64_t acc = 0;
    for (int i = threadIdx.x; i < n; i += 64) acc += M[(int64_t)i * ld + p[i]];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    for (int r = threadIdx.x; r < group; r += 64) Eout[rep0 + r] = -acc / 2;
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
hipError_t launch_set_limits(InstState *state, const int64_t *limits, int count, hipStream_t s) {
    hipLaunchKernelGGL(set_limits_kernel, dim3((count + 63) / 64), dim3(64), 0, s, state, limits,
                       count);
    return hipGetLastError();
}

hipError_t launch_halt(InstState *state, int count, hipStream_t s) {
    hipLaunchKernelGGL(halt_kernel, dim3((count + 63) / 64), dim3(64), 0, s, state, count);
    return hipGetLastError();
}

__global__ void narrow16_kernel(const int32_t *__restrict__ src, int16_t *__restrict__ dst,
                                int64_t count) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += (int64_t)gridDim.x * blockDim.x)
        dst[k] = (int16_t)max(-32768, min(32767, src[k]));
}

hipError_t launch_narrow16(const int32_t *src, int16_t *dst, int64_t count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, 4096);
    hipLaunchKernelGGL(narrow16_kernel, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, count);
    return hipGetLastError();
}

hipError_t launch_force(const int32_t *F, const int32_t *D, const int32_t *perm, int n, int ld,
                        int nrep, int32_t *Iout, hipStream_t s) {
    const int tiles = (n + 15) / 16;
    hipLaunchKernelGGL(force_kernel, dim3(tiles, tiles, nrep), dim3(256), 0, s, F, D, perm, n, ld,
                       Iout);
    return hipGetLastError();
}

int force_digit_planes(int64_t max_abs) {
    // signed base-256 digits in [-128, 127]: P digits hold [-128 S, 127 S], S = (256^P - 1) / 255
    if (max_abs <= 127) return 1;
    if (max_abs <= 32639) return 2;
    if (max_abs <= 8355711) return 3;
    return 4;
}

hipError_t launch_force_mfma(const int32_t *F, const int32_t *D, const int32_t *Ft,
                             const int32_t *Dt, const int32_t *perm, int n, int ld, int nrep,
                             int group, int planes_d, int planes_f, int32_t *Iout, hipStream_t s) {
    if (group < 1 || nrep % group || planes_d < 1 || planes_d > kFmMaxPlanes || planes_f < 1 ||
        planes_f > kFmMaxPlanes || ld % 4 || ld < n + 1)
        return hipErrorInvalidValue;
    const int tiles = (n + kFmTile - 1) / kFmTile;
    hipLaunchKernelGGL(force_mfma_kernel, dim3(tiles, tiles, nrep / group), dim3(256), 0, s, F, D,
                       Ft, Dt, perm, n, ld, group, planes_d, planes_f, Iout);
    return hipGetLastError();
}

hipError_t launch_energy_from_force(const int32_t *I, const int32_t *perm, int n, int ld, int nrep,
                                   int group, int64_t *Eout, hipStream_t s) {
    if (group < 1 || nrep % group) return hipErrorInvalidValue;
    hipLaunchKernelGGL(energy_from_force_kernel, dim3(nrep / group), dim3(64), 0, s, I, perm, n, ld,
                       group, Eout);
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
