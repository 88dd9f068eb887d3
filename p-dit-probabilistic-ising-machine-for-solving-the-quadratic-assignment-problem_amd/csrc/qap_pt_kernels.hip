// qap_pt_kernels.hip -- set-up and verification kernels (the sweep and exchange
// kernels are in qap_pt_sweep.hip).
//
//   force_kernel     I = -(D G^T + D^T G') for a permutation, 16x16 tiles through LDS
//                    (CPU_PIM.cpp:164-181); used at set-up and by the verifier.
//   energy_kernel    E = sum_ij D[i][j] F[p(i)][p(j)] (CPU_PIM.cpp:184-192).
#include "qap_pt_device.h"
#include "qap_pt_launch.h"

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
