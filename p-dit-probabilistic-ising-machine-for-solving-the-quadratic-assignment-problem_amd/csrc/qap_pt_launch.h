// qap_pt_launch.h -- host-callable launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <mutex>

#include "qap_pt_device.h"

namespace qap {

// LDS available to one workgroup on gfx950 (160 KiB per CU).
constexpr int kMaxLdsBytes = 160 * 1024;
// sites per round of the speculative sweep kernel in latency-bound pools that admit it
// (sixteen waves, four per SIMD, the most a workgroup holds: n = 64, 16 replicas on
// T 1..100: 74.2 M swap-attempts/s against 68.4 M with eight)
constexpr int kSpecSites = 16;

// Lets kernel K take up to kMaxLdsBytes of dynamic LDS.  Set once per process and kernel
// instantiation; callers on several threads (a pool's owner, the serving worker) are
// serialised by a per-kernel mutex, and a failure is returned, never remembered as done,
// so the next launch tries again.
template <auto K>
hipError_t allow_full_lds() {
    static std::atomic<bool> done{false};
    static std::mutex mu;
    if (done.load(std::memory_order_acquire)) return hipSuccess;
    std::lock_guard<std::mutex> lock(mu);
    if (done.load(std::memory_order_relaxed)) return hipSuccess;
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(K),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             kMaxLdsBytes);
    if (e == hipSuccess) done.store(true, std::memory_order_release);
    return e;
}

// Where a replica's force matrix lives during a sweep launch.
enum IMode : int {
    IM_LDS32 = 0,  // int32 in LDS
    IM_LDS16 = 1,  // uint16 holding -I in LDS (F, D >= 0 and bound <= 65535)
    IM_HBM = 2,    // int32 in HBM, updated in place
    IM_LAZY = 3,   // int32 in HBM, rank-2 updates held in LDS and applied in batches
                   // (symmetric instances; 16-bit update vectors)
    IM_ONESHOT = 4,  // one-shot parallel-proposal steps, int32 in HBM (QAP_PT_ONE_SHOT)
    IM_HBM64 = 5,    // int64 in HBM, updated in place (bounds beyond int32)
};

struct SweepShape {
    int nt;      // threads per workgroup (multiple of 64)
    int cpl;     // candidates per thread: n <= nt * cpl
    int im;      // IMode
    bool sym;    // symmetric, zero-diagonal F and D
    // every instance of the launch has sparse flow: the kernel built without the dense
    // update path runs (set per bucket; any instance is still updated exactly there)
    bool sparse = false;
    // > 0: spec sites evaluated at once, one wave each (sweep_spec_kernel: symmetric,
    // n <= 64, int32 LDS), the first moved one committed; 0: one site at a time
    int spec = 0;
};

// whether a kernel instance exists for nt threads x cpl candidates under placement im
bool sweep_shape_supported(int nt, int cpl, int im);
int sweep_lds_bytes(int n, int ld_s, int im);
// LDS of a workgroup of shape sh for size n (the speculative kernel's layout when sh.spec)
int shape_lds_bytes(int n, int ld_s, const SweepShape &sh);
// LDS of the lazy kernel for size n, and whether it holds at least 8 pending updates
int lazy_lds_bytes(int n);
// LDS of a lazy workgroup for size n holding at most max_pending (a multiple of 8) updates
int lazy_lds_bytes(int n, int max_pending);
bool lazy_fits(int n);
// LDS of the one-shot kernel for size n, and whether one matrix pass holds 32 update vectors
int oneshot_lds_bytes(int n);
bool oneshot_fits(int n);

hipError_t launch_sweep(const PoolDev &P, const SweepShape &sh, const int32_t *list, int nblocks,
                        int lds_bytes, hipStream_t s);
// `intervals` exchange intervals of an LDS-placement bucket in one cooperative launch: the
// force matrices stay in LDS and the exchange steps between intervals run in the kernel
// (the first interval's step must have run before the launch).  Needs PoolDev::sync;
// hipErrorCooperativeLaunchTooLarge when the workgroups cannot all be resident.
hipError_t launch_sweep_persistent(const PoolDev &P, const SweepShape &sh, const int32_t *list,
                                   int nblocks, int lds_bytes, int intervals, hipStream_t s);
// whether a persistent launch of nblocks workgroups of this shape fits the device at once
bool sweep_persistent_fits(const SweepShape &sh, int nblocks, int lds_bytes);
hipError_t launch_sweep64(const PoolDev &P, const SweepShape &sh, const int32_t *list, int nblocks,
                          int lds_bytes, hipStream_t s);
int sweep64_lds_bytes(int n);
hipError_t launch_force64(const int32_t *F, const int32_t *D, const int32_t *perm, int n, int ld,
                          int nrep, int64_t *Iout, hipStream_t s);
hipError_t launch_count_mismatch64(const int64_t *a, const int64_t *b, int n, int ld,
                                   int64_t total_rows, unsigned long long *out, hipStream_t s);
// Flush helpers of the lazy kernel: with `helpers` > 0 the launch has nrep replica
// workgroups and helpers x nrep helper workgroups (cooperative, all resident); a replica's
// flush is split by rows between it and its helpers, which read the pending updates from
// `uv` (per replica block: U then V, uv_stride int16 elements) and synchronise through
// `words`: [0] fault (a wait that gave up), then per replica block r [1 + 3 r] flush
// requests, [2 + 3 r] pending count, [3 + 3 r] helper-wave arrivals.  Words after [0] must
// be zero at the start of every launch.
struct FlushHelp {
    int16_t *uv = nullptr;
    uint32_t *words = nullptr;
    int64_t uv_stride = 0;
    int nrep = 0;
    int helpers = 0;
};
// int16 elements of pending updates (U and V) a lazy workgroup of size n holds with
// lds_bytes of LDS
int64_t lazy_pending_elems(int n, int lds_bytes);
// Whether nblocks x (1 + helpers) workgroups of this lazy shape are resident at once.
bool lazy_helpers_fit(const SweepShape &sh, int nblocks, int helpers, int lds_bytes);
// A lazy launch with flush helpers (cooperative); grid = fh.nrep * (1 + fh.helpers).
hipError_t launch_sweep_lazy_helped(const PoolDev &P, const SweepShape &sh, const int32_t *list,
                                    int lds_bytes, const FlushHelp &fh, hipStream_t s);
hipError_t launch_sweep_lazy(const PoolDev &P, const SweepShape &sh, const int32_t *list,
                             int nblocks, int lds_bytes, hipStream_t s);
hipError_t launch_oneshot(const PoolDev &P, const SweepShape &sh, const int32_t *list, int nblocks,
                          int lds_bytes, hipStream_t s);
// max_rt: the largest replicas x ladders of an instance in the pool (staged in LDS up to
// 2048)
hipError_t launch_exchange(const PoolDev &P, int count, int max_rt, hipStream_t s);
hipError_t launch_set_limits(InstState *state, const int64_t *limits, int count, hipStream_t s);
hipError_t launch_halt(InstState *state, int count, hipStream_t s);
// Set-up force matrices of nrep replicas on the int8 matrix cores: replicas g .. g + group - 1
// (g a multiple of group) start from the permutation of replica g, built once and stored to
// each.  planes_d / planes_f: signed base-256 digits of max|D| / max|F|
// (force_digit_planes, 1..4); the entries must fit int32 (force bound).  ld % 4 == 0,
// ld > n.
int force_digit_planes(int64_t max_abs);
hipError_t launch_force_mfma(const int32_t *F, const int32_t *D, const int32_t *Ft,
                             const int32_t *Dt, const int32_t *perm, int n, int ld, int nrep,
                             int group, int planes_d, int planes_f, int32_t *Iout, hipStream_t s);
hipError_t launch_force(const int32_t *F, const int32_t *D, const int32_t *perm, int n, int ld,
                        int nrep, int32_t *Iout, hipStream_t s);
// Set-up costs of nrep replicas from their force matrices (E = -sum_i I[i][p(i)] / 2),
// once per group of replicas sharing a permutation.
hipError_t launch_energy_from_force(const int32_t *I, const int32_t *perm, int n, int ld, int nrep,
                                   int group, int64_t *Eout, hipStream_t s);
hipError_t launch_energy(const int32_t *F, const int32_t *D, const int32_t *perm, int n, int nrep,
                         int64_t *Eout, hipStream_t s);
// dst[k] = src[k] saturated to int16, k < count (the 16-bit copies of F and D that the lazy
// kernel reads; every entry of an instance placed there fits).
hipError_t launch_narrow16(const int32_t *src, int16_t *dst, int64_t count, hipStream_t s);
hipError_t launch_count_mismatch(const int32_t *a, const int32_t *b, int n, int ld,
                                 int64_t total_rows, unsigned long long *out, hipStream_t s);

}  // namespace qap
