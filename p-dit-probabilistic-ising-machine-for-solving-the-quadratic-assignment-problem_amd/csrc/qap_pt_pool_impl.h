// qap_pt_pool_impl.h -- the pool's host-side structure, shared by the batch entry points
// (qap_pt_pool.cpp) and the slot mode the serving loop runs on (qap_pt_slots.cpp).
// Private to the library.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <vector>

#include "../../include/qap_pt.h"
#include "qap_pt_launch.h"

namespace qap {

// Everything one instance needs on the device, built on the host (prepare_instance).
// Set-up force build of one instance: signed base-256 digits of max|D| and max|F| for the
// matrix cores, and how many consecutive replicas share each starting permutation.
struct ForceBuild {
    int planes_d = 1, planes_f = 1, group = 1;
};

struct HostInstance {
    InstDesc d{};
    SweepShape shape{};
    ForceBuild fb{};
    std::vector<int32_t> F, D, Ft, Dt, Fd, Dd, cls, span, perm, pinv, pos;
    std::vector<float> beta2f;
    std::vector<double> c20;
};

// Enqueues the set-up force matrices and costs of an instance's RT replicas (int32
// placements): the int8 matrix-core build and costs read off the force matrices, or with
// QAP_PT_FORCE_BUILD=valu the vector-ALU build and the O(n^2) cost sums.
hipError_t enqueue_force_build(const qap_pt_pool *p, const InstDesc &d, const ForceBuild &fb,
                               hipStream_t s);

struct Bucket {
    SweepShape shape;
    int lds_bytes = 0;
    std::vector<int32_t> reps;
    int32_t *d_list = nullptr;
};

}  // namespace qap

struct qap_pt_pool {
    int device = 0;
    int count = 0;
    int max_rt = 0;  // largest replicas x ladders of an instance (exchange staging)
    hipStream_t stream = nullptr;
    std::vector<qap_pt_params> params;
    std::vector<qap::InstDesc> desc;
    std::vector<qap::SweepShape> shapes;
    std::vector<qap::ForceBuild> fbuild;  // per instance (slot)
    std::vector<qap::Bucket> buckets;
    std::vector<int64_t> limits;  // mega-iteration limit per instance (host mirror)
    int total_reps = 0, total_ladders = 0;
    int64_t total_p = 0, total_I = 0, total_I64 = 0, total_mat = 0, total_vec = 0, total_beta = 0;
    qap::PoolDev P{};
    // owned device buffers
    qap::InstDesc *d_inst = nullptr;
    qap::InstState *d_state = nullptr;
    int32_t *d_rep_inst = nullptr, *d_F = nullptr, *d_D = nullptr, *d_Ft = nullptr, *d_Dt = nullptr;
    int16_t *d_F16 = nullptr, *d_D16 = nullptr;  // 16-bit copies of F and D
    int32_t *d_Fdiag = nullptr, *d_Ddiag = nullptr, *d_fclass = nullptr, *d_fspan = nullptr;
    float *d_beta2f = nullptr;
    double *d_c20 = nullptr;
    int32_t *d_perm = nullptr, *d_pinv = nullptr, *d_best_perm = nullptr, *d_I = nullptr;
    int64_t *d_I64 = nullptr;  // int64 force matrices of instances beyond the int32 bounds
    int64_t *d_E = nullptr, *d_bestE = nullptr, *d_limits = nullptr;
    uint64_t *d_accepted = nullptr, *d_cells = nullptr;
    int32_t *d_pos = nullptr, *d_order = nullptr, *d_tss = nullptr;
    uint32_t *d_sync = nullptr;  // persistent launches: fault word + per instance 2 words
    uint64_t *d_xslot = nullptr; // persistent launches: tagged cost words, 4 per replica
    // latency-bound pool of LDS placements: exchange intervals run as persistent
    // cooperative launches (replicas stay in LDS, exchange steps in the kernel)
    bool persistent = false;
    // latency-bound pool of one lazy bucket: every sweep launch brings `helpers` flush
    // helper workgroups per replica (cooperative launches, no graphs)
    int helpers = 0;
    int16_t *d_uv = nullptr;       // pending updates handed to the helpers
    uint32_t *d_fwords = nullptr;  // fault word + 3 words per replica block
    int64_t uv_stride = 0;
    int64_t *d_ladder_best = nullptr;
    std::vector<hipStream_t> side;     // one per extra bucket (default priority)
    hipEvent_t ev_fork = nullptr;
    std::vector<hipEvent_t> ev_join;
    // captured exchange intervals by depth: a run replays the deepest graph that fits and
    // then one graph of the remainder, so no replay carries idle intervals
    struct CapturedRun {
        int depth;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<CapturedRun> graphs;
    // timed runs: summed device time (ms) and count of each bucket's sweep launches
    std::vector<double> bucket_ms;
    std::vector<int64_t> bucket_launches;
    double seconds = 0.0;
    // slot mode: `count` fixed-capacity slots that instances are loaded into and released
    // from while the pool runs (qap_pt_slots.cpp)
    bool slotted = false;
    int n_cap = 0, r_cap = 0, t_cap = 0, rt_cap = 0;
    std::vector<uint8_t> used;
    std::vector<qap::HostInstance> staged;  // host sources of a slot's pending uploads
    std::vector<qap::InstState> st_load;         // per slot: the state uploaded by its load
    std::vector<std::vector<int64_t>> fill64;  // per slot: INT64_MAX / zero upload sources
    bool lists_dirty = false;
    // arena pool: every device buffer is a 256-byte-aligned piece of a region the library
    // does not allocate (arena != nullptr); such a pool runs on the caller's stream
    // (own_stream false), launches without graphs, events or side streams
    // (plain_launch), and frees nothing
    char *arena = nullptr;
    size_t arena_size = 0, arena_used = 0;
    bool own_stream = true;
    bool plain_launch = false;
};


namespace qap {

#define QAP_HIP_TRY(expr)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return ::qap::fail(e_ == hipErrorOutOfMemory ? QAP_PT_E_NOMEM : QAP_PT_E_DEVICE,    \
                               "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                               __LINE__);                                                       \
    } while (0)

template <typename T>
hipError_t dalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void **)p, count * sizeof(T));
}

// Device buffer of `count` elements for pool p: hipMalloc, or for an arena pool the next
// 256-byte-aligned piece of the arena.  An arena that is too small only advances
// arena_used past arena_size (the caller checks once, after the last piece, and reports
// the exact size needed) and hands out nullptr.
template <typename T>
hipError_t palloc(qap_pt_pool *p, T **ptr, size_t count) {
    if (!p->arena) return dalloc(ptr, count);
    *ptr = nullptr;
    if (count == 0) count = 1;
    const size_t off = (p->arena_used + 255) & ~(size_t)255;
    p->arena_used = off + count * sizeof(T);
    if (p->arena_used <= p->arena_size) *ptr = reinterpret_cast<T *>(p->arena + off);
    return hipSuccess;
}

// A pool whose buffers come from `arena` (size bytes, device memory the caller keeps) and
// whose work goes to `stream`; QAP_PT_E_NOMEM, before any device work, when the instances
// need more than `size`.
int pool_create_in_arena(int32_t count, const qap_problem *problems, const qap_pt_params *params,
                         void *arena, size_t size, hipStream_t stream, qap_pt_pool **out);

int validate(const qap_problem &pb, const qap_pt_params &pr, int k);
// throughput: the instance runs in a pool with more replicas than the device has CUs (the
// placement is then chosen by tuned full-chip rate rather than by per-replica latency)
void prepare_instance(const qap_problem &pb, const qap_pt_params &pr, bool throughput,
                      HostInstance &h);
// Row stride of a force matrix in HBM for size n.
int hbm_stride(int n);
int device_cus();
bool throughput_regime(int64_t replicas);
// Threads x candidates per thread for placement s.im at size n (tuned table, size rule,
// QAP_PT_SHAPE override); sets s.nt and s.cpl.
void shape_for_size(int n, bool sparse, SweepShape &s, bool latency = false);
// Orders a bucket's workgroups: larger n first, groups of 8 instances interleaved
// replica by replica (consecutive workgroup ids go round-robin over the 8 XCDs, so each
// instance stays on one XCD and its F and D stay in that XCD's L2).
void order_bucket(const qap_pt_pool *p, Bucket &B, const std::function<int(int32_t)> &inst_of);
// Fills the device-pointer table P from the pool's buffers.
void bind_pool(qap_pt_pool *p);
void pool_free(qap_pt_pool *p);

}  // namespace qap
