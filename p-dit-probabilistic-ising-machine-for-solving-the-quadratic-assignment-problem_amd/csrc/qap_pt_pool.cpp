// qap_pt_pool.cpp -- host side of the C-ABI (include/qap_pt.h): validation, device
// layout, set-up, hipGraph-captured sweep+exchange steps, results.
//
// Device layout (all instances of a pool concatenated, structure of arrays):
//   F, D, Ft, Dt      int32 n*n per instance (transposes feed the asymmetric path)
//   Fdiag, Ddiag      int32 n per instance
//   perm, pinv,
//   best_perm         int32 n per replica      (replica = instance-major, ladder, rung)
//   I                 int32 n*ld per replica   (ld = n rounded up to 4: 16-byte rows)
//   E, bestE          int64 per replica; accepted uint64; pos int32 (ladder position)
//   order             int32 per replica slot (position -> replica, per ladder)
//   tss, ladder_best  per ladder
//   beta2f, c20       per ladder position (float exp2-units beta, double 20/beta)
#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/qap_pt.h"
#include "qap_pt_flow.h"
#include "qap_pt_launch.h"
#include "qap_pt_internal.h"
#include "qap_pt_pool_impl.h"

using namespace qap;

namespace qap {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

}  // namespace qap

namespace {

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(QAP_PT_E_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                   \
    } while (0)

// Reference ln -> exp2 conversion constant (CPU_PIM.cpp:528, GPU_PIM.cu:792-793).
constexpr double kLog2e = 1.44269504089;

struct Bounds {
    int64_t force, delta, maxD, maxF;
};

Bounds compute_bounds(int n, const int32_t *F, const int32_t *D) {
    // |I[i][a]| <= min(maxF * max_i sum_j |D_ij|+|D_ji|, maxD * max_a sum_c |F_ac|+|F_ca|)
    int64_t maxD = 0, maxF = 0, sD = 0, sF = 0;
    for (int64_t x = 0; x < (int64_t)n * n; ++x) {
        maxD = std::max<int64_t>(maxD, std::llabs((long long)D[x]));
        maxF = std::max<int64_t>(maxF, std::llabs((long long)F[x]));
    }
    for (int i = 0; i < n; ++i) {
        int64_t rd = 0, rf = 0;
        for (int j = 0; j < n; ++j) {
            rd += std::llabs((long long)D[(int64_t)i * n + j]) + std::llabs((long long)D[(int64_t)j * n + i]);
            rf += std::llabs((long long)F[(int64_t)i * n + j]) + std::llabs((long long)F[(int64_t)j * n + i]);
        }
        sD = std::max(sD, rd);
        sF = std::max(sF, rf);
    }
    // products in 128 bits: with entries near 2^31 the bounds pass 2^63 (saturated at
    // INT64_MAX; such an instance is refused)
    const __int128 force = std::min((__int128)maxF * sD, (__int128)maxD * sF);
    const __int128 delta = 4 * force + 16 * (__int128)maxD * maxF;
    const __int128 top = INT64_MAX;
    Bounds b;
    b.force = (int64_t)std::min(force, top);
    b.delta = (int64_t)std::min(delta, top);
    b.maxD = maxD;
    b.maxF = maxF;
    return b;
}

// 0: force matrices, deltas and update terms provably fit int32; 1: they fit int64 only
// (4 B + 16 maxD maxF < 2^62 and the cost n^2 maxD maxF < 2^62); -1: neither.
int width_class(int n, const Bounds &b) {
    if (b.delta < ((int64_t)1 << 30) && b.force + 8 * b.maxD * b.maxF < ((int64_t)1 << 31)) return 0;
    if (b.delta < ((int64_t)1 << 62) &&
        (__int128)n * n * b.maxD * b.maxF < ((__int128)1 << 62))
        return 1;
    return -1;
}

bool is_sym_zero_diag(int n, const int32_t *M) {
    for (int i = 0; i < n; ++i) {
        if (M[(int64_t)i * n + i] != 0) return false;
        for (int j = i + 1; j < n; ++j)
            if (M[(int64_t)i * n + j] != M[(int64_t)j * n + i]) return false;
    }
    return true;
}

int round4i(int x) { return (x + 3) & ~3; }
}  // namespace

namespace qap {

// Row strides of the force matrix.  Every row has at least one padding column (index
// n): lanes without a column of their own in the rank-2 update read and rewrite it, so
// full row groups need no per-lane predication.  Both LDS copies have 16-byte rows (an
// odd number of 4-word quads for int32, of 8-column octets for uint16) so the rank-2
// update runs on 128-bit LDS accesses.
int hbm_stride(int n) { return round4i(n + 1); }

int device_cus() {
    // read once per process (thread-safe initialisation: pools may be created on several
    // host threads)
    static const int cus = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            return v;
        return 256;  // MI355X, for shape reports without a device
    }();
    return cus;
}

bool throughput_regime(int64_t replicas) { return replicas > device_cus(); }

// Threads x candidates per thread for placement s.im at size n: the tuned table's row for
// (n, flow class, placement), else a size rule; QAP_PT_SHAPE="nt,cpl" overrides both.
void shape_for_size(int n, bool sparse, SweepShape &s, bool latency) {
    if (n <= 64) {
        s.nt = 64;
        s.cpl = 1;
    } else if (n <= 128) {
        s.nt = 128;
        s.cpl = 1;
    } else if (n <= 256) {
        s.nt = 256;
        s.cpl = 1;
    } else if (n <= 512) {
        s.nt = 256;
        s.cpl = 2;
    } else {
        s.nt = 256;
        s.cpl = 4;
    }
    if (s.im == IM_LAZY && n > 256) {  // eight waves keep more of the flush's loads in flight
        s.nt = 512;
        s.cpl = n <= 512 ? 1 : 2;
    }
    auto take = [&](int nt, int cpl) {
        if (sweep_shape_supported(nt, cpl, s.im) && nt * cpl >= n) {
            s.nt = nt;
            s.cpl = cpl;
        }
    };
    // a latency-bound pool's own row first, else the full-chip row
    const TunedShape *t = latency ? tuned_shape(n, sparse, s.im, true) : nullptr;
    if (!t) t = tuned_shape(n, sparse, s.im, false);
    if (t) take(t->nt, t->cpl);
    if (const char *env = std::getenv("QAP_PT_SHAPE")) {
        int nt = 0, cpl = 0;
        if (std::sscanf(env, "%d,%d", &nt, &cpl) == 2) take(nt, cpl);
    }
}

int lds_stride(int n, int im) {
    if (im == IM_LDS32) {  // 16-byte rows for 128-bit updates; odd number of 4-word quads
        int q = (n + 1 + 3) / 4;
        if (q % 2 == 0) ++q;
        return 4 * q;
    }
    int q8 = (n + 1 + 7) / 8;  // uint16: 16-byte rows, an odd number of 8-column octets
    if (q8 % 2 == 0) ++q8;
    return 8 * q8;
}
}  // namespace qap
namespace {

// Threads x candidates per thread and force-matrix placement for size n.
// Placement: QAP_PT_IMODE=0|1|2|3 when that one fits (LDS budget, uint16 range, lazy
// conditions); else, for a throughput-bound pool, the fitting placement with the highest
// rate in the full-chip table among those it rates; else the first of lds32 -> lds16 ->
// lazy -> hbm that fits.  Shape: QAP_PT_SHAPE="nt,cpl" (tests and the autotuner), else
// the row for (n, flow class, placement) of the pool's regime table (latency-bound pools:
// the latency table, then the full-chip one), else a size rule.
SweepShape choose_shape(int n, bool sym, bool nonneg, int64_t force_bound, bool sparse,
                        bool small_uv, bool throughput, int *ld_s) {
    SweepShape s{};
    s.sym = sym;
    const bool fit32 = sweep_lds_bytes(n, lds_stride(n, IM_LDS32), IM_LDS32) <= kMaxLdsBytes;
    const bool fit16 = nonneg && force_bound <= 65535 &&
                       sweep_lds_bytes(n, lds_stride(n, IM_LDS16), IM_LDS16) <= kMaxLdsBytes;
    // lazy: symmetric, update vectors within 16 bits, room for 8 pending updates
    const bool lazy = sym && small_uv && lazy_fits(n);
    const char *im_env = std::getenv("QAP_PT_IMODE");
    const int want = im_env ? std::atoi(im_env) : -1;
    if (want == IM_HBM) s.im = IM_HBM;
    else if (want == IM_LDS16 && fit16) s.im = IM_LDS16;
    else if (want == IM_LAZY && lazy) s.im = IM_LAZY;
    else if (fit32) s.im = IM_LDS32;
    else if (fit16) s.im = IM_LDS16;
    else if (lazy) s.im = IM_LAZY;
    else s.im = IM_HBM;
    if (want < 0 && throughput) {
        // A pool with more replicas than CUs is throughput-bound: take the fastest rated
        // placement that fits (a smaller LDS footprint lets more replicas share a CU).  A
        // size bucket is rated at its largest n, so an LDS placement that fits only below
        // that size has no rate there; it drops out (it needs the most LDS, the slowest
        // at full load: n = 161..186 in an n <= 256 bank ran the int32 LDS kernel at a
        // third of the lazy kernel's rate).  A latency-bound pool keeps the order above
        // (force matrix in LDS whenever it fits): with a CU per replica the placements'
        // rates follow the acceptance rate and the flow structure more than n (the
        // latency table rates lazy above the LDS placements from n = 96 on its random
        // binary-flow instances, while the grey-pattern n = 256 instance runs 57 M
        // swap-attempts/s in uint16 LDS against 23 M lazy), so that table sets shapes only.
        const int cand[4] = {fit32 ? IM_LDS32 : -1, fit16 ? IM_LDS16 : -1, lazy ? IM_LAZY : -1,
                             IM_HBM};
        double best_rate = 0.0;
        int best_im = -1;
        for (int im : cand) {
            if (im < 0) continue;
            const TunedShape *t = tuned_shape(n, sparse, im);
            if (t && t->rate > best_rate) {
                best_rate = t->rate;
                best_im = im;
            }
        }
        if (best_im >= 0) s.im = best_im;
    }
    *ld_s = lds_stride(n, s.im);
    shape_for_size(n, sparse, s, !throughput);
    // A latency-bound pool of symmetric instances that fit one wave in int32 LDS evaluates
    // several sites at once, a wave each (sweep_spec_kernel): a round costs about what one
    // site does, and advances to the first site whose update moved a facility, so mostly
    // rejected updates (cold replicas) run up to that many times faster.
    // QAP_PT_SPEC=0|2|4|8|16 sets the sites per round (0: one at a time).
    if (!throughput && s.im == IM_LDS32 && sym && n <= 64) {
        int sp = kSpecSites;
        if (const char *e = std::getenv("QAP_PT_SPEC")) sp = std::atoi(e);
        s.spec = (sp == 2 || sp == 4 || sp == 8 || sp == 16) ? sp : 0;
    }
    return s;
}

const char *placement_name(int im) {
    switch (im) {
    case IM_LDS32: return "lds32";
    case IM_LDS16: return "lds16";
    case IM_LAZY: return "lazy";
    case IM_ONESHOT: return "oneshot";
    case IM_HBM64: return "hbm64";
    default: return "hbm";
    }
}

// Initial permutation (CPU_PIM.cpp:134-161): identity, then for i = 0..n-1 swap
// positions i and j, j uniform in [0,n) from Philox ctr = {i, stream, 0, tag}.
void init_perm(int n, uint64_t seed, uint32_t stream, uint32_t tag, int32_t *perm) {
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = 0; i < n; ++i) {
        const u32x4 x = philox4x32_10(u32x4{(uint32_t)i, stream, 0u, tag}, (uint32_t)seed,
                                      (uint32_t)(seed >> 32));
        const int j = (int)(((uint64_t)x.x * (uint64_t)n) >> 32);
        if (i != j) std::swap(perm[i], perm[j]);
    }
}


}  // namespace

namespace qap {

int validate(const qap_problem &pb, const qap_pt_params &pr, int k) {
    if (pb.n < 1 || pb.n > QAP_PT_MAX_N)
        return fail(QAP_PT_E_INVAL, "instance %d: n=%d outside [1,%d]", k, pb.n, QAP_PT_MAX_N);
    if (!pb.F || !pb.D) return fail(QAP_PT_E_INVAL, "instance %d: F or D is NULL", k);
    if (pr.R < 1 || pr.T < 1) return fail(QAP_PT_E_INVAL, "instance %d: R and T must be >= 1", k);
    if ((int64_t)pr.R * pr.T > (1 << 24))
        return fail(QAP_PT_E_INVAL, "instance %d: R*T too large", k);
    if (!pr.beta_ln) return fail(QAP_PT_E_INVAL, "instance %d: beta_ln is NULL", k);
    for (int l = 0; l < pr.R; ++l) {
        if (!(pr.beta_ln[l] > 0.0) || !(pr.beta_ln[l] < 1e30))
            return fail(QAP_PT_E_INVAL, "instance %d: beta_ln[%d] must be finite and > 0", k, l);
        if (l && pr.beta_ln[l] < pr.beta_ln[l - 1])
            return fail(QAP_PT_E_INVAL, "instance %d: beta_ln must be ascending", k);
    }
    if (pr.sweeps_per_exchange < 1)
        return fail(QAP_PT_E_INVAL, "instance %d: sweeps_per_exchange must be >= 1", k);
    if (pr.sweeps < 0 || pr.sweeps % pr.sweeps_per_exchange)
        return fail(QAP_PT_E_INVAL, "instance %d: sweeps must be >= 0 and a multiple of "
                                    "sweeps_per_exchange", k);
    if (pr.sweeps > (int64_t)0xffffffffLL)
        return fail(QAP_PT_E_INVAL, "instance %d: sweeps must be < 2^32", k);
    if (pr.kick_after < 0) return fail(QAP_PT_E_INVAL, "instance %d: kick_after < 0", k);
    if (pr.timeout_s < 0) return fail(QAP_PT_E_INVAL, "instance %d: timeout_s < 0", k);
    if (pr.init_perm) {
        std::vector<char> seen(pb.n, 0);
        for (int i = 0; i < pb.n; ++i) {
            const int v = pr.init_perm[i];
            if (v < 0 || v >= pb.n || seen[v])
                return fail(QAP_PT_E_INVAL, "instance %d: init_perm is not a permutation", k);
            seen[v] = 1;
        }
    }
    const Bounds b = compute_bounds(pb.n, pb.F, pb.D);
    if (pr.flags & ~(QAP_PT_INIT_PER_REPLICA | QAP_PT_ONE_SHOT))
        return fail(QAP_PT_E_INVAL, "instance %d: unknown flags 0x%x", k, pr.flags);
    if (pr.flags & QAP_PT_ONE_SHOT) {
        // the one-shot kernel gives every location a lane (n <= 256 threads) and splits the
        // 16-bit update vectors into two signed bytes for the matrix cores
        if (pb.n > 256 || !oneshot_fits(pb.n))
            return fail(QAP_PT_E_INVAL, "instance %d: one-shot steps need n <= 256 (n=%d)", k,
                        pb.n);
        if (2 * b.maxD > 32639 || 4 * b.maxF > 32639)
            return fail(QAP_PT_E_INVAL,
                        "instance %d: one-shot steps need 2 max|D| and 4 max|F| <= 32639", k);
    }
    const int width = width_class(pb.n, b);
    if (width < 0)
        return fail(QAP_PT_E_OVERFLOW,
                    "instance %d: |force| bound %lld / |delta| bound %lld exceed the int64 range "
                    "the sweep kernels prove exact",
                    k, (long long)b.force, (long long)b.delta);
    if (width > 0 && (pr.flags & QAP_PT_ONE_SHOT))
        return fail(QAP_PT_E_OVERFLOW,
                    "instance %d: one-shot steps need force bounds within int32 (|delta| bound "
                    "%lld)", k, (long long)b.delta);
    return QAP_PT_OK;
}


void pool_free(qap_pt_pool *p) {
    if (!p) return;
    (void)hipSetDevice(p->device);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    for (auto &g : p->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    if (p->arena) {  // nothing here was allocated by the pool
        delete p;
        return;
    }
    for (auto &b : p->buckets)
        if (b.d_list) (void)hipFree(b.d_list);
    void *bufs[] = {p->d_inst,  p->d_state,     p->d_rep_inst, p->d_F,      p->d_D, p->d_F16, p->d_D16,
                    p->d_fclass, p->d_fspan,
                    p->d_Ft,    p->d_Dt,        p->d_Fdiag,    p->d_Ddiag,  p->d_beta2f,
                    p->d_c20,   p->d_perm,      p->d_pinv,     p->d_best_perm, p->d_I, p->d_I64,
                    p->d_E,     p->d_bestE,     p->d_limits,   p->d_accepted, p->d_cells, p->d_pos,
                    p->d_order, p->d_tss,       p->d_ladder_best, p->d_sync, p->d_xslot,
                    p->d_uv,    p->d_fwords};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (auto st : p->side) (void)hipStreamDestroy(st);
    for (auto ev : p->ev_join) (void)hipEventDestroy(ev);
    if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    if (p->stream && p->own_stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

}  // namespace qap
namespace {


// Side streams and fork/join events for buckets beyond the first (created once per pool).
int ensure_side_streams(qap_pt_pool *p) {
    const int nb = (int)p->buckets.size();
    while ((int)p->side.size() < nb - 1) {
        hipStream_t st = nullptr;
        HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        p->side.push_back(st);
        hipEvent_t ev = nullptr;
        HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        p->ev_join.push_back(ev);
    }
    if (nb > 1 && !p->ev_fork) HIP_TRY(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
    return QAP_PT_OK;
}

// One exchange interval: exchange, then every sweep bucket.  Buckets (different kernel
// instantiations) run concurrently on side streams that fork from and join back into the
// pool stream.  With `ev` (timed runs) each bucket's launch is bracketed by
// ev[2 b], ev[2 b + 1] recorded on the stream it is launched on.
hipError_t enqueue_interval(qap_pt_pool *p, const hipEvent_t *ev) {
    const int nb = (int)p->buckets.size();
    hipError_t e = launch_exchange(p->P, p->count, p->max_rt, p->stream);
    if (e == hipSuccess && nb > 1) e = hipEventRecord(p->ev_fork, p->stream);
    auto sweep = [&](int b, hipStream_t st) {
        const Bucket &B = p->buckets[b];
        if (e == hipSuccess && ev) e = hipEventRecord(ev[2 * b], st);
        if (e == hipSuccess)
            e = launch_sweep(p->P, B.shape, B.d_list, (int)B.reps.size(), B.lds_bytes, st);
        if (e == hipSuccess && ev) e = hipEventRecord(ev[2 * b + 1], st);
    };
    for (int b = 1; b < nb && e == hipSuccess; ++b) {
        e = hipStreamWaitEvent(p->side[b - 1], p->ev_fork, 0);
        sweep(b, p->side[b - 1]);
        if (e == hipSuccess) e = hipEventRecord(p->ev_join[b - 1], p->side[b - 1]);
    }
    if (nb > 0) sweep(0, p->stream);
    for (int b = 1; b < nb && e == hipSuccess; ++b)
        e = hipStreamWaitEvent(p->stream, p->ev_join[b - 1], 0);
    return e;
}

// The executable graph of `depth` exchange intervals, captured on first use.  At most four
// depths are kept (a pool runs one or two: its step and a remainder).
int graph_for(qap_pt_pool *p, int depth, hipGraphExec_t *out) {
    for (auto &g : p->graphs)
        if (g.depth == depth) {
            *out = g.exec;
            return QAP_PT_OK;
        }
    if (p->graphs.size() >= 4) {
        (void)hipStreamSynchronize(p->stream);
        (void)hipGraphExecDestroy(p->graphs.front().exec);
        (void)hipGraphDestroy(p->graphs.front().graph);
        p->graphs.erase(p->graphs.begin());
    }
    const int rc = ensure_side_streams(p);
    if (rc) return rc;
    HIP_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int m = 0; m < depth && e == hipSuccess; ++m) e = enqueue_interval(p, nullptr);
    if (e != hipSuccess) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(p->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        return fail(QAP_PT_E_DEVICE, "kernel launch during capture failed: %s",
                    hipGetErrorString(e));
    }
    qap_pt_pool::CapturedRun g{depth, nullptr, nullptr};
    HIP_TRY(hipStreamEndCapture(p->stream, &g.graph));
    e = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g.graph);
        return fail(QAP_PT_E_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    p->graphs.push_back(g);
    *out = g.exec;
    return QAP_PT_OK;
}

// Enqueues `intervals` exchange intervals on the pool stream: replays of the deepest graph
// (up to kGraphDepth intervals) and one of the remainder, or, for an arena pool, the same
// launches one by one.
constexpr int kGraphDepth = 32;
// exchange intervals per persistent launch at most (bounds one launch's duration)
constexpr int kPersistMaxIntervals = 4096;

// One persistent launch per bucket for `intervals` intervals, after the exchange step of
// the first one.  Returns hipErrorCooperativeLaunchTooLarge (nothing launched) when a
// bucket's workgroups cannot all be resident.
hipError_t enqueue_persistent(qap_pt_pool *p, int intervals, const hipEvent_t *ev) {
    hipError_t e = launch_exchange(p->P, p->count, p->max_rt, p->stream);
    for (size_t b = 0; b < p->buckets.size() && e == hipSuccess; ++b) {
        const Bucket &B = p->buckets[b];
        if (ev) e = hipEventRecord(ev[2 * b], p->stream);
        if (e == hipSuccess)
            e = launch_sweep_persistent(p->P, B.shape, B.d_list, (int)B.reps.size(), B.lds_bytes,
                                        intervals, p->stream);
        if (e == hipSuccess && ev) e = hipEventRecord(ev[2 * b + 1], p->stream);
    }
    return e;
}

// One exchange interval of a helped lazy bucket: exchange step, cleared flush words, the
// cooperative sweep launch with its helpers (events around it when `ev`).
hipError_t enqueue_helped(qap_pt_pool *p, const hipEvent_t *ev) {
    const Bucket &B = p->buckets[0];
    FlushHelp fh;
    fh.uv = p->d_uv;
    fh.words = p->d_fwords;
    fh.uv_stride = p->uv_stride;
    fh.nrep = (int)B.reps.size();
    fh.helpers = p->helpers;
    hipError_t e = launch_exchange(p->P, p->count, p->max_rt, p->stream);
    if (e == hipSuccess)
        e = hipMemsetAsync(p->d_fwords + 1, 0, sizeof(uint32_t) * 3 * (size_t)fh.nrep, p->stream);
    if (e == hipSuccess && ev) e = hipEventRecord(ev[0], p->stream);
    if (e == hipSuccess) e = launch_sweep_lazy_helped(p->P, B.shape, B.d_list, B.lds_bytes, fh, p->stream);
    if (e == hipSuccess && ev) e = hipEventRecord(ev[1], p->stream);
    return e;
}

int enqueue_intervals(qap_pt_pool *p, int64_t intervals) {
    if (p->helpers > 0) {
        for (int64_t m = 0; m < intervals; ++m) {
            const hipError_t e = enqueue_helped(p, nullptr);
            if (e != hipSuccess)
                return fail(QAP_PT_E_DEVICE, "helped lazy sweep launch failed: %s",
                            hipGetErrorString(e));
        }
        return QAP_PT_OK;
    }
    if (p->persistent) {
        for (int64_t m = 0; m < intervals; m += kPersistMaxIntervals) {
            const int n = (int)std::min<int64_t>(intervals - m, kPersistMaxIntervals);
            const hipError_t e = enqueue_persistent(p, n, nullptr);
            if (e != hipSuccess)
                return fail(QAP_PT_E_DEVICE, "persistent sweep launch failed: %s",
                            hipGetErrorString(e));
        }
        return QAP_PT_OK;
    }
    if (p->plain_launch) {
        // no side streams, events or graph objects: every bucket in turn on the pool stream
        for (int64_t x = 0; x < intervals; ++x) {
            HIP_TRY(launch_exchange(p->P, p->count, p->max_rt, p->stream));
            for (auto &b : p->buckets)
                HIP_TRY(launch_sweep(p->P, b.shape, b.d_list, (int)b.reps.size(), b.lds_bytes,
                                     p->stream));
        }
        return QAP_PT_OK;
    }
    for (int64_t m = 0; m < intervals;) {
        const int depth = (int)std::min<int64_t>(kGraphDepth, intervals - m);
        hipGraphExec_t ex = nullptr;
        const int rc = graph_for(p, depth, &ex);
        if (rc) return rc;
        HIP_TRY(hipGraphLaunch(ex, p->stream));
        m += depth;
    }
    return QAP_PT_OK;
}

}  // namespace

namespace qap {

// Everything one instance needs on the device, built on the host: descriptor (offsets
// left to the caller), kernel shape, matrices and their transposes, diagonals, facility
// classes and flow spans, per-rung betas, and the initial permutations of every replica
// (CPU_PIM.cpp:109-162).  Ladder positions start as the identity (order == pos).
void prepare_instance(const qap_problem &pb, const qap_pt_params &pr, bool throughput,
                      HostInstance &h) {
    const int n = pb.n;
    const int64_t nn = (int64_t)n * n;
    InstDesc &d = h.d;
    const bool sym = is_sym_zero_diag(n, pb.F) && is_sym_zero_diag(n, pb.D);
    bool nonneg = true;
    for (int64_t x = 0; x < nn && nonneg; ++x) nonneg = pb.F[x] >= 0 && pb.D[x] >= 0;
    const Bounds bnd = compute_bounds(n, pb.F, pb.D);
    int ld_s = 0;
    const bool sparse = offdiag_density(n, pb.F) < kSparseFlow;
    // lazy updates hold u (|u| <= 2 max|D|) and v (|v| <= 4 max|F|) in 16 bits, split into
    // two signed bytes for the matrix cores (the high byte of 32640..32767 would not fit)
    const bool small_uv = 2 * bnd.maxD <= 32639 && 4 * bnd.maxF <= 32639;
    if (pr.flags & QAP_PT_ONE_SHOT) {
        // a location per thread, a candidate slot per lane (n <= 64 cpl), force matrix in
        // HBM; eight waves (two per SIMD) hide the proposal loads and scans, 1.2-1.6x
        // over four at n = 64 .. 256.  QAP_PT_SHAPE="nt,cpl" (nt 256 or 512) for tests
        h.shape = SweepShape{512, n <= 64 ? 1 : (n <= 128 ? 2 : 4), IM_ONESHOT, sym};
        if (const char *env = std::getenv("QAP_PT_SHAPE")) {
            int nt = 0, cpl = 0;
            if (std::sscanf(env, "%d,%d", &nt, &cpl) == 2 && sweep_shape_supported(nt, cpl, IM_ONESHOT) &&
                nt >= n && 64 * cpl >= n) {
                h.shape.nt = nt;
                h.shape.cpl = cpl;
            }
        }
    } else if (width_class(n, bnd) > 0) {
        // beyond int32: int64 force matrix in HBM, updated in place (sizes by the size rule)
        h.shape = SweepShape{};
        h.shape.im = IM_HBM64;
        h.shape.sym = sym;
        shape_for_size(n, sparse, h.shape);
    } else {
        h.shape = choose_shape(n, sym, nonneg, bnd.force, sparse, small_uv, throughput, &ld_s);
    }
    d = InstDesc{};
    d.n = n;
    d.ld = hbm_stride(n);
    d.ld_s = ld_s;
    d.sparse_upd = sparse ? 1 : 0;
    // u = distance differences (<= 2 max|D|), v = flow differences (<= 4 max|F|)
    d.mul24 = (2 * bnd.maxD < (1 << 23) && 4 * bnd.maxF < (1 << 23)) ? 1 : 0;
    d.R = pr.R;
    d.T = pr.T;
    d.spe = pr.sweeps_per_exchange;
    d.kick_after = pr.kick_after;
    d.has_goal = pr.has_goal ? 1 : 0;
    d.sym = sym ? 1 : 0;
    // Persistent launches run the exchange step in every workgroup from 32-bit tagged costs
    // when a ladder set fits one wave and every cost fits: |E| <= max|F| sum|D|.
    // QAP_PT_XMIRROR=0 keeps the counter protocol (one workgroup runs each step).
    {
        __int128 sumD = 0;
        for (int64_t x = 0; x < nn; ++x) sumD += std::llabs((long long)pb.D[x]);
        const bool fits = (__int128)bnd.maxF * sumD < (__int128)INT32_MAX;
        const char *env = std::getenv("QAP_PT_XMIRROR");
        d.xmirror = (pr.R * pr.T <= 64 && fits && !(env && env[0] == '0')) ? 1 : 0;
    }
    d.goal = pr.goal;
    d.seed_lo = (uint32_t)pr.seed;
    d.seed_hi = (uint32_t)(pr.seed >> 32);

    h.F.assign(pb.F, pb.F + nn);
    h.D.assign(pb.D, pb.D + nn);
    h.Ft.resize(nn);
    h.Dt.resize(nn);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            h.Ft[(int64_t)i * n + j] = pb.F[(int64_t)j * n + i];
            h.Dt[(int64_t)i * n + j] = pb.D[(int64_t)j * n + i];
        }
    h.Fd.resize(n);
    h.Dd.resize(n);
    h.span.resize(n);
    for (int i = 0; i < n; ++i) {
        h.Fd[i] = pb.F[(int64_t)i * n + i];
        h.Dd[i] = pb.D[(int64_t)i * n + i];
        h.span[i] = facility_span(n, pb.F, i);
    }
    // class id (identical flow row and column) in the low 16 bits, twin class id (flow
    // rows equal outside the two swapped columns; symmetric instances) in the high 16
    h.cls = facility_classes(n, pb.F);
    const std::vector<int32_t> twin = facility_twins(n, pb.F, sym);
    for (int i = 0; i < n; ++i) h.cls[i] |= twin[i] << 16;
    for (int l = 0; l < pr.R; ++l) {
        const double b2 = pr.beta_ln[l] * kLog2e;
        h.beta2f.push_back((float)b2);
        h.c20.push_back(20.0 / b2);
    }
    // replicas sharing a start: all of them with init_perm, each alone with per-replica
    // starts, else the R replicas of a ladder
    h.fb.planes_d = force_digit_planes(bnd.maxD);
    h.fb.planes_f = force_digit_planes(bnd.maxF);
    h.fb.group = pr.init_perm ? pr.R * pr.T : ((pr.flags & QAP_PT_INIT_PER_REPLICA) ? 1 : pr.R);
    std::vector<int32_t> pi(n), inv(n);
    for (int t = 0; t < pr.T; ++t) {
        if (!(pr.flags & QAP_PT_INIT_PER_REPLICA)) {
            if (pr.init_perm) std::copy(pr.init_perm, pr.init_perm + n, pi.begin());
            else init_perm(n, pr.seed, (uint32_t)t, 1u, pi.data());
        }
        for (int l = 0; l < pr.R; ++l) {
            if (pr.flags & QAP_PT_INIT_PER_REPLICA) {
                if (pr.init_perm) std::copy(pr.init_perm, pr.init_perm + n, pi.begin());
                else init_perm(n, pr.seed, (uint32_t)(t * pr.R + l), 2u, pi.data());
            }
            h.perm.insert(h.perm.end(), pi.begin(), pi.end());
            for (int i = 0; i < n; ++i) inv[pi[i]] = i;
            h.pinv.insert(h.pinv.end(), inv.begin(), inv.end());
            h.pos.push_back(l);
        }
    }
}

hipError_t enqueue_force_build(const qap_pt_pool *p, const InstDesc &d, const ForceBuild &fb,
                               hipStream_t s) {
    const int RT = d.R * d.T;
    const char *env = std::getenv("QAP_PT_FORCE_BUILD");
    if (env && std::strcmp(env, "valu") == 0) {
        const hipError_t e = launch_force(p->d_F + d.mat_off, p->d_D + d.mat_off,
                                          p->d_perm + d.p_base, d.n, d.ld, RT, p->d_I + d.I_base, s);
        if (e != hipSuccess) return e;
        return launch_energy(p->d_F + d.mat_off, p->d_D + d.mat_off, p->d_perm + d.p_base, d.n, RT,
                             p->d_E + d.rep_base, s);
    }
    const hipError_t e = launch_force_mfma(p->d_F + d.mat_off, p->d_D + d.mat_off,
                                           p->d_Ft + d.mat_off, p->d_Dt + d.mat_off,
                                           p->d_perm + d.p_base, d.n, d.ld, RT, fb.group,
                                           fb.planes_d, fb.planes_f, p->d_I + d.I_base, s);
    if (e != hipSuccess) return e;
    return launch_energy_from_force(p->d_I + d.I_base, p->d_perm + d.p_base, d.n, d.ld, RT,
                                    fb.group, p->d_E + d.rep_base, s);
}

void order_bucket(const qap_pt_pool *p, Bucket &B, const std::function<int(int32_t)> &inst_of) {
    std::map<int, std::vector<int32_t>> by_inst;
    for (int32_t g : B.reps) by_inst[inst_of(g)].push_back(g);
    std::vector<int> insts;
    for (auto &kv : by_inst) insts.push_back(kv.first);
    std::stable_sort(insts.begin(), insts.end(),
                     [&](int x, int y) { return p->desc[x].n > p->desc[y].n; });
    std::vector<int32_t> order_list;
    for (size_t g0 = 0; g0 < insts.size(); g0 += 8) {
        const size_t g1 = std::min(insts.size(), g0 + 8);
        size_t longest = 0;
        for (size_t x = g0; x < g1; ++x) longest = std::max(longest, by_inst[insts[x]].size());
        for (size_t r = 0; r < longest; ++r)
            for (size_t x = g0; x < g1; ++x) {
                const auto &v = by_inst[insts[x]];
                if (r < v.size()) order_list.push_back(v[r]);
            }
    }
    B.reps.swap(order_list);
}

void bind_pool(qap_pt_pool *p) {
    PoolDev &P = p->P;
    P.inst = p->d_inst;
    P.state = p->d_state;
    P.rep_inst = p->d_rep_inst;
    P.F = p->d_F;
    P.D = p->d_D;
    P.F16 = p->d_F16;
    P.D16 = p->d_D16;
    P.Ft = p->d_Ft;
    P.Dt = p->d_Dt;
    P.Fdiag = p->d_Fdiag;
    P.Ddiag = p->d_Ddiag;
    P.fclass = p->d_fclass;
    P.fspan = p->d_fspan;
    P.beta2f = p->d_beta2f;
    P.c20 = p->d_c20;
    P.perm = p->d_perm;
    P.pinv = p->d_pinv;
    P.best_perm = p->d_best_perm;
    P.I = p->d_I;
    P.I64 = p->d_I64;
    P.E = p->d_E;
    P.bestE = p->d_bestE;
    P.accepted = p->d_accepted;
    P.cells = p->d_cells;
    P.fcells = p->d_cells + p->total_reps;
    P.pos = p->d_pos;
    P.order = p->d_order;
    P.tss = p->d_tss;
    P.ladder_best = p->d_ladder_best;
    P.sync = p->d_sync;
    P.xslot = p->d_xslot;
}

}  // namespace qap
namespace {


struct ArenaSpec {
    char *base;
    size_t size;
    hipStream_t stream;
};

int pool_create_impl(int32_t count, const qap_problem *problems, const qap_pt_params *params,
                     qap_pt_pool **out, const ArenaSpec *arena = nullptr) {
    if (!out) return fail(QAP_PT_E_INVAL, "out is NULL");
    *out = nullptr;
    if (count < 1 || !problems || !params) return fail(QAP_PT_E_INVAL, "empty instance list");
    for (int k = 0; k < count; ++k) {
        const int rc = validate(problems[k], params[k], k);
        if (rc) return rc;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(QAP_PT_E_DEVICE, "no HIP device visible");

    qap_pt_pool *p = new qap_pt_pool();
    (void)hipGetDevice(&p->device);
    if (arena) {
        p->arena = arena->base;
        p->arena_size = arena->size;
        p->own_stream = false;
        p->plain_launch = true;
        p->stream = arena->stream;
    }
    p->count = count;
    p->params.assign(params, params + count);
    p->desc.resize(count);
    p->shapes.resize(count);
    p->fbuild.resize(count);
    p->limits.assign(count, 0);

    std::vector<int32_t> hF, hD, hFt, hDt, hFd, hDd, hCls, hSpan, rep_inst, perm, pinv, pos, order;
    std::vector<float> beta2f;
    std::vector<double> c20;
    std::map<std::tuple<int, int, int, int, int>, int> bucket_of;
    auto append = [](auto &dst, const auto &src) { dst.insert(dst.end(), src.begin(), src.end()); };
    int64_t all_reps = 0;
    for (int k = 0; k < count; ++k) all_reps += (int64_t)params[k].R * params[k].T;
    const bool throughput = throughput_regime(all_reps);
    for (int k = 0; k < count; ++k) {
        HostInstance h;
        prepare_instance(problems[k], params[k], throughput, h);
        InstDesc &d = p->desc[k];
        d = h.d;
        p->shapes[k] = h.shape;
        p->fbuild[k] = h.fb;
        const int n = d.n, RT = d.R * d.T;
        d.mat_off = p->total_mat;
        d.vec_off = p->total_vec;
        d.p_base = p->total_p;
        d.I_base = p->total_I;
        d.rep_base = p->total_reps;
        d.ladder_base = p->total_ladders;
        d.beta_off = (int32_t)p->total_beta;
        p->total_mat += (int64_t)n * n;
        p->total_vec += n;
        p->total_p += (int64_t)RT * n;
        if (h.shape.im == IM_HBM64) {  // int64 force matrices live apart
            d.I64_base = p->total_I64;
            p->total_I64 += (int64_t)RT * n * d.ld;
        } else {
            p->total_I += (int64_t)RT * n * d.ld;
        }
        p->total_reps += RT;
        p->total_ladders += d.T;
        p->total_beta += d.R;
        p->max_rt = std::max(p->max_rt, RT);
        append(hF, h.F);
        append(hD, h.D);
        append(hFt, h.Ft);
        append(hDt, h.Dt);
        append(hFd, h.Fd);
        append(hDd, h.Dd);
        append(hCls, h.cls);
        append(hSpan, h.span);
        append(beta2f, h.beta2f);
        append(c20, h.c20);
        append(perm, h.perm);
        append(pinv, h.pinv);
        append(pos, h.pos);
        append(order, h.pos);
        for (int r = 0; r < RT; ++r) rep_inst.push_back(k);
        const SweepShape &sh = h.shape;
        const auto key = std::make_tuple(sh.nt, sh.cpl, sh.im, (int)sh.sym, sh.spec);
        auto it = bucket_of.find(key);
        int bi;
        if (it == bucket_of.end()) {
            bi = (int)p->buckets.size();
            bucket_of[key] = bi;
            p->buckets.emplace_back();
            p->buckets[bi].shape = sh;
            p->buckets[bi].shape.sparse = true;
        } else {
            bi = it->second;
        }
        Bucket &B = p->buckets[bi];
        B.shape.sparse = B.shape.sparse && d.sparse_upd != 0;  // all-sparse buckets only
        B.lds_bytes = std::max(B.lds_bytes, shape_lds_bytes(n, d.ld_s, sh));
        for (int r = 0; r < RT; ++r) B.reps.push_back(d.rep_base + r);
    }

    for (auto &B : p->buckets) order_bucket(p, B, [&](int32_t g) { return rep_inst[g]; });
    // Persistent launches: a latency-bound pool (a CU per replica) whose force matrices all
    // sit in LDS keeps them there across exchange intervals, the exchange steps running in
    // the kernel; QAP_PT_PERSIST=0 runs every interval as its own launches instead.
    {
        const char *env = std::getenv("QAP_PT_PERSIST");
        p->persistent = !arena && !throughput && !(env && env[0] == '0') &&
                        p->buckets.size() == 1 &&
                        sweep_persistent_fits(p->buckets[0].shape, (int)p->buckets[0].reps.size(),
                                              p->buckets[0].lds_bytes);
    }
    // Flush helpers: a latency-bound pool whose replicas all run the lazy kernel leaves
    // CUs idle; helper workgroups can take shares of every flush.  Handing rows between
    // workgroups costs agent-scope fences, which write back or invalidate a whole XCD's L2
    // (shared with the other replicas there), so they pay only where flushes are largest:
    // one helper per replica from n = 768 (n = 1024: 13.6 -> 12.3 ms per sweep launch;
    // n = 512 lost 17 % with one helper and 37 % with three).  QAP_PT_FLUSH_HELPERS=k sets
    // k (0..3) for any lazy latency-bound pool.
    if (!arena && !throughput && p->buckets.size() == 1 && p->buckets[0].shape.im == IM_LAZY) {
        const Bucket &B = p->buckets[0];
        const int nb = (int)B.reps.size();
        int n_max = 0;
        for (int k = 0; k < count; ++k) n_max = std::max(n_max, p->desc[k].n);
        int h = n_max >= 768 ? 1 : 0;
        if (const char *env = std::getenv("QAP_PT_FLUSH_HELPERS")) h = std::max(0, std::atoi(env));
        h = std::min({h, 3, device_cus() / std::max(nb, 1) - 1});
        while (h > 0 && !lazy_helpers_fit(B.shape, nb, h, B.lds_bytes)) --h;
        if (h > 0) {
            int64_t stride = 0;
            for (int k = 0; k < count; ++k)
                stride = std::max(stride, lazy_pending_elems(p->desc[k].n, B.lds_bytes));
            p->uv_stride = (stride + 7) & ~(int64_t)7;  // 16-byte aligned per replica
            p->helpers = h;
        }
    }

    auto bail = [&](int rc) {
        pool_free(p);
        return rc;
    };
#define POOL_TRY(expr)                                                                           \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return bail(fail(e_ == hipErrorOutOfMemory ? QAP_PT_E_NOMEM : QAP_PT_E_DEVICE,       \
                             "%s failed: %s", #expr, hipGetErrorString(e_)));                    \
    } while (0)

    if (p->own_stream) POOL_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    const int RTall = p->total_reps;
    POOL_TRY(palloc(p, &p->d_inst, count));
    POOL_TRY(palloc(p, &p->d_state, count));
    POOL_TRY(palloc(p, &p->d_limits, count));
    POOL_TRY(palloc(p, &p->d_rep_inst, RTall));
    POOL_TRY(palloc(p, &p->d_F, p->total_mat));
    POOL_TRY(palloc(p, &p->d_D, p->total_mat));
    POOL_TRY(palloc(p, &p->d_F16, p->total_mat));
    POOL_TRY(palloc(p, &p->d_D16, p->total_mat));
    POOL_TRY(palloc(p, &p->d_Ft, p->total_mat));
    POOL_TRY(palloc(p, &p->d_Dt, p->total_mat));
    POOL_TRY(palloc(p, &p->d_Fdiag, p->total_vec));
    POOL_TRY(palloc(p, &p->d_Ddiag, p->total_vec));
    POOL_TRY(palloc(p, &p->d_fclass, p->total_vec));
    POOL_TRY(palloc(p, &p->d_fspan, p->total_vec));
    POOL_TRY(palloc(p, &p->d_beta2f, p->total_beta));
    POOL_TRY(palloc(p, &p->d_c20, p->total_beta));
    POOL_TRY(palloc(p, &p->d_perm, p->total_p));
    POOL_TRY(palloc(p, &p->d_pinv, p->total_p));
    POOL_TRY(palloc(p, &p->d_best_perm, p->total_p));
    POOL_TRY(palloc(p, &p->d_I, p->total_I));
    POOL_TRY(palloc(p, &p->d_I64, p->total_I64));
    POOL_TRY(palloc(p, &p->d_E, RTall));
    POOL_TRY(palloc(p, &p->d_bestE, RTall));
    POOL_TRY(palloc(p, &p->d_accepted, RTall));
    POOL_TRY(palloc(p, &p->d_cells, 2 * (size_t)RTall));  // update and flush counts
    POOL_TRY(palloc(p, &p->d_pos, RTall));
    POOL_TRY(palloc(p, &p->d_order, RTall));
    POOL_TRY(palloc(p, &p->d_tss, p->total_ladders));
    POOL_TRY(palloc(p, &p->d_ladder_best, p->total_ladders));
    POOL_TRY(palloc(p, &p->d_sync, 1 + 2 * (size_t)count));
    POOL_TRY(palloc(p, &p->d_xslot, 4 * (size_t)RTall));
    for (auto &b : p->buckets) POOL_TRY(palloc(p, &b.d_list, b.reps.size()));
    if (p->helpers > 0) {  // pending updates and flush words of the helped lazy bucket
        const size_t nb = p->buckets[0].reps.size();
        POOL_TRY(dalloc(&p->d_uv, nb * (size_t)p->uv_stride));
        POOL_TRY(dalloc(&p->d_fwords, 1 + 3 * nb));
        POOL_TRY(hipMemset(p->d_fwords, 0, sizeof(uint32_t) * (1 + 3 * nb)));
    }
    if (p->arena && p->arena_used > p->arena_size)
        return bail(fail(QAP_PT_E_NOMEM,
                         "the instances need %zu bytes of device memory; the fixed scratch "
                         "region holds %zu",
                         p->arena_used, p->arena_size));
    for (auto &b : p->buckets)
        POOL_TRY(hipMemcpy(b.d_list, b.reps.data(), b.reps.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice));

    std::vector<InstState> st(count);
    for (auto &s : st) {
        std::memset(&s, 0, sizeof(s));
        s.global_best = INT64_MAX;
    }
    std::vector<int64_t> bestE(RTall, INT64_MAX), ladder_best(p->total_ladders, INT64_MAX);
    std::vector<uint64_t> acc(RTall, 0);
    std::vector<int32_t> tss(p->total_ladders, 0);
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    };
    POOL_TRY(up(p->d_inst, p->desc.data(), sizeof(InstDesc) * count));
    POOL_TRY(up(p->d_state, st.data(), sizeof(InstState) * count));
    POOL_TRY(up(p->d_rep_inst, rep_inst.data(), 4 * rep_inst.size()));
    POOL_TRY(up(p->d_F, hF.data(), 4 * hF.size()));
    POOL_TRY(up(p->d_D, hD.data(), 4 * hD.size()));
    POOL_TRY(up(p->d_Ft, hFt.data(), 4 * hFt.size()));
    POOL_TRY(up(p->d_Dt, hDt.data(), 4 * hDt.size()));
    POOL_TRY(up(p->d_Fdiag, hFd.data(), 4 * hFd.size()));
    POOL_TRY(up(p->d_Ddiag, hDd.data(), 4 * hDd.size()));
    POOL_TRY(up(p->d_fclass, hCls.data(), 4 * hCls.size()));
    POOL_TRY(up(p->d_fspan, hSpan.data(), 4 * hSpan.size()));
    POOL_TRY(up(p->d_beta2f, beta2f.data(), 4 * beta2f.size()));
    POOL_TRY(up(p->d_c20, c20.data(), 8 * c20.size()));
    POOL_TRY(up(p->d_perm, perm.data(), 4 * perm.size()));
    POOL_TRY(up(p->d_pinv, pinv.data(), 4 * pinv.size()));
    POOL_TRY(up(p->d_best_perm, perm.data(), 4 * perm.size()));
    POOL_TRY(up(p->d_bestE, bestE.data(), 8 * bestE.size()));
    POOL_TRY(up(p->d_accepted, acc.data(), 8 * acc.size()));
    POOL_TRY(up(p->d_cells, acc.data(), 8 * acc.size()));
    POOL_TRY(up(p->d_cells + RTall, acc.data(), 8 * acc.size()));
    POOL_TRY(up(p->d_pos, pos.data(), 4 * pos.size()));
    POOL_TRY(up(p->d_order, order.data(), 4 * order.size()));
    POOL_TRY(up(p->d_tss, tss.data(), 4 * tss.size()));
    POOL_TRY(up(p->d_ladder_best, ladder_best.data(), 8 * ladder_best.size()));
    POOL_TRY(hipMemsetAsync(p->d_I, 0, sizeof(int32_t) * std::max<int64_t>(p->total_I, 1),
                            p->stream));
    POOL_TRY(hipMemsetAsync(p->d_I64, 0, sizeof(int64_t) * std::max<int64_t>(p->total_I64, 1),
                            p->stream));
    POOL_TRY(hipMemsetAsync(p->d_sync, 0, sizeof(uint32_t) * (1 + 2 * (size_t)count), p->stream));
    // tag 0 is never awaited: the first exchange step inside a launch has megas_done >= 1
    POOL_TRY(hipMemsetAsync(p->d_xslot, 0, sizeof(uint64_t) * 4 * (size_t)RTall, p->stream));
    POOL_TRY(launch_narrow16(p->d_F, p->d_F16, p->total_mat, p->stream));
    POOL_TRY(launch_narrow16(p->d_D, p->d_D16, p->total_mat, p->stream));

    bind_pool(p);

    // Force matrices (CPU_PIM.cpp:164-181) and costs (:184-192) on the device.
    for (int k = 0; k < count; ++k) {
        const InstDesc &d = p->desc[k];
        const int RT = d.R * d.T;
        if (p->shapes[k].im == IM_HBM64) {
            POOL_TRY(launch_force64(p->d_F + d.mat_off, p->d_D + d.mat_off, p->d_perm + d.p_base,
                                    d.n, d.ld, RT, p->d_I64 + d.I64_base, p->stream));
            POOL_TRY(launch_energy(p->d_F + d.mat_off, p->d_D + d.mat_off, p->d_perm + d.p_base,
                                   d.n, RT, p->d_E + d.rep_base, p->stream));
        } else {
            POOL_TRY(enqueue_force_build(p, d, p->fbuild[k], p->stream));
        }
    }
    POOL_TRY(hipStreamSynchronize(p->stream));
#undef POOL_TRY
    *out = p;
    return QAP_PT_OK;
}

int pool_run_impl(qap_pt_pool *p, int64_t sweeps, double *kernel_ms, int64_t *launches) {
    if (!p) return fail(QAP_PT_E_INVAL, "pool is NULL");
    HIP_TRY(hipSetDevice(p->device));
    int64_t max_new = 0;
    bool any_goal = false;
    double timeout = 0.0;
    for (int k = 0; k < p->count; ++k) {
        const qap_pt_params &pr = p->params[k];
        int64_t megas;
        if (sweeps < 0) {
            megas = pr.sweeps / pr.sweeps_per_exchange - p->limits[k];
            if (megas < 0) megas = 0;
        } else {
            if (sweeps % pr.sweeps_per_exchange)
                return fail(QAP_PT_E_INVAL, "instance %d: sweeps %lld not a multiple of %d", k,
                            (long long)sweeps, pr.sweeps_per_exchange);
            megas = sweeps / pr.sweeps_per_exchange;
        }
        if ((p->limits[k] + megas) * pr.sweeps_per_exchange > (int64_t)0xffffffffLL)
            return fail(QAP_PT_E_INVAL, "instance %d: total sweeps would exceed 2^32", k);
        p->limits[k] += megas;
        max_new = std::max(max_new, megas);
        any_goal = any_goal || pr.has_goal;
        if (pr.timeout_s > 0 && (timeout == 0.0 || pr.timeout_s < timeout)) timeout = pr.timeout_s;
    }
    HIP_TRY(hipMemcpyAsync(p->d_limits, p->limits.data(), sizeof(int64_t) * p->count,
                           hipMemcpyHostToDevice, p->stream));
    HIP_TRY(launch_set_limits(p->d_state, p->d_limits, p->count, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));

    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&]() {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    if (kernel_ms) {
        // Timed mode: the graph's launches issued directly, bucket concurrency included,
        // every sweep launch bracketed by events on the stream it is launched on.
        if (p->plain_launch) return fail(QAP_PT_E_INVAL, "timed runs need a pool of its own");
        const int rc0 = ensure_side_streams(p);
        if (rc0) return rc0;
        const int nb = (int)p->buckets.size();
        p->bucket_ms.resize(nb, 0.0);
        p->bucket_launches.resize(nb, 0);
        // persistent pools: one launch per bucket covers up to kPersistMaxIntervals
        // intervals, bracketed by events on the pool stream
        const int chunk = p->persistent ? kPersistMaxIntervals : 64;
        std::vector<hipEvent_t> ev(2 * (p->persistent ? 1 : chunk) * std::max(nb, 1), nullptr);
        int rc = QAP_PT_OK;
        for (auto &e : ev)
            if (hipEventCreate(&e) != hipSuccess) rc = fail(QAP_PT_E_DEVICE, "hipEventCreate failed");
        double total = 0.0;
        int64_t nl = 0;
        int64_t m = 0;
        while (m < max_new && rc == QAP_PT_OK) {
            int used = 0;
            if (p->persistent) {
                const int n = (int)std::min<int64_t>(max_new - m, chunk);
                const hipError_t e = enqueue_persistent(p, n, ev.data());
                if (e != hipSuccess) {
                    rc = fail(QAP_PT_E_DEVICE, "persistent sweep launch failed: %s",
                              hipGetErrorString(e));
                    break;
                }
                used = 1;
                m += n;
            }
            for (; !p->persistent && used < chunk && m < max_new; ++used, ++m) {
                const hipError_t e = p->helpers > 0 ? enqueue_helped(p, ev.data() + 2 * used * nb)
                                                    : enqueue_interval(p, ev.data() + 2 * used * nb);
                if (e != hipSuccess) {
                    rc = fail(QAP_PT_E_DEVICE, "interval launch failed: %s", hipGetErrorString(e));
                    break;
                }
            }
            if (hipStreamSynchronize(p->stream) != hipSuccess) {
                rc = fail(QAP_PT_E_DEVICE, "stream synchronize failed");
                break;
            }
            for (int u = 0; u < used; ++u)
                for (int b = 0; b < nb; ++b) {
                    float ms = 0.f;
                    (void)hipEventElapsedTime(&ms, ev[2 * (u * nb + b)], ev[2 * (u * nb + b) + 1]);
                    total += ms;
                    p->bucket_ms[b] += ms;
                    p->bucket_launches[b] += 1;
                }
            nl += (int64_t)used * nb;
            if (timeout > 0 && elapsed() > timeout) {
                (void)launch_halt(p->d_state, p->count, p->stream);
                break;
            }
        }
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (rc) return rc;
        *kernel_ms = total;
        if (launches) *launches = nl;
    } else {
        // goal or timeout: the host looks at the instances after every kGraphDepth intervals
        const int64_t batch = (any_goal || timeout > 0) ? kGraphDepth : std::max<int64_t>(max_new, 1);
        for (int64_t m = 0; m < max_new; m += batch) {
            const int rc = enqueue_intervals(p, std::min<int64_t>(batch, max_new - m));
            if (rc) return rc;
            if (any_goal || timeout > 0) {
                HIP_TRY(hipStreamSynchronize(p->stream));
                if (timeout > 0 && elapsed() > timeout) {
                    HIP_TRY(launch_halt(p->d_state, p->count, p->stream));
                    break;
                }
                if (any_goal) {
                    std::vector<InstState> st(p->count);
                    HIP_TRY(hipMemcpy(st.data(), p->d_state, sizeof(InstState) * p->count,
                                      hipMemcpyDeviceToHost));
                    bool all_done = true;
                    for (int k = 0; k < p->count; ++k)
                        all_done = all_done && (st[k].halted || st[k].megas_done >= p->limits[k]);
                    if (all_done) break;
                }
            }
        }
        if (launches) *launches = 0;
    }
    HIP_TRY(hipStreamSynchronize(p->stream));
    p->seconds += elapsed();
    if (p->persistent || p->helpers > 0) {
        uint32_t fault = 0;
        HIP_TRY(hipMemcpy(&fault, p->persistent ? p->d_sync : p->d_fwords, sizeof(fault),
                          hipMemcpyDeviceToHost));
        if (fault)
            return fail(QAP_PT_E_DEVICE, "a cooperative sweep launch gave up waiting for the "
                                         "other workgroups of its replica or ladder");
    }
    return QAP_PT_OK;
}

int pool_result_impl(qap_pt_pool *p, int inst, qap_pt_result *res) {
    if (!p || !res) return fail(QAP_PT_E_INVAL, "NULL argument");
    if (inst < 0 || inst >= p->count) return fail(QAP_PT_E_INVAL, "instance %d out of range", inst);
    HIP_TRY(hipSetDevice(p->device));
    const InstDesc &d = p->desc[inst];
    const int RT = d.R * d.T, n = d.n;
    std::vector<int64_t> E(RT), bestE(RT);
    std::vector<uint64_t> acc(RT), cells(RT), fcells(RT);
    InstState st;
    HIP_TRY(hipMemcpy(E.data(), p->d_E + d.rep_base, 8 * RT, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(bestE.data(), p->d_bestE + d.rep_base, 8 * RT, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(acc.data(), p->d_accepted + d.rep_base, 8 * RT, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cells.data(), p->d_cells + d.rep_base, 8 * RT, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(fcells.data(), p->d_cells + p->total_reps + d.rep_base, 8 * RT,
                      hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&st, p->d_state + inst, sizeof(st), hipMemcpyDeviceToHost));
    int bi = -1;
    int64_t best = INT64_MAX;
    for (int r = 0; r < RT; ++r)
        if (bestE[r] < best) {
            best = bestE[r];
            bi = r;
        }
    if (bi < 0) {  // no accepted move anywhere: the best initial state
        for (int r = 0; r < RT; ++r)
            if (bi < 0 || E[r] < best) {
                best = E[r];
                bi = r;
            }
    }
    res->best_cost = best;
    if (res->best_perm)
        HIP_TRY(hipMemcpy(res->best_perm, p->d_best_perm + d.p_base + (int64_t)bi * n, 4 * n,
                          hipMemcpyDeviceToHost));
    res->sweeps_done = st.megas_done * d.spe;
    res->attempts = (uint64_t)res->sweeps_done * (uint64_t)n * (uint64_t)RT;
    res->accepted = 0;
    for (auto a : acc) res->accepted += a;
    res->update_cells = 0;
    for (auto c : cells) res->update_cells += c;
    res->flush_cells = 0;
    for (auto c : fcells) res->flush_cells += c;
    res->exchanges = st.exchanges;
    res->kicks = st.kicks;
    res->seconds = p->seconds;
    return QAP_PT_OK;
}

}  // namespace

namespace qap {

int pool_create_internal(int32_t count, const qap_problem *problems, const qap_pt_params *params,
                         qap_pt_pool **out) {
    try {
        return pool_create_impl(count, problems, params, out);
    } catch (const std::bad_alloc &) {
        return fail(QAP_PT_E_NOMEM, "host allocation failed");
    }
}

int pool_launch_async(qap_pt_pool *p, hipEvent_t done) {
    HIP_TRY(hipSetDevice(p->device));
    int64_t max_new = 0;
    for (int k = 0; k < p->count; ++k) {
        const qap_pt_params &pr = p->params[k];
        const int64_t megas = pr.sweeps / pr.sweeps_per_exchange - p->limits[k];
        if (megas > 0) {
            p->limits[k] += megas;
            max_new = std::max(max_new, megas);
        }
    }
    HIP_TRY(hipMemcpyAsync(p->d_limits, p->limits.data(), sizeof(int64_t) * p->count,
                           hipMemcpyHostToDevice, p->stream));
    HIP_TRY(launch_set_limits(p->d_state, p->d_limits, p->count, p->stream));
    if (max_new > 0) {
        const int rc = enqueue_intervals(p, max_new);
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(done, p->stream));
    return QAP_PT_OK;
}

int pool_result_internal(qap_pt_pool *p, int inst, qap_pt_result *res) {
    try {
        return pool_result_impl(p, inst, res);
    } catch (...) {
        return fail(QAP_PT_E_NOMEM, "host exception in result");
    }
}

void pool_destroy_internal(qap_pt_pool *p) { pool_free(p); }

int pool_create_in_arena(int32_t count, const qap_problem *problems, const qap_pt_params *params,
                         void *arena, size_t size, hipStream_t stream, qap_pt_pool **out) {
    const ArenaSpec a{static_cast<char *>(arena), size, stream};
    try {
        return pool_create_impl(count, problems, params, out, &a);
    } catch (const std::bad_alloc &) {
        return fail(QAP_PT_E_NOMEM, "host allocation failed");
    }
}

bool wide_instance(const qap_problem &pb) {
    return width_class(pb.n, compute_bounds(pb.n, pb.F, pb.D)) > 0;
}

int64_t pool_device_bytes(int32_t count, const qap_problem *problems, const qap_pt_params *params) {
    int64_t bytes = 0;
    for (int k = 0; k < count; ++k) {
        const int64_t n = problems[k].n, RT = (int64_t)params[k].R * params[k].T;
        const int64_t ld = hbm_stride((int)n);
        bytes += 4 * (4 * n * n + 4 * n) + 4 * n * n;  // F, D, Ft, Dt, 16-bit F and D, diagonals, classes, spans
        bytes += (wide_instance(problems[k]) ? 8 : 4) * RT * n * ld;  // force matrices
        bytes += 4 * 3 * RT * n;                      // perm, pinv, best_perm
        bytes += RT * (8 * 5 + 4 * 3);                // E, bestE, accepted, 2 x cells, pos, order, rep_inst
        bytes += RT * 8 * 4;                          // tagged cost words of mirrored exchange steps
        bytes += 24 * params[k].T + 16 * params[k].R + sizeof(InstDesc) + sizeof(InstState) + 8;
    }
    return bytes;
}

}  // namespace qap

extern "C" {

const char *qap_pt_last_error(void) { return qap::g_last_error.c_str(); }

int qap_pt_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int qap_pt_pool_create(int32_t count, const qap_problem *problems, const qap_pt_params *params,
                       qap_pt_pool **out) {
    try {
        return pool_create_impl(count, problems, params, out);
    } catch (const std::bad_alloc &) {
        return fail(QAP_PT_E_NOMEM, "host allocation failed");
    } catch (...) {
        return fail(QAP_PT_E_INVAL, "unexpected host exception");
    }
}

int qap_pt_pool_run(qap_pt_pool *pool, int64_t sweeps, double *kernel_ms, int64_t *launches) {
    try {
        return pool_run_impl(pool, sweeps, kernel_ms, launches);
    } catch (...) {
        return fail(QAP_PT_E_NOMEM, "host exception in run");
    }
}

int qap_pt_pool_result(qap_pt_pool *pool, int32_t inst, qap_pt_result *res) {
    try {
        return pool_result_impl(pool, inst, res);
    } catch (...) {
        return fail(QAP_PT_E_NOMEM, "host exception in result");
    }
}

int qap_pt_pool_best_costs(qap_pt_pool *p, int64_t *best_costs) {
    if (!p || !best_costs) return fail(QAP_PT_E_INVAL, "NULL argument");
    try {
        HIP_TRY(hipSetDevice(p->device));
        std::vector<int64_t> E(p->total_reps), bestE(p->total_reps);
        HIP_TRY(hipMemcpy(E.data(), p->d_E, 8 * (size_t)p->total_reps, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(bestE.data(), p->d_bestE, 8 * (size_t)p->total_reps, hipMemcpyDeviceToHost));
        for (int k = 0; k < p->count; ++k) {
            const InstDesc &d = p->desc[k];
            const int RT = d.R * d.T;
            int64_t best = INT64_MAX, init = INT64_MAX;
            for (int r = 0; r < RT; ++r) {
                best = std::min(best, bestE[d.rep_base + r]);
                init = std::min(init, E[d.rep_base + r]);
            }
            // as qap_pt_pool_result: with no accepted move anywhere, the best initial cost
            best_costs[k] = best != INT64_MAX ? best : init;
        }
        return QAP_PT_OK;
    } catch (const std::bad_alloc &) {
        return fail(QAP_PT_E_NOMEM, "host allocation failed");
    }
}

int qap_pt_pool_state(qap_pt_pool *p, int32_t inst, int32_t *perm, int64_t *E, int64_t *bestE,
                      int32_t *pos, uint64_t *accepted, int32_t *best_perms) {
    if (!p) return fail(QAP_PT_E_INVAL, "pool is NULL");
    if (inst < 0 || inst >= p->count) return fail(QAP_PT_E_INVAL, "instance %d out of range", inst);
    HIP_TRY(hipSetDevice(p->device));
    const InstDesc &d = p->desc[inst];
    const int64_t RT = (int64_t)d.R * d.T, n = d.n;
    if (perm)
        HIP_TRY(hipMemcpy(perm, p->d_perm + d.p_base, 4 * RT * n, hipMemcpyDeviceToHost));
    if (best_perms)
        HIP_TRY(hipMemcpy(best_perms, p->d_best_perm + d.p_base, 4 * RT * n, hipMemcpyDeviceToHost));
    if (E) HIP_TRY(hipMemcpy(E, p->d_E + d.rep_base, 8 * RT, hipMemcpyDeviceToHost));
    if (bestE) HIP_TRY(hipMemcpy(bestE, p->d_bestE + d.rep_base, 8 * RT, hipMemcpyDeviceToHost));
    if (pos) HIP_TRY(hipMemcpy(pos, p->d_pos + d.rep_base, 4 * RT, hipMemcpyDeviceToHost));
    if (accepted)
        HIP_TRY(hipMemcpy(accepted, p->d_accepted + d.rep_base, 8 * RT, hipMemcpyDeviceToHost));
    return QAP_PT_OK;
}

int qap_pt_pool_verify(qap_pt_pool *p, int64_t *force_mismatches, int64_t *cost_mismatches) {
    if (!p) return fail(QAP_PT_E_INVAL, "pool is NULL");
    HIP_TRY(hipSetDevice(p->device));
    int64_t fm = 0, cm = 0;
    for (int k = 0; k < p->count; ++k) {
        const InstDesc &d = p->desc[k];
        const int RT = d.R * d.T;
        int32_t *scratch = nullptr;
        int64_t *Es = nullptr;
        unsigned long long *cnt = nullptr;
        const bool wide = p->shapes[k].im == IM_HBM64;
        const size_t words = (size_t)RT * d.n * d.ld * (wide ? 2 : 1);
        if (hipMalloc((void **)&scratch, 4 * words) != hipSuccess ||
            hipMalloc((void **)&Es, 8 * RT) != hipSuccess ||
            hipMalloc((void **)&cnt, 8) != hipSuccess) {
            if (scratch) (void)hipFree(scratch);
            if (Es) (void)hipFree(Es);
            return fail(QAP_PT_E_NOMEM, "verify scratch allocation failed");
        }
        (void)hipMemsetAsync(scratch, 0, 4 * words, p->stream);
        (void)hipMemsetAsync(cnt, 0, 8, p->stream);
        if (wide) {
            int64_t *s64 = reinterpret_cast<int64_t *>(scratch);
            (void)launch_force64(p->d_F + d.mat_off, p->d_D + d.mat_off, p->d_perm + d.p_base, d.n,
                                 d.ld, RT, s64, p->stream);
            (void)launch_count_mismatch64(s64, p->d_I64 + d.I64_base, d.n, d.ld, (int64_t)RT * d.n,
                                          cnt, p->stream);
        } else {
            (void)launch_force(p->d_F + d.mat_off, p->d_D + d.mat_off, p->d_perm + d.p_base, d.n,
                               d.ld, RT, scratch, p->stream);
            (void)launch_count_mismatch(scratch, p->d_I + d.I_base, d.n, d.ld, (int64_t)RT * d.n,
                                        cnt, p->stream);
        }
        (void)launch_energy(p->d_F + d.mat_off, p->d_D + d.mat_off, p->d_perm + d.p_base, d.n, RT, Es,
                      p->stream);
        unsigned long long c = 0;
        std::vector<int64_t> e1(RT), e2(RT);
        hipError_t e = hipStreamSynchronize(p->stream);
        if (e == hipSuccess) e = hipMemcpy(&c, cnt, 8, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(e1.data(), Es, 8 * RT, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(e2.data(), p->d_E + d.rep_base, 8 * RT, hipMemcpyDeviceToHost);
        (void)hipFree(scratch);
        (void)hipFree(Es);
        (void)hipFree(cnt);
        if (e != hipSuccess) return fail(QAP_PT_E_DEVICE, "verify failed: %s", hipGetErrorString(e));
        fm += (int64_t)c;
        for (int r = 0; r < RT; ++r) cm += e1[r] != e2[r];
    }
    if (force_mismatches) *force_mismatches = fm;
    if (cost_mismatches) *cost_mismatches = cm;
    return QAP_PT_OK;
}

int qap_pt_pool_destroy(qap_pt_pool *pool) {
    pool_free(pool);
    return QAP_PT_OK;
}

int qap_pt_solve_batch_hip(int32_t count, const qap_problem *problems,
                           const qap_pt_params *params, qap_pt_result *results) {
    if (!results) return fail(QAP_PT_E_INVAL, "results is NULL");
    qap_pt_pool *p = nullptr;
    int rc = qap_pt_pool_create(count, problems, params, &p);
    if (rc) return rc;
    rc = qap_pt_pool_run(p, -1, nullptr, nullptr);
    for (int k = 0; rc == QAP_PT_OK && k < count; ++k) rc = qap_pt_pool_result(p, k, &results[k]);
    pool_free(p);
    return rc;
}

int qap_pt_solve_hip(const qap_problem *prob, const qap_pt_params *par, qap_pt_result *res) {
    if (!prob || !par || !res) return fail(QAP_PT_E_INVAL, "NULL argument");
    return qap_pt_solve_batch_hip(1, prob, par, res);
}

int qap_pt_pool_shape(qap_pt_pool *p, int32_t inst, char *buf, int32_t buflen) {
    if (!p || !buf || buflen < 1) return fail(QAP_PT_E_INVAL, "bad argument");
    if (inst < 0 || inst >= p->count) return fail(QAP_PT_E_INVAL, "instance %d out of range", inst);
    const SweepShape &s = p->shapes[inst];
    const InstDesc &d = p->desc[inst];
    char spec[16] = "";
    if (s.spec > 0) std::snprintf(spec, sizeof(spec), ",spec%d", s.spec);
    std::snprintf(buf, buflen, "%dx%d,%s,%s%s,%s", s.nt, s.cpl,
                  placement_name(s.im),
                  s.sym ? "sym" : "general", spec, d.sparse_upd ? "sparse-flow" : "dense-flow");
    return QAP_PT_OK;
}

int qap_pt_pool_bucket(qap_pt_pool *p, int32_t b, char *shape, int32_t buflen,
                       int64_t *replicas, double *kernel_ms, int64_t *launches) {
    if (!p) return fail(QAP_PT_E_INVAL, "pool is NULL");
    if (b < 0 || b >= (int32_t)p->buckets.size())
        return fail(QAP_PT_E_INVAL, "bucket %d out of range (%d buckets)", b,
                    (int)p->buckets.size());
    const Bucket &B = p->buckets[b];
    if (shape && buflen > 0) {
        char spec[16] = "";
        if (B.shape.spec > 0) std::snprintf(spec, sizeof(spec), ",spec%d", B.shape.spec);
        std::snprintf(shape, buflen, "%dx%d,%s,%s%s%s", B.shape.nt, B.shape.cpl,
                      placement_name(B.shape.im), B.shape.sym ? "sym" : "general", spec,
                      B.shape.sparse ? ",sparse-flow" : "");
    }
    if (replicas) *replicas = (int64_t)B.reps.size();
    if (kernel_ms) *kernel_ms = b < (int32_t)p->bucket_ms.size() ? p->bucket_ms[b] : 0.0;
    if (launches) *launches = b < (int32_t)p->bucket_launches.size() ? p->bucket_launches[b] : 0;
    return QAP_PT_OK;
}

int qap_pt_kernel_shape(int32_t n, int64_t replicas, char *buf, int32_t buflen) {
    if (n < 1 || n > QAP_PT_MAX_N || replicas < 1 || !buf || buflen < 1)
        return fail(QAP_PT_E_INVAL, "bad argument");
    int ld_s = 0;
    const SweepShape s = choose_shape(n, true, true, (int64_t)1 << 40, false, true,
                                      throughput_regime(replicas), &ld_s);
    std::snprintf(buf, buflen, "%dx%d,%s", s.nt, s.cpl,
                  placement_name(s.im));
    return QAP_PT_OK;
}

}  // extern "C"
