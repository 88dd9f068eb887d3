// qap_pt_slots.cpp -- slot mode of the pool: the device state the serving loop keeps
// resident.  A slotted pool has `count` fixed-capacity instance slots (n <= n_cap,
// R <= r_cap rungs, T <= t_cap ladders, R*T <= rt_cap replicas), allocated once.  An
// instance is loaded into a free slot (uploads, force matrices and costs enqueued on the
// pool stream), runs in the same launches as every other loaded instance -- one exchange
// launch over all slots and one sweep launch per kernel shape per exchange interval --
// and is released when its run has ended.  Slot i's regions sit at fixed offsets, so
// loading never moves another instance; an empty or finished slot has megas_done >=
// mega_limit and is skipped by the exchange and sweep kernels.
//
// All slot calls other than slots_launch expect the pool stream to be idle (the previous
// launch finished): the serving loop loads and releases between launches.
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <map>
#include <tuple>

#include "qap_pt_internal.h"
#include "qap_pt_pool_impl.h"

namespace qap {
namespace {

constexpr int kBankLazyPending = 32;  // pending rank-2 updates a lazy bank slot holds

struct SlotLayout {
    int64_t mat, vec, beta, p, I, reps, ladders;
};

SlotLayout layout(int n_cap, int r_cap, int t_cap, int rt_cap) {
    SlotLayout L;
    L.mat = (int64_t)n_cap * n_cap;
    L.vec = n_cap;
    L.beta = r_cap;
    L.p = (int64_t)rt_cap * n_cap;
    L.I = (int64_t)rt_cap * n_cap * hbm_stride(n_cap);
    L.reps = rt_cap;
    L.ladders = t_cap;
    return L;
}

}  // namespace

int64_t slots_device_bytes(int32_t slots, int32_t n_cap, int32_t r_cap, int32_t t_cap,
                           int32_t rt_cap) {
    const SlotLayout L = layout(n_cap, r_cap, t_cap, rt_cap);
    int64_t per = 4 * (4 * L.mat + 4 * L.vec) + (4 + 8) * L.beta + 4 * 3 * L.p + 4 * L.I +
                  L.reps * (8 * 4 + 4 * 3) + L.ladders * (4 + 8) + sizeof(InstDesc) +
                  sizeof(InstState) + 8;
    return per * slots + 4 * (int64_t)slots * rt_cap;  // + one replica list per shape, roughly
}

int slots_create(int32_t slots, int32_t n_cap, int32_t r_cap, int32_t t_cap, int32_t rt_cap,
                 qap_pt_pool **out) {
    if (!out) return fail(QAP_PT_E_INVAL, "out is NULL");
    *out = nullptr;
    if (slots < 1 || n_cap < 1 || n_cap > QAP_PT_MAX_N || r_cap < 1 || t_cap < 1 || rt_cap < 1)
        return fail(QAP_PT_E_INVAL, "invalid slot capacity");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(QAP_PT_E_DEVICE, "no HIP device visible");
    qap_pt_pool *p = new qap_pt_pool();
    (void)hipGetDevice(&p->device);
    p->slotted = true;
    p->count = slots;
    p->n_cap = n_cap;
    p->r_cap = r_cap;
    p->t_cap = t_cap;
    p->rt_cap = rt_cap;
    p->max_rt = rt_cap;
    p->params.assign(slots, qap_pt_params{});
    p->desc.assign(slots, InstDesc{});
    p->shapes.assign(slots, SweepShape{});
    p->limits.assign(slots, 0);
    p->used.assign(slots, 0);
    p->staged.resize(slots);
    p->st_load.assign(slots, InstState{});
    p->fill64.resize(slots);
    const SlotLayout L = layout(n_cap, r_cap, t_cap, rt_cap);
    p->total_mat = L.mat * slots;
    p->total_vec = L.vec * slots;
    p->total_beta = L.beta * slots;
    p->total_p = L.p * slots;
    p->total_I = L.I * slots;
    p->total_reps = (int)(L.reps * slots);
    p->total_ladders = (int)(L.ladders * slots);

    auto bail = [&](int rc) {
        pool_free(p);
        return rc;
    };
#define SLOT_TRY(expr)                                                                            \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return bail(fail(e_ == hipErrorOutOfMemory ? QAP_PT_E_NOMEM : QAP_PT_E_DEVICE,        \
                             "%s failed: %s", #expr, hipGetErrorString(e_)));                     \
    } while (0)
    SLOT_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    const int64_t R = p->total_reps;
    SLOT_TRY(dalloc(&p->d_inst, slots));
    SLOT_TRY(dalloc(&p->d_state, slots));
    SLOT_TRY(dalloc(&p->d_limits, slots));
    SLOT_TRY(dalloc(&p->d_rep_inst, R));
    SLOT_TRY(dalloc(&p->d_F, p->total_mat));
    SLOT_TRY(dalloc(&p->d_D, p->total_mat));
    SLOT_TRY(dalloc(&p->d_F16, p->total_mat));
    SLOT_TRY(dalloc(&p->d_D16, p->total_mat));
    SLOT_TRY(dalloc(&p->d_Ft, p->total_mat));
    SLOT_TRY(dalloc(&p->d_Dt, p->total_mat));
    SLOT_TRY(dalloc(&p->d_Fdiag, p->total_vec));
    SLOT_TRY(dalloc(&p->d_Ddiag, p->total_vec));
    SLOT_TRY(dalloc(&p->d_fclass, p->total_vec));
    SLOT_TRY(dalloc(&p->d_fspan, p->total_vec));
    SLOT_TRY(dalloc(&p->d_beta2f, p->total_beta));
    SLOT_TRY(dalloc(&p->d_c20, p->total_beta));
    SLOT_TRY(dalloc(&p->d_perm, p->total_p));
    SLOT_TRY(dalloc(&p->d_pinv, p->total_p));
    SLOT_TRY(dalloc(&p->d_best_perm, p->total_p));
    SLOT_TRY(dalloc(&p->d_I, p->total_I));
    SLOT_TRY(dalloc(&p->d_E, R));
    SLOT_TRY(dalloc(&p->d_bestE, R));
    SLOT_TRY(dalloc(&p->d_accepted, R));
    SLOT_TRY(dalloc(&p->d_cells, 2 * R));  // update and flush counts
    SLOT_TRY(dalloc(&p->d_pos, R));
    SLOT_TRY(dalloc(&p->d_order, R));
    SLOT_TRY(dalloc(&p->d_tss, p->total_ladders));
    SLOT_TRY(dalloc(&p->d_ladder_best, p->total_ladders));
    // empty slots: n = 0 and megas_done == mega_limit == 0, so every kernel skips them
    std::vector<int32_t> rep_inst(R);
    for (int64_t g = 0; g < R; ++g) rep_inst[g] = (int32_t)(g / rt_cap);
    SLOT_TRY(hipMemcpy(p->d_rep_inst, rep_inst.data(), 4 * R, hipMemcpyHostToDevice));
    SLOT_TRY(hipMemcpy(p->d_inst, p->desc.data(), sizeof(InstDesc) * slots, hipMemcpyHostToDevice));
    SLOT_TRY(hipMemcpy(p->d_state, p->st_load.data(), sizeof(InstState) * slots,
                       hipMemcpyHostToDevice));
#undef SLOT_TRY
    bind_pool(p);
    *out = p;
    return QAP_PT_OK;
}

// (slots hold int32 force matrices: an instance beyond the int32 bounds runs in a pool of
// its own)
bool slots_fit(const qap_pt_pool *p, const qap_problem &pb, const qap_pt_params &pr) {
    return p && p->slotted && pb.n <= p->n_cap && pr.R <= p->r_cap && pr.T <= p->t_cap &&
           (int64_t)pr.R * pr.T <= p->rt_cap && !wide_instance(pb);
}

int slots_free(const qap_pt_pool *p) {
    for (int s = 0; s < p->count; ++s)
        if (!p->used[s]) return s;
    return -1;
}

int slots_in_use(const qap_pt_pool *p) {
    int k = 0;
    for (uint8_t u : p->used) k += u;
    return k;
}

int slots_load(qap_pt_pool *p, int slot, const qap_problem &pb, const qap_pt_params &pr) {
    if (!p || !p->slotted || slot < 0 || slot >= p->count || p->used[slot])
        return fail(QAP_PT_E_INVAL, "no such free slot");
    int rc = validate(pb, pr, slot);
    if (rc) return rc;
    if (!slots_fit(p, pb, pr)) return fail(QAP_PT_E_INVAL, "instance exceeds the slot capacity");
    QAP_HIP_TRY(hipSetDevice(p->device));
    HostInstance &h = p->staged[slot];
    h = HostInstance{};
    // a bank is sized to run full: its slots' placement follows the full-chip rates, and
    // every slot takes the shape of the bank's size class for its placement, so a bank
    // runs one launch per placement and interval (per-size shapes split the n <= 256 bank
    // into two lazy launches that the hardware queues largely serialised: the served
    // stream ran 178 M attempts/s against 241 M with one shape)
    prepare_instance(pb, pr, throughput_regime((int64_t)p->count * p->rt_cap), h);
    if (h.shape.im != IM_ONESHOT) shape_for_size(p->n_cap, h.d.sparse_upd != 0, h.shape);
    h.shape.spec = 0;  // bank slots share the one-site kernels (their LDS is sized for them)
    const SlotLayout L = layout(p->n_cap, p->r_cap, p->t_cap, p->rt_cap);
    InstDesc &d = h.d;
    const int n = d.n, RT = d.R * d.T;
    d.mat_off = L.mat * slot;
    d.vec_off = L.vec * slot;
    d.beta_off = (int32_t)(L.beta * slot);
    d.p_base = L.p * slot;
    d.I_base = L.I * slot;
    d.rep_base = (int32_t)(L.reps * slot);
    d.ladder_base = (int32_t)(L.ladders * slot);
    p->desc[slot] = d;
    p->shapes[slot] = h.shape;
    p->params[slot] = pr;
    p->params[slot].beta_ln = nullptr;  // the rungs live in the slot (beta2f, c20)
    p->params[slot].init_perm = nullptr;
    p->limits[slot] = pr.sweeps / pr.sweeps_per_exchange;

    // upload sources must outlive the asynchronous copies: they stay in the slot record
    std::vector<int64_t> &f = p->fill64[slot];
    f.assign((size_t)2 * RT + 2 * d.T, 0);
    int64_t *bestE = f.data(), *zeros = f.data() + RT, *lbest = f.data() + 2 * RT;
    std::fill(bestE, bestE + RT, INT64_MAX);
    std::fill(lbest, lbest + d.T, INT64_MAX);  // ladder bests; tss uses the zero tail below
    InstState &st = p->st_load[slot];
    std::memset(&st, 0, sizeof(st));
    st.global_best = INT64_MAX;
    st.mega_limit = p->limits[slot];

    hipStream_t s = p->stream;
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    };
    const int64_t nn = (int64_t)n * n;
    QAP_HIP_TRY(up(p->d_F + d.mat_off, h.F.data(), 4 * nn));
    QAP_HIP_TRY(up(p->d_D + d.mat_off, h.D.data(), 4 * nn));
    QAP_HIP_TRY(launch_narrow16(p->d_F + d.mat_off, p->d_F16 + d.mat_off, nn, s));
    QAP_HIP_TRY(launch_narrow16(p->d_D + d.mat_off, p->d_D16 + d.mat_off, nn, s));
    QAP_HIP_TRY(up(p->d_Ft + d.mat_off, h.Ft.data(), 4 * nn));
    QAP_HIP_TRY(up(p->d_Dt + d.mat_off, h.Dt.data(), 4 * nn));
    QAP_HIP_TRY(up(p->d_Fdiag + d.vec_off, h.Fd.data(), 4 * n));
    QAP_HIP_TRY(up(p->d_Ddiag + d.vec_off, h.Dd.data(), 4 * n));
    QAP_HIP_TRY(up(p->d_fclass + d.vec_off, h.cls.data(), 4 * n));
    QAP_HIP_TRY(up(p->d_fspan + d.vec_off, h.span.data(), 4 * n));
    QAP_HIP_TRY(up(p->d_beta2f + d.beta_off, h.beta2f.data(), 4 * d.R));
    QAP_HIP_TRY(up(p->d_c20 + d.beta_off, h.c20.data(), 8 * d.R));
    QAP_HIP_TRY(up(p->d_perm + d.p_base, h.perm.data(), 4 * (int64_t)RT * n));
    QAP_HIP_TRY(up(p->d_pinv + d.p_base, h.pinv.data(), 4 * (int64_t)RT * n));
    QAP_HIP_TRY(up(p->d_best_perm + d.p_base, h.perm.data(), 4 * (int64_t)RT * n));
    QAP_HIP_TRY(up(p->d_bestE + d.rep_base, bestE, 8 * RT));
    QAP_HIP_TRY(up(p->d_accepted + d.rep_base, zeros, 8 * RT));
    QAP_HIP_TRY(up(p->d_cells + d.rep_base, zeros, 8 * RT));
    QAP_HIP_TRY(up(p->d_cells + p->total_reps + d.rep_base, zeros, 8 * RT));
    QAP_HIP_TRY(up(p->d_pos + d.rep_base, h.pos.data(), 4 * RT));
    QAP_HIP_TRY(up(p->d_order + d.rep_base, h.pos.data(), 4 * RT));
    QAP_HIP_TRY(up(p->d_tss + d.ladder_base, zeros, 4 * d.T));
    QAP_HIP_TRY(up(p->d_ladder_best + d.ladder_base, lbest, 8 * d.T));
    QAP_HIP_TRY(up(p->d_inst + slot, &p->desc[slot], sizeof(InstDesc)));
    // force matrices (CPU_PIM.cpp:164-181) and costs (:184-192), then the state that
    // makes the slot live at the next exchange step
    QAP_HIP_TRY(enqueue_force_build(p, d, h.fb, s));
    QAP_HIP_TRY(up(p->d_state + slot, &st, sizeof(InstState)));
    p->used[slot] = 1;
    p->lists_dirty = true;
    return QAP_PT_OK;
}

int slots_launch(qap_pt_pool *p, int intervals, hipEvent_t done) {
    if (!p || !p->slotted) return fail(QAP_PT_E_INVAL, "not a slotted pool");
    QAP_HIP_TRY(hipSetDevice(p->device));
    if (p->lists_dirty) {
        // one replica list per kernel shape over the loaded slots (device lists sized for
        // every replica of every slot, allocated on first use of a shape)
        std::map<std::tuple<int, int, int, int>, int> key_of;
        for (int b = 0; b < (int)p->buckets.size(); ++b) {
            const SweepShape &sh = p->buckets[b].shape;
            key_of[std::make_tuple(sh.nt, sh.cpl, sh.im, (int)sh.sym)] = b;
            p->buckets[b].reps.clear();
            p->buckets[b].lds_bytes = 0;
            p->buckets[b].shape.sparse = true;
        }
        for (int s = 0; s < p->count; ++s) {
            if (!p->used[s]) continue;
            const SweepShape &sh = p->shapes[s];
            const InstDesc &d = p->desc[s];
            const auto key = std::make_tuple(sh.nt, sh.cpl, sh.im, (int)sh.sym);
            auto it = key_of.find(key);
            int b;
            if (it == key_of.end()) {
                b = (int)p->buckets.size();
                key_of[key] = b;
                p->buckets.emplace_back();
                p->buckets[b].shape = sh;
                p->buckets[b].shape.sparse = true;
                QAP_HIP_TRY(dalloc(&p->buckets[b].d_list, (size_t)p->total_reps));
            } else {
                b = it->second;
            }
            Bucket &B = p->buckets[b];
            B.shape.sparse = B.shape.sparse && d.sparse_upd != 0;  // all-sparse buckets only
            // lazy slots hold up to kBankLazyPending pending updates.  24 instead of 32 let
            // a third workgroup share a CU at n <= 256 while the kernel kept six rows of the
            // site in LDS (served stream 245 -> 264 M attempts/s); with those rows in
            // registers 32 fit three workgroups there as well, and 32 measured 266-270 ->
            // 282 M on the served stream)
            const int lds = sh.im == IM_LAZY ? std::min(sweep_lds_bytes(d.n, d.ld_s, sh.im),
                                                        lazy_lds_bytes(d.n, kBankLazyPending))
                                             : sweep_lds_bytes(d.n, d.ld_s, sh.im);
            B.lds_bytes = std::max(B.lds_bytes, lds);
            for (int r = 0; r < d.R * d.T; ++r) B.reps.push_back(d.rep_base + r);
        }
        const int rt_cap = p->rt_cap;
        for (auto &B : p->buckets) {
            order_bucket(p, B, [rt_cap](int32_t g) { return g / rt_cap; });
            if (!B.reps.empty())
                QAP_HIP_TRY(hipMemcpyAsync(B.d_list, B.reps.data(), 4 * B.reps.size(),
                                           hipMemcpyHostToDevice, p->stream));
        }
        p->lists_dirty = false;
    }
    for (int m = 0; m < intervals; ++m) {
        QAP_HIP_TRY(launch_exchange(p->P, p->count, p->max_rt, p->stream));
        for (auto &B : p->buckets)
            if (!B.reps.empty())
                QAP_HIP_TRY(launch_sweep(p->P, B.shape, B.d_list, (int)B.reps.size(), B.lds_bytes,
                                         p->stream));
    }
    if (done) QAP_HIP_TRY(hipEventRecord(done, p->stream));
    return QAP_PT_OK;
}

int slots_finished(qap_pt_pool *p, std::vector<int> &done, std::vector<InstState> &state) {
    done.clear();
    QAP_HIP_TRY(hipSetDevice(p->device));
    state.resize(p->count);
    QAP_HIP_TRY(hipMemcpy(state.data(), p->d_state, sizeof(InstState) * p->count,
                          hipMemcpyDeviceToHost));
    for (int s = 0; s < p->count; ++s)
        if (p->used[s] && (state[s].halted || state[s].megas_done >= p->limits[s]))
            done.push_back(s);
    return QAP_PT_OK;
}

int slots_halt(qap_pt_pool *p, int slot) {
    if (!p || !p->slotted || slot < 0 || slot >= p->count || !p->used[slot])
        return fail(QAP_PT_E_INVAL, "no such loaded slot");
    const int32_t one = 1;
    QAP_HIP_TRY(hipMemcpy(reinterpret_cast<char *>(p->d_state + slot) + offsetof(InstState, halted),
                          &one, sizeof(one), hipMemcpyHostToDevice));
    return QAP_PT_OK;
}

int slots_release(qap_pt_pool *p, int slot) {
    if (!p || !p->slotted || slot < 0 || slot >= p->count || !p->used[slot])
        return fail(QAP_PT_E_INVAL, "no such loaded slot");
    p->used[slot] = 0;
    p->staged[slot] = HostInstance{};
    p->fill64[slot].clear();
    p->lists_dirty = true;
    return QAP_PT_OK;
}

}  // namespace qap
