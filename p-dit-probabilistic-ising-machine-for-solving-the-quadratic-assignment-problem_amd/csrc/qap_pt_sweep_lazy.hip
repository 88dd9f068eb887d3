// qap_pt_sweep_lazy.hip -- p-dit sweeps of symmetric instances whose force matrix stays in
// HBM (n too large for LDS), with the rank-2 updates accumulated lazily.
//
// The force matrix in HBM holds I_base; the rank-2 updates of the last K accepted moves,
// I += u_k v_k^T (u = distance differences, v = flow differences, CPU_PIM.cpp:263-269),
// wait in LDS as 16-bit vectors stored k-contiguous per row: U[x][k], V[c][k].  Every
// entry the update rule reads (CPU_PIM.cpp:220-232) is I_base + sum_k U[x][k] V[c][k],
// formed as 16-bit dot products:
//   row i      I[i][b] for lane b:  I_base[i][b] + dot(U[i], V[b])
//   column a   I[j][a] for lane b:  I_base[j][a] + dot(U[j], V[a])     (j = pinv(b))
// (the I_base entries prefetched into lane b's registers one site ahead),
//   diagonal   dg[c] = I[pinv(c)][c], kept exact move by move; I[i][a] = dg[a].
// When K reaches the capacity the LDS holds (kmax <= 32, up to 64 from n = 384) the
// pending updates are applied to I_base in one pass over the matrix, so kmax accepted moves
// cost one read and one write of I instead of kmax; every launch ends with such a flush, so
// I in HBM is exact between launches.  The flush is an integer matrix product U V^T on the matrix cores: each 16-bit
// entry splits into two signed bytes, and the int8 products (exact in int32) are
// recombined with shifts.  All sums are 32-bit two's complement: the entries themselves
// fit (the host checks the force bound), so wrapped partial sums still give the exact
// entries.
//
// One workgroup per replica as in qap_pt_sweep.hip (candidate b owned by thread
// b % NT); the heat-bath pick, Philox counters and exchange semantics are the same, so
// trajectories are identical to the in-LDS and in-HBM kernels.
#include "qap_pt_device.h"
#include "qap_pt_launch.h"
#include "qap_pt_mfma_update.h"
#include "qap_pt_sweep_common.h"

namespace qap {

namespace {

// reduction slots: [0,16) per-wave maxima and [16,80) per-wave prefix sums of the pick,
// [96,100) the winner (facility, delta, location, I[j][a]), [104,108) the next site's
// p^-1-indexed entries two candidates trade on a move, [140,156) per-wave update counts
constexpr int kLazyRed = 160;
constexpr int kLazyMaxK = 64;
// Below n = 384 at most 32 pending updates: the LDS of 32 leaves room for more workgroups
// per CU, which the full-chip loads use.  From n = 384 on one workgroup holds its CU's LDS
// anyway, and up to 64 pending updates (all 64 fit at n = 512) halve the flushes per sweep;
// only the 512-thread shapes take more than 32 (the smaller ones keep their register
// budget), so those instances run with 512 threads.
constexpr int kLazyLargeN = 384;
__host__ __device__ constexpr int lazy_cap(int n) { return n >= kLazyLargeN ? kLazyMaxK : 32; }
// p, pinv, cls, rnd, dg, drow x2 -- round8(n) words each
constexpr int kLazyArrays = 7;

__host__ __device__ inline int lazy_fixed_words(int n) { return kLazyArrays * round8(n) + kLazyRed; }

// Pending-update capacity (kmax, a multiple of 8: rows are read 8 entries at a time) and
// U / V row stride (kstride int16) that `lds_bytes` of LDS give an instance of size n.
// The stride is an odd number of 16-byte groups, so the 128-bit reads of consecutive
// rows by consecutive lanes fall in distinct LDS banks.
struct LazyLayout {
    int kmax, kstride;
};

__host__ __device__ inline LazyLayout lazy_layout(int n, int lds_bytes) {
    const int avail = lds_bytes - 4 * lazy_fixed_words(n);
    const int g = avail > 0 ? avail / (4 * round8(n) * 8) : 0;  // 8-entry groups that fit
    const int cap = lazy_cap(n) / 8;
    int kg = g < cap ? g : cap;
    int sg = kg | 1;
    if (sg > g) sg = --kg;  // kg was even and only kg groups fit: drop to kg - 1 (odd)
    return LazyLayout{8 * (kg > 0 ? kg : 0), 8 * (sg > 0 ? sg : 0)};
}

// sum_k x[k] y[k] over packed signed 16-bit entries, 8 per 128-bit word, on the packed
// dot-product instruction (v_dot2c_i32_i16: lo*lo + hi*hi + acc, wrapping in 32 bits
// without the clamp: exactly the 32-bit sum the update rule needs).  The pairs are taken
// as shuffles of the word viewed as 8 shorts: bit-casting the word's int elements one by
// one to short2 in an unrolled loop was compiled to four copies of the first pair (a
// device check of this helper against integer arithmetic found it; the shuffle, memcpy and
// short2-pointer forms are exact).
typedef short lazy_s2 __attribute__((ext_vector_type(2)));
typedef short lazy_s8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ int32_t dot8(v4h x, v4h y, int32_t acc) {
    const lazy_s8 a = __builtin_bit_cast(lazy_s8, x), b = __builtin_bit_cast(lazy_s8, y);
    acc = __builtin_amdgcn_sdot2(__builtin_shufflevector(a, a, 0, 1),
                                 __builtin_shufflevector(b, b, 0, 1), acc, false);
    acc = __builtin_amdgcn_sdot2(__builtin_shufflevector(a, a, 2, 3),
                                 __builtin_shufflevector(b, b, 2, 3), acc, false);
    acc = __builtin_amdgcn_sdot2(__builtin_shufflevector(a, a, 4, 5),
                                 __builtin_shufflevector(b, b, 4, 5), acc, false);
    acc = __builtin_amdgcn_sdot2(__builtin_shufflevector(a, a, 6, 7),
                                 __builtin_shufflevector(b, b, 6, 7), acc, false);
    return acc;
}

}  // namespace

int lazy_lds_bytes(int n, int max_pending) {
    // max_pending entries (a multiple of 8, at most lazy_cap(n)) in an odd number of 8-entry
    // groups
    const int groups = (max_pending < lazy_cap(n) ? max_pending : lazy_cap(n)) / 8;
    const int want = 4 * lazy_fixed_words(n) + 4 * round8(n) * 8 * (groups | 1);
    return want < kMaxLdsBytes ? want : kMaxLdsBytes;
}

int lazy_lds_bytes(int n) { return lazy_lds_bytes(n, lazy_cap(n)); }

bool lazy_fits(int n) { return lazy_layout(n, lazy_lds_bytes(n)).kmax >= 8; }

// Workgroups of up to 256 threads run where many replicas share a CU (the full-chip
// batched and served loads): three waves per SIMD instead of two (at most 168 VGPRs, no
// spills) with two row tiles of the flush in flight instead of four measured +5.5 % on
// 4096 replicas of n = 30..150 and +2.8 % on the served stream, with n = 512 (512
// threads, unchanged) equal.
template <int NT> constexpr int kLazyWaves = NT >= 512 ? 2 : 3;
template <int NT> constexpr int kLazyFlushDepth = NT >= 512 ? kUpdDepth : 2;

// Longest a flush helper or a replica waits at a flush: 30 s of the 100 MHz real-time
// counter (the launch is cooperative, so the waits end; the limit turns a fault into a
// reported error instead of a hang).
constexpr uint64_t kFlushWaitTicks = 3000000000ull;
constexpr uint32_t kFlushStop = 0xffffffffu;

// Rows [lo, hi) of I_base += U V^T for K pending updates: one 32-entry chunk on the
// matrix cores per pass, two when K > 32 (512-thread shapes only).
template <int NT>
__device__ __forceinline__ void lazy_flush_part(int32_t *Ig, int ld, int n, const int16_t *U,
                                                const int16_t *V, int kpad, int K, int lane,
                                                int wid, int lo, int hi) {
    constexpr int NW = NT / kWave;
    if constexpr (NT >= 512) {
        if (K > 32) {
            mfma_rank_update<NW, 2, kLazyFlushDepth<NT>>(Ig, ld, n, U, V, kpad, K, lane, wid, lo, hi);
            return;
        }
    }
    mfma_rank_update<NW, 1, kLazyFlushDepth<NT>>(Ig, ld, n, U, V, kpad, K, lane, wid, lo, hi);
}

// Rows [lo, hi) of part p of a flush split into `parts` runs of whole 16-row tiles.
__device__ __forceinline__ void flush_rows(int n, int parts, int p, int &lo, int &hi) {
    const int ntile = (n + 15) / 16;
    const int per = (ntile + parts - 1) / parts;
    lo = min(n, 16 * per * p);
    hi = min(n, 16 * per * (p + 1));
}

// HELP: built with the flush-helper protocol (cooperative launches of a pool that brings
// helper workgroups); plain launches use the build without it, whose smaller register
// footprint spills fewer scalar registers in the sweep loop (+1.2 % on the 512-instance
// batch).
template <int NT, int CPL, bool HELP>
__global__ __launch_bounds__(NT, kLazyWaves<NT>) void sweep_lazy_kernel(PoolDev P, const int32_t *__restrict__ rep_list,
                                                        int lds_bytes, FlushHelp fh) {
    constexpr int NW = NT / kWave;
    extern __shared__ __align__(16) int32_t smem[];

    // blocks from fh.nrep on are flush helpers: helper h of replica block r is block
    // fh.nrep (1 + h) + r (the same XCD as r when fh.nrep is a multiple of 8)
    const bool helper = HELP && fh.helpers > 0 && (int)blockIdx.x >= fh.nrep;
    const int rb = helper ? ((int)blockIdx.x - fh.nrep) % fh.nrep : (int)blockIdx.x;
    const int g = rep_list[rb];
    const int inst = P.rep_inst[g];
    if (!P.state[inst].active) return;  // the replica and its helpers alike
    const InstDesc d = P.inst[inst];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int n = d.n, ld = d.ld, n4 = round8(n);
    const int local = g - d.rep_base;
    const LazyLayout lay = lazy_layout(n, lds_bytes);
    // more than 32 pending updates: 512-thread shapes only (two 32-entry chunks per flush)
    const int kmax = NT >= 512 ? lay.kmax : min(lay.kmax, 32);
    const int kpad = lay.kstride;  // U / V row stride in int16
    if (kmax < 8) return;          // the host places an instance here only when 8 updates fit
    uint32_t *req_w = fh.words + 1 + 3 * rb, *k_w = req_w + 1, *done_w = req_w + 2;
    const int16_t *uvg = fh.uv + fh.uv_stride * rb;

    if (HELP && helper) {
        // ---- flush helper: wait for a request, update this helper's rows, arrive ----
        int32_t *Ig = P.I + d.I_base + (int64_t)local * n * ld;
        uint32_t *slot = reinterpret_cast<uint32_t *>(smem);
        int lo = 0, hi = 0;
        flush_rows(n, fh.helpers + 1, 1 + ((int)blockIdx.x - fh.nrep) / fh.nrep, lo, hi);
        uint32_t seen = 0;
        for (;;) {
            if (tid == 0) {
                const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
                uint32_t v;
                // relaxed polls: an acquire on every poll would invalidate the caches the
                // replicas beside this helper read from; one acquire fence follows the change
                while ((v = __hip_atomic_load(req_w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ==
                       seen) {
                    __builtin_amdgcn_s_sleep(8);
                    if (__builtin_amdgcn_s_memrealtime() - t0 > kFlushWaitTicks) {
                        __hip_atomic_store(fh.words, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        v = kFlushStop;
                        break;
                    }
                }
                slot[0] = v;
                slot[1] = __hip_atomic_load(k_w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __syncthreads();
            const uint32_t v = slot[0];
            const int K = (int)slot[1];
            __syncthreads();
            if (v == kFlushStop) return;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            if (lo < hi)
                lazy_flush_part<NT>(Ig, ld, n, uvg, uvg + (size_t)n4 * kpad, kpad, K, lane, wid, lo,
                                    hi);
            // every wave releases its own rows and counts itself
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            if (lane == 0)
                __hip_atomic_fetch_add(done_w, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            seen = v;
        }
    }

    int32_t *p_s = smem;
    int32_t *pinv_s = p_s + n4;
    int32_t *cls_s = pinv_s + n4;
    uint32_t *rnd_s = reinterpret_cast<uint32_t *>(cls_s + n4);
    int32_t *dg_s = reinterpret_cast<int32_t *>(rnd_s + n4);
    int32_t *drow_s = dg_s + n4;  // 2 buffers: D rows of this site and the next, for u
    int32_t *red_s = drow_s + 2 * n4;
    int16_t *U_s = reinterpret_cast<int16_t *>(red_s + kLazyRed);  // [n4][kpad]
    int16_t *V_s = U_s + (size_t)n4 * kpad;                         // [n4][kpad]

    // the 16-bit copies: every entry of an instance placed here fits (the host checks that
    // 2 max|D| and 4 max|F| do), and half the bytes keep more of the rows in L2
    const int16_t *__restrict__ F = P.F16 + d.mat_off;
    const int16_t *__restrict__ D = P.D16 + d.mat_off;
    int32_t *pg = P.perm + d.p_base + (int64_t)local * n;
    int32_t *pinvg = P.pinv + d.p_base + (int64_t)local * n;
    int32_t *bestpg = P.best_perm + d.p_base + (int64_t)local * n;
    int32_t *Ig = P.I + d.I_base + (int64_t)local * n * ld;

    auto zero_pending = [&]() {
        v4h *u4 = reinterpret_cast<v4h *>(U_s);
        const int words = n4 * kpad / 8 * 2;  // U and V are contiguous
        for (int x = tid; x < words; x += NT) u4[x] = v4h{0, 0, 0, 0};
    };

    // ---- stage the replica --------------------------------------------------
    for (int x = tid; x < n; x += NT) {
        p_s[x] = pg[x];
        pinv_s[x] = pinvg[x];
        cls_s[x] = P.fclass[d.vec_off + x];
    }
    zero_pending();
    lds_sync<NW>();
    int a_cur = p_s[0];
    for (int x = tid; x < n; x += NT) {
        dg_s[x] = Ig[(int64_t)pinv_s[x] * ld + x];
        drow_s[x] = D[x];
    }
    lds_sync<NW>();

    int64_t E = P.E[g];
    int64_t bestE = P.bestE[g];
    uint64_t acc = P.accepted[g];
    uint64_t cells = 0;
    uint64_t fcells = 0;  // entries written by flushes (the same count in every thread)
    const int pos = P.pos[g];
    const float beta2f = P.beta2f[d.beta_off + pos];
    const double c20 = P.c20[d.beta_off + pos];
    const uint32_t sweep0 = P.state[inst].sweep_base;

    int32_t pv[CPL];
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
        const int b = tid + q * NT;
        pv[q] = (b < n) ? pinv_s[b] : 0;
    }
    int buf = 0, K = 0, a_reg = a_cur;
    bool stale = false;  // the site's I_base entries in registers are out of date

    // The entries of a site that candidate b = tid + q NT reads go straight to its lane's
    // registers: I_base[site][b], I_base[pinv(b)][fac], D[site][pinv(b)] and F[fac][b]
    // (fac = p(site)); the next site's arrive while this site's pick runs.  A move changes
    // pinv at two candidates only, a and b, whose prefetched pinv-indexed entries then
    // trade places.
    auto load_site = [&](int site, int fac, int32_t(&r)[CPL], int32_t(&c)[CPL],
                         int32_t(&dd)[CPL], int32_t(&f)[CPL]) {
#pragma unroll
        for (int q = 0; q < CPL; ++q) {
            const int bc = min(tid + q * NT, n - 1);
            const int jq = pv[q];
            r[q] = Ig[(int64_t)site * ld + bc];
            c[q] = Ig[(int64_t)jq * ld + fac];
            dd[q] = D[(int64_t)site * n + jq];
            f[q] = F[(int64_t)fac * n + bc];
        }
    };
    int32_t cr[CPL], cc[CPL], cd[CPL], cf[CPL];  // this site's
    load_site(0, a_cur, cr, cc, cd, cf);

    // I_base += U V^T over the whole matrix on the matrix cores, then no pending update is
    // left.  With helpers the rows are split: U, V and K go to global memory, a request
    // wakes the helpers, this workgroup updates the first run of rows, and every helper
    // wave's arrival is awaited before the matrix is read again.
    uint32_t fseq = 0;  // flushes requested from the helpers in this launch
    auto flush = [&](bool keep_prefetch) {
        if (K == 0) return;
        fcells += (uint64_t)n * (uint64_t)n;
        if (HELP && fh.helpers > 0) {
            int16_t *uvw = fh.uv + fh.uv_stride * rb;
            const v4h *src = reinterpret_cast<const v4h *>(U_s);
            v4h *dst = reinterpret_cast<v4h *>(uvw);
            const int words = n4 * kpad / 8 * 2;  // U and V are contiguous, 8 int16 per word
            for (int x = tid; x < words; x += NT) dst[x] = src[x];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __syncthreads();
            ++fseq;
            if (tid == 0) {
                __hip_atomic_store(k_w, (uint32_t)K, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(req_w, fseq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
            int lo = 0, hi = 0;
            flush_rows(n, fh.helpers + 1, 0, lo, hi);
            lazy_flush_part<NT>(Ig, ld, n, U_s, V_s, kpad, K, lane, wid, lo, hi);
            if (tid == 0) {
                const uint32_t want = fseq * (uint32_t)(fh.helpers * (NT / kWave));
                const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
                while (__hip_atomic_load(done_w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != want) {
                    __builtin_amdgcn_s_sleep(1);
                    if (__builtin_amdgcn_s_memrealtime() - t0 > kFlushWaitTicks) {
                        __hip_atomic_store(fh.words, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
            }
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        } else {
            lazy_flush_part<NT>(Ig, ld, n, U_s, V_s, kpad, K, lane, wid, 0, n);
        }
        __syncthreads();  // the matrix writes reach every wave before it is read again
        zero_pending();
        K = 0;
        stale = !keep_prefetch;
        lds_sync<NW>();
    };

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
            if (a != a_cur || stale) {  // the last move put another facility here: reload
                a_cur = a;
                load_site(i, a, cr, cc, cd, cf);
                stale = false;
            }
            const int32_t *drow = drow_s + buf * n4;  // D[i][.]
            const int groups = K / 8 + ((K & 7) != 0);
            const v4h *Ui = reinterpret_cast<const v4h *>(U_s + (size_t)i * kpad);
            const v4h *Va = reinterpret_cast<const v4h *>(V_s + (size_t)a * kpad);

            // (1) deltas of every candidate from I_base plus the pending updates, without
            // the term -I[i][a] = -dg[a] that all share: the pick subtracts it after its
            // first barrier, which also publishes the diagonal entries the last move wrote
            int32_t dE[CPL], rv[CPL], cv[CPL];
            int lmax = INT32_MIN;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const int b = tid + q * NT;
                dE[q] = INT32_MIN;
                rv[q] = 0;
                cv[q] = 0;
                if (b < n) {
                    const int j = pv[q];
                    int32_t r = cr[q], c = cc[q];
                    const v4h *Vb = reinterpret_cast<const v4h *>(V_s + (size_t)b * kpad);
                    const v4h *Uj = reinterpret_cast<const v4h *>(U_s + (size_t)j * kpad);
                    for (int gg = 0; gg < groups; ++gg) {
                        r = dot8(Ui[gg], Vb[gg], r);
                        c = dot8(Uj[gg], Va[gg], c);
                    }
                    rv[q] = r;  // I[i][b]
                    cv[q] = c;  // I[j][a]
                    const int32_t v = r + c - dg_s[b] - 4 * cd[q] * cf[q];
                    dE[q] = v;
                    lmax = max(lmax, v);
                }
            }
            // prefetch for the next site: its distance row (to LDS, for u) first, then
            // the entries its candidates read (registers)
            const int inext = (i + 1 == n) ? 0 : i + 1;
            const int a_nx = p_s[inext];
            int32_t dn[CPL], nr[CPL], nc[CPL], nd[CPL], nf[CPL];
#pragma unroll
            for (int q = 0; q < CPL; ++q) dn[q] = D[(int64_t)inext * n + min(tid + q * NT, n - 1)];
            load_site(inext, a_nx, nr, nc, nd, nf);

            // (2)-(3) heat-bath pick; the winner commits the permutation swap
            const Pick pk =
                heat_bath_pick<NT, CPL>(dE, lmax, pv, ri, beta2f, c20, a, i, red_s, tid, dg_s + a);
            int32_t my_rv = 0, my_cv = 0;
#pragma unroll
            for (int q = 0; q < CPL; ++q)
                if (pk.owner && pk.b == tid + q * NT) {
                    my_rv = rv[q];
                    my_cv = cv[q];
                }
            if (pk.owner && pk.b != a) {
                p_s[i] = pk.b;
                p_s[pk.j] = a;
            }
            {
                int32_t *dnb = drow_s + (buf ^ 1) * n4;
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const int bb = tid + q * NT;
                    if (bb < n) dnb[bb] = dn[q];
                }
            }
            if (pk.owner) {
                red_s[96] = pk.b;
                red_s[97] = pk.d;
                red_s[98] = pk.j;
                red_s[99] = my_cv;  // I[j][a] before the move
            }
            if (pk.S == 0 && tid == 0) {  // every weight truncated to zero: stay
                red_s[96] = a;
                red_s[97] = 0;
                red_s[98] = i;
                red_s[99] = 0;
            }
            lds_sync<NW>();
            const int b = red_s[96], j = red_s[98];
            const int32_t dsel = red_s[97], cv_w = red_s[99];
            // candidates a and b trade inverse-permutation entries, and so their prefetched
            // pinv-indexed entries of the next site (handed over in LDS, read after the
            // next barrier)
            auto trade_prefetch = [&]() {
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const int bb = tid + q * NT;
                    if (bb == a) {
                        nc[q] = red_s[106];
                        nd[q] = red_s[107];
                    } else if (bb == b) {
                        nc[q] = red_s[104];
                        nd[q] = red_s[105];
                    }
                }
            };
            if (b != a) {
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const int bb = tid + q * NT;
                    if (bb == a) {
                        red_s[104] = nc[q];
                        red_s[105] = nd[q];
                    } else if (bb == b) {
                        red_s[106] = nc[q];
                        red_s[107] = nd[q];
                    }
                    if (bb == b) pv[q] = i;
                    else if (bb == a) pv[q] = j;
                }
                E -= dsel;  // (4) commit (CPU_PIM.cpp:252-283)
                ++acc;
                if (((cls_s[a] ^ cls_s[b]) & 0xffff) != 0) {  // flow class (low 16 bits)
                    // the move's u (rows) and v (columns) become pending update K
                    uint32_t cnt = 0;
#pragma unroll
                    for (int q = 0; q < CPL; ++q) {
                        const int c = tid + q * NT;
                        const int cc = min(c, n - 1);
                        const int32_t fb = F[(int64_t)b * n + cc];
                        const int32_t dj = D[(int64_t)j * n + cc];
                        const int32_t v = 2 * (fb - cf[q]);   // 2 (F[b][c] - F[a][c])
                        if (c < n) {
                            U_s[(size_t)c * kpad + K] = (int16_t)(dj - drow[c]);  // D[j][c] - D[i][c]
                            V_s[(size_t)c * kpad + K] = (int16_t)v;
                        }
                        cnt += (uint32_t)__popcll(__ballot(c < n && v != 0));
                    }
                    cells += (uint64_t)n * cnt;
                    lds_sync<NW>();
                    trade_prefetch();
                    // diagonal entries after the move: the swapped facilities take the
                    // entries at their new locations, the others add u[pinv(c)] v[c]
#pragma unroll
                    for (int q = 0; q < CPL; ++q) {
                        const int c = tid + q * NT;
                        if (c < n) {
                            const int32_t vc = V_s[(size_t)c * kpad + K];
                            if (c == b) dg_s[c] = my_rv + (int32_t)U_s[(size_t)i * kpad + K] * vc;
                            else if (c == a) dg_s[c] = cv_w + (int32_t)U_s[(size_t)j * kpad + K] * vc;
                            else dg_s[c] += (int32_t)U_s[(size_t)pv[q] * kpad + K] * vc;
                        }
                    }
                    ++K;
                    // no barrier: each lane wrote its own diagonal entries, and the one
                    // another lane reads (dg[a] of the next site) waits for the pick's
                    // first barrier
                    if (K == kmax) {
                        // the next site's prefetched I_base entries take the pending
                        // updates the flush applies, instead of being read again:
                        // I[inext][x] += U[inext] . V[x], I[pinv(x)][a_nx] += U[pinv(x)] . V[a_nx]
                        const v4h *Un = reinterpret_cast<const v4h *>(U_s + (size_t)inext * kpad);
                        const v4h *Vn = reinterpret_cast<const v4h *>(V_s + (size_t)a_nx * kpad);
                        const int kg = K / 8 + ((K & 7) != 0);
#pragma unroll
                        for (int q = 0; q < CPL; ++q) {
                            const int x = tid + q * NT;
                            if (x < n) {
                                int32_t r = nr[q], c = nc[q];
                                const v4h *Vx = reinterpret_cast<const v4h *>(V_s + (size_t)x * kpad);
                                const v4h *Uj = reinterpret_cast<const v4h *>(U_s + (size_t)pv[q] * kpad);
                                for (int gg = 0; gg < kg; ++gg) {
                                    r = dot8(Un[gg], Vx[gg], r);
                                    c = dot8(Uj[gg], Vn[gg], c);
                                }
                                nr[q] = r;
                                nc[q] = c;
                            }
                        }
                        flush(true);
                    }
                } else {
                    // identical flow rows and columns: I's columns a and b are equal, so the
                    // diagonal entries of the two facilities trade places
                    if (tid == 0) {
                        const int32_t t = dg_s[a];
                        dg_s[a] = dg_s[b];
                        dg_s[b] = t;
                    }
                    lds_sync<NW>();
                    trade_prefetch();
                }
                if (E < bestE) {  // CPU_PIM.cpp:271-272, plus the permutation
                    bestE = E;
                    for (int x = tid; x < n; x += NT) bestpg[x] = p_s[x];
                }
            }
            buf ^= 1;
#pragma unroll
            for (int q = 0; q < CPL; ++q) {  // the next site's entries become current
                cr[q] = nr[q];
                cc[q] = nc[q];
                cd[q] = nd[q];
                cf[q] = nf[q];
            }
            a_cur = a_nx;
            a_reg = (b != a && j == inext) ? a : a_nx;
        }
    }
    flush(false);
    if (HELP && fh.helpers > 0 && tid == 0)  // the helpers leave
        __hip_atomic_store(req_w, kFlushStop, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);

    if (lane == 0) reinterpret_cast<uint64_t *>(red_s + 140)[wid] = cells;
    lds_sync<NW>();
    for (int x = tid; x < n; x += NT) pg[x] = p_s[x];
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
        const int bb = tid + q * NT;
        if (bb < n) pinvg[bb] = pv[q];
    }
    if (tid == 0) {
        uint64_t tot = cells;
        if constexpr (NW > 1) {
            const uint64_t *wc = reinterpret_cast<const uint64_t *>(red_s + 140);
#pragma unroll
            for (int w = 1; w < NW; ++w) tot += wc[w];
        }
        P.E[g] = E;
        P.bestE[g] = bestE;
        P.accepted[g] = acc;
        P.cells[g] += tot;
        P.fcells[g] += fcells;
    }
}

// mode 0: plain launch of nblocks; 1: cooperative launch of fh.nrep (1 + fh.helpers)
// blocks; 2: nothing launched, hipSuccess when that many blocks are resident at once
template <int NT, int CPL>
hipError_t launch_lazy_t(const PoolDev &P, const int32_t *list, int nblocks, int lds_bytes,
                         const FlushHelp &fh, int mode, hipStream_t s) {
    if (mode == 0) {
        const hipError_t e = allow_full_lds<sweep_lazy_kernel<NT, CPL, false>>();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((sweep_lazy_kernel<NT, CPL, false>), dim3(nblocks), dim3(NT), lds_bytes,
                           s, P, list, lds_bytes, fh);
        return hipGetLastError();
    }
    auto k = sweep_lazy_kernel<NT, CPL, true>;
    const hipError_t e = allow_full_lds<sweep_lazy_kernel<NT, CPL, true>>();
    if (e != hipSuccess) return e;
    const int blocks = fh.nrep * (1 + fh.helpers);
    if (mode == 2) {
        int per_cu = 0, dev = 0, cus = 0;
        hipError_t q = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(k), NT, (size_t)lds_bytes);
        if (q == hipSuccess) q = hipGetDevice(&dev);
        if (q == hipSuccess) q = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (q != hipSuccess) return q;
        return (int64_t)per_cu * cus >= blocks ? hipSuccess : hipErrorCooperativeLaunchTooLarge;
    }
    PoolDev pd = P;
    const int32_t *lp = list;
    int lb = lds_bytes;
    FlushHelp f = fh;
    void *args[] = {&pd, &lp, &lb, &f};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k), dim3(blocks), dim3(NT),
                                      args, (unsigned)lds_bytes, s);
}

static hipError_t lazy_dispatch(const PoolDev &P, const SweepShape &sh, const int32_t *list,
                                int nblocks, int lds_bytes, const FlushHelp &fh, int mode,
                                hipStream_t s) {
    if (!sh.sym) return hipErrorInvalidValue;
    switch (sh.nt * 8 + sh.cpl) {
    case 64 * 8 + 1: return launch_lazy_t<64, 1>(P, list, nblocks, lds_bytes, fh, mode, s);
    case 64 * 8 + 2: return launch_lazy_t<64, 2>(P, list, nblocks, lds_bytes, fh, mode, s);
    case 64 * 8 + 4: return launch_lazy_t<64, 4>(P, list, nblocks, lds_bytes, fh, mode, s);
    case 128 * 8 + 1: return launch_lazy_t<128, 1>(P, list, nblocks, lds_bytes, fh, mode, s);
    case 128 * 8 + 2: return launch_lazy_t<128, 2>(P, list, nblocks, lds_bytes, fh, mode, s);
    case 256 * 8 + 1: return launch_lazy_t<256, 1>(P, list, nblocks, lds_bytes, fh, mode, s);
    case 256 * 8 + 2: return launch_lazy_t<256, 2>(P, list, nblocks, lds_bytes, fh, mode, s);
    case 256 * 8 + 4: return launch_lazy_t<256, 4>(P, list, nblocks, lds_bytes, fh, mode, s);
    case 512 * 8 + 1: return launch_lazy_t<512, 1>(P, list, nblocks, lds_bytes, fh, mode, s);
    case 512 * 8 + 2: return launch_lazy_t<512, 2>(P, list, nblocks, lds_bytes, fh, mode, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_sweep_lazy(const PoolDev &P, const SweepShape &sh, const int32_t *list,
                             int nblocks, int lds_bytes, hipStream_t s) {
    return lazy_dispatch(P, sh, list, nblocks, lds_bytes, FlushHelp{}, 0, s);
}

hipError_t launch_sweep_lazy_helped(const PoolDev &P, const SweepShape &sh, const int32_t *list,
                                    int lds_bytes, const FlushHelp &fh, hipStream_t s) {
    if (fh.helpers < 1 || fh.nrep < 1 || !fh.uv || !fh.words) return hipErrorInvalidValue;
    return lazy_dispatch(P, sh, list, fh.nrep, lds_bytes, fh, 1, s);
}

bool lazy_helpers_fit(const SweepShape &sh, int nblocks, int helpers, int lds_bytes) {
    FlushHelp fh;
    fh.nrep = nblocks;
    fh.helpers = helpers;
    return lazy_dispatch(PoolDev{}, sh, nullptr, nblocks, lds_bytes, fh, 2, nullptr) == hipSuccess;
}

int64_t lazy_pending_elems(int n, int lds_bytes) {
    return 2 * (int64_t)round8(n) * lazy_layout(n, lds_bytes).kstride;
}

}  // namespace qap
