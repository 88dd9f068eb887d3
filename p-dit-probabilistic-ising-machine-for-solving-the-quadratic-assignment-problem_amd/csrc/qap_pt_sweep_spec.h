// qap_pt_sweep_spec.h -- p-dit sweeps that evaluate several sites at once (included by
// qap_pt_sweep.hip, whose exchange step it shares).
//
// A rejected update leaves the replica's state -- permutation, force matrix, cost -- as it
// was, and every site's uniform is a Philox word of (site, sweep, replica), independent of
// what happened before.  So the updates of sites i0, i0+1, ..., i0+S-1 can be evaluated
// together against the current state: wave w evaluates site i0+w (all n <= 64 candidates,
// one per lane, the same heat-bath pick as the one-wave sweep kernel), and after one
// barrier the first site whose pick moved a facility is committed; the sites before it
// were rejections, the sites after it are evaluated again in the next round, from the
// committed state.  The chain is exactly the one site-by-site update (CPU_PIM.cpp:195-290),
// and when most updates are rejected (cold replicas) a round advances up to S sites in the
// time of one.
//
// Symmetric instances with n <= 64 whose force matrix fits LDS as int32 (the in-LDS int32
// placement); D and F are held in LDS as well, so no global memory is read between
// launches' staging and write-back.  Exchange intervals, persistent launches and the
// in-kernel exchange step (persistent_exchange) follow sweep_kernel.
#pragma once

namespace qap {

constexpr int kSpecRes = 4;  // words per wave in the round's result slots: b, j, delta, a

// The result slots are double-buffered by round parity: a round that commits nothing ends
// without a barrier, and a wave that starts the next round may write its slot while another
// still reads this round's; it writes that buffer again only after the next round's
// barrier, which every wave reaches after its reads.
__host__ __device__ inline int spec_lds_words(int n, int ld_s, int S) {
    const int n4 = round8(n);
    return 4 * n4 + 2 * kSpecRes * S + kRedWords + 2 * n * n + n * ld_s;
}

template <int S>
__global__ __launch_bounds__(S * kWave) void sweep_spec_kernel(PoolDev P,
                                                               const int32_t *__restrict__ rep_list,
                                                               int max_iv) {
    constexpr int NT = S * kWave;
    extern __shared__ __align__(16) int32_t smem[];

    const int g = rep_list[blockIdx.x];
    const int inst = P.rep_inst[g];
    if (!P.state[inst].active) return;
    const InstDesc d = P.inst[inst];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int n = d.n, ld = d.ld, ld_s = d.ld_s, n4 = round8(n);
    const int local = g - d.rep_base;

    int32_t *p_s = smem;
    int32_t *pinv_s = p_s + n4;
    int32_t *cls_s = pinv_s + n4;
    uint32_t *rnd_s = reinterpret_cast<uint32_t *>(cls_s + n4);
    int32_t *res_s = reinterpret_cast<int32_t *>(rnd_s + n4);  // [2][S][kSpecRes]
    int32_t *red_s = res_s + 2 * kSpecRes * S;
    int32_t *D_s = red_s + kRedWords;  // D and F, n x n row-major
    int32_t *F_s = D_s + n * n;
    int32_t *I_s = F_s + n * n;        // force matrix, row stride ld_s

    int32_t *pg = P.perm + d.p_base + (int64_t)local * n;
    int32_t *pinvg = P.pinv + d.p_base + (int64_t)local * n;
    int32_t *bestpg = P.best_perm + d.p_base + (int64_t)local * n;
    int32_t *Ig = P.I + d.I_base + (int64_t)local * n * ld;

    // ---- stage the replica, D and F into LDS ------------------------------------
    for (int x = tid; x < n; x += NT) {
        p_s[x] = pg[x];
        pinv_s[x] = pinvg[x];
        cls_s[x] = P.fclass[d.vec_off + x];
    }
    for (int e = tid; e < n * n; e += NT) {
        D_s[e] = P.D[d.mat_off + e];
        F_s[e] = P.F[d.mat_off + e];
    }
    for (int e = tid; e < n * n; e += NT) {
        const int x = e / n, c = e - x * n;
        I_s[x * ld_s + c] = Ig[(int64_t)x * ld + c];
    }
    lds_sync<S>();

    int64_t E = P.E[g];
    int64_t bestE = P.bestE[g];
    uint64_t acc = P.accepted[g];
    uint64_t cells = 0;  // counted by thread 0
    int pos = P.pos[g];
    float beta2f = P.beta2f[d.beta_off + pos];
    double c20 = P.c20[d.beta_off + pos];
    uint32_t sweep0 = P.state[inst].sweep_base;
    const bool mirror = max_iv > 1 && d.xmirror && P.xslot;  // as sweep_kernel
    XMirror xm{};
    int xstat = kXmRun;
    if (mirror && wid == 0) xm = xm_load(P, inst, d, lane);
    XBeta bt{0.f, 0.0};
    if (mirror) bt = xm_beta_load(P, d, lane);
    uint32_t gen0 = 0;
    if (max_iv > 1 && !mirror)
        gen0 = __hip_atomic_load(&P.sync[2 + 2 * inst], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

    // every wave keeps lane b's location pinv(b) (the candidates are the same in each)
    int32_t pv[1];
    pv[0] = lane < n ? pinv_s[lane] : 0;

    bool pre = false;  // rnd_s already holds the Philox words of sweep pre_sweep
    uint32_t pre_sweep = 0;
    for (int iv = 0; iv < max_iv; ++iv) {
        if (iv > 0) {  // exchange step between intervals (persistent launch)
            // while wave 0 runs a mirrored step, wave 1 draws the next interval's first sweep
            // (every read of rnd_s is behind the last round's barrier; the step's barrier
            // publishes the words)
            pre_sweep = sweep0 + (uint32_t)d.spe;
            pre = mirror && S > 1;
            if (pre && wid == 1)
                for (int x = lane; x < n; x += kWave)
                    rnd_s[x] = philox4x32_10(u32x4{(uint32_t)x, pre_sweep, (uint32_t)local, 0u},
                                             d.seed_lo, d.seed_hi)
                                   .x;
            xstat = persistent_exchange(P, d, inst, g, local, E, bestE, gen0 + (uint32_t)iv,
                                        mirror, xm, red_s, tid, &pos, &sweep0);
            if (xstat != kXmRun) break;
            if (mirror) {
                xm_beta_at(bt, pos, beta2f, c20);
            } else {
                beta2f = P.beta2f[d.beta_off + pos];
                c20 = P.c20[d.beta_off + pos];
            }
        }
        for (int s = 0; s < d.spe; ++s) {
            const uint32_t sweep = sweep0 + (uint32_t)s;
            if (!(pre && s == 0 && sweep == pre_sweep)) {
                lds_sync<S>();
                for (int x = tid; x < n; x += NT)
                    rnd_s[x] = philox4x32_10(u32x4{(uint32_t)x, sweep, (uint32_t)local, 0u},
                                             d.seed_lo, d.seed_hi)
                                   .x;
                lds_sync<S>();
            }

            for (int i0 = 0, par = 0; i0 < n; par ^= 1) {
                int32_t *rs_s = res_s + par * (kSpecRes * S);  // this round's slots
                // (1)-(3) wave w: site i0 + w against the current state (one-wave pick)
                const int i = i0 + wid;
                if (i < n) {  // wave-uniform
                    const int a = p_s[i];
                    const int b = lane;
                    int32_t dE[1] = {INT32_MIN};
                    int lmax = 0;
                    if (b < n) {
                        const int j = pv[0];
                        const int32_t force = I_s[i * ld_s + b] + I_s[j * ld_s + a] -
                                              I_s[i * ld_s + a] - I_s[j * ld_s + b];
                        dE[0] = force - 4 * D_s[i * n + j] * F_s[a * n + b];
                        lmax = max(lmax, dE[0]);
                    }
                    const Pick pk =
                        heat_bath_pick<kWave, 1>(dE, lmax, pv, rnd_s[i], beta2f, c20, a, i, red_s, lane);
                    const uint64_t bal = __ballot(pk.owner);
                    int bw = a, jw = i;
                    int32_t dw = 0;
                    if (bal) {
                        const int src = __ffsll((unsigned long long)bal) - 1;
                        bw = __builtin_amdgcn_readlane(pk.b, src);
                        jw = __builtin_amdgcn_readlane(pk.j, src);
                        dw = __builtin_amdgcn_readlane(pk.d, src);
                    }
                    if (lane == 0) {
                        rs_s[kSpecRes * wid + 0] = bw;
                        rs_s[kSpecRes * wid + 1] = jw;
                        rs_s[kSpecRes * wid + 2] = dw;
                        rs_s[kSpecRes * wid + 3] = a;
                    }
                }
                lds_sync<S>();
                // the first of this round's sites whose pick moved a facility: lane w reads
                // site i0+w's slot (one 16-byte LDS read), a ballot finds the first move
                v4i rs = {0, 0, 0, 0};
                if (lane < S) rs = reinterpret_cast<const v4i *>(rs_s)[lane];
                const uint64_t moved = __ballot(lane < S && i0 + lane < n && rs.x != rs.w);
                if (moved == 0) {  // every site of the round stayed
                    i0 += S;
                    continue;
                }
                const int ws = __ffsll((unsigned long long)moved) - 1;
                const int ic = i0 + ws;
                const int b = __builtin_amdgcn_readlane(rs.x, ws);
                const int j = __builtin_amdgcn_readlane(rs.y, ws);
                const int32_t dsel = __builtin_amdgcn_readlane(rs.z, ws);
                const int a = __builtin_amdgcn_readlane(rs.w, ws);
                // (4) commit (CPU_PIM.cpp:252-283).  Every read of I and p this round came
                // before the barrier above, and the next round's slots are the other buffer,
                // so the commit writes without a barrier of its own.
                E -= dsel;
                ++acc;
                if (lane == b) pv[0] = ic;
                else if (lane == a) pv[0] = j;
                if (tid == 0) {
                    p_s[ic] = b;
                    p_s[j] = a;
                }
                if (((cls_s[a] ^ cls_s[b]) & 0xffff) != 0) {  // identical flow: v = 0
                    // I += u v^T, u = D[j] - D[ic] (rows), v = 2 (F[b] - F[a]) (columns):
                    // column c = lane, rows wid, wid + S, ...
                    const int c = min(lane, n - 1);
                    const int32_t vc = 2 * (F_s[b * n + c] - F_s[a * n + c]);
                    if (wid == 0) {  // nonzero v columns (n <= 64: one wave holds every column)
                        const uint64_t nzm = __ballot(lane < n && vc != 0);
                        if (lane == 0) cells += (uint64_t)n * (uint64_t)__popcll(nzm);
                    }
                    for (int x = wid; x < n; x += S) {
                        const int32_t ux = D_s[j * n + x] - D_s[ic * n + x];
                        if (lane < n) I_s[x * ld_s + lane] += ux * vc;
                    }
                }
                lds_sync<S>();  // the updated matrix and permutation before the next round
                if (E < bestE) {  // CPU_PIM.cpp:271-272, plus the permutation
                    bestE = E;
                    for (int x = tid; x < n; x += NT) bestpg[x] = p_s[x];
                }
                i0 = ic + 1;
            }
        }
    }  // intervals

    lds_sync<S>();
    for (int x = tid; x < n; x += NT) pg[x] = p_s[x];
    if (wid == 0 && lane < n) pinvg[lane] = pv[0];
    for (int e = tid; e < n * n; e += NT) {
        const int x = e / n, c = e - x * n;
        Ig[(int64_t)x * ld + c] = I_s[x * ld_s + c];
    }
    if (tid == 0) {
        P.E[g] = E;
        P.bestE[g] = bestE;
        P.accepted[g] = acc;
        P.cells[g] += cells;
    }
    if (mirror && local == 0 && wid == 0 && xstat != kXmFault)
        xm_store(P, inst, d, xm, xstat == kXmRun, lane);
}

}  // namespace qap
