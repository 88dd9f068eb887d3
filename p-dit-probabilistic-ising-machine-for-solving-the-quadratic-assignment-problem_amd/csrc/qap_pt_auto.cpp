// qap_pt_auto.cpp -- automatic beta-ladder selection (CPU-PIM/CPU_PIM.cpp:366-476).
//
// The reference runs, for each exponent b in [-30, 1], five trials of every ladder with
// all replicas at beta2 = 2^b for 10 exchange intervals from fresh permutations, and
// averages the final cost over trials, ladders and replicas (:394-421).  Here the
// 32 x 5 trials are independent instances of one batched pool (split into several pools
// only when their device footprint exceeds kAutoBudget), so the whole scan costs about
// as much as one short run of the batch.  The selection rule over the averages
// (:431-473) is qap_pt_auto_select.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "qap_pt_internal.h"

namespace {

constexpr double kLog2e = 1.44269504089;  // the reference's ln -> exp2 factor (:528)
constexpr double kAutoBudget = 32.0 * (1ull << 30);

int auto_beta_impl(const qap_problem *prob, const qap_pt_params *base, double lower_prop,
                   double higher_prop, double *avg_energy, double *beta0_ln, double *ratio) {
    if (!prob || !base || !beta0_ln || !ratio) return qap::fail(QAP_PT_E_INVAL, "auto_beta: NULL argument");
    if (prob->n < 1 || prob->n > QAP_PT_MAX_N || !prob->F || !prob->D)
        return qap::fail(QAP_PT_E_INVAL, "auto_beta: n must be in [1, %d] with F and D set", QAP_PT_MAX_N);
    if (base->R < 1 || base->T < 1 || base->sweeps_per_exchange < 1)
        return qap::fail(QAP_PT_E_INVAL, "auto_beta: R, T and sweeps_per_exchange must be >= 1");
    const int steps = QAP_PT_AUTO_STEPS, trials = QAP_PT_AUTO_TRIALS;
    const int R = base->R, T = base->T, RT = R * T;
    std::vector<std::vector<double>> ladders(steps);
    for (int i = 0; i < steps; ++i)
        ladders[i].assign(R, std::exp2(QAP_PT_AUTO_MIN_EXP + (double)i) / kLog2e);
    std::vector<qap_problem> probs;
    std::vector<qap_pt_params> pars;
    for (int i = 0; i < steps; ++i)
        for (int t = 0; t < trials; ++t) {
            probs.push_back(*prob);
            qap_pt_params p = *base;
            p.beta_ln = ladders[i].data();
            p.sweeps = (int64_t)QAP_PT_AUTO_MEGAS * base->sweeps_per_exchange;
            p.has_goal = 0;
            p.timeout_s = 0.0;
            p.flags = base->flags & (QAP_PT_INIT_PER_REPLICA | QAP_PT_ONE_SHOT);
            p.init_perm = nullptr;  // every trial starts from its own shuffle (setup(), :130-161)
            // A distinct Philox key per trial; the base seed keeps the scan reproducible.
            p.seed = base->seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(1 + i * trials + t));
            pars.push_back(p);
        }
    const int total = steps * trials;
    const int64_t one = qap::pool_device_bytes(1, probs.data(), pars.data());
    int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(total, (int64_t)(kAutoBudget / (double)one)));
    std::vector<double> avg(steps, 0.0);
    std::vector<int64_t> E(RT);
    for (int k0 = 0; k0 < total; k0 += chunk) {
        const int cnt = std::min(chunk, total - k0);
        qap_pt_pool *pool = nullptr;
        int rc = qap_pt_pool_create(cnt, probs.data() + k0, pars.data() + k0, &pool);
        if (rc) return rc;
        rc = qap_pt_pool_run(pool, -1, nullptr, nullptr);
        for (int k = 0; k < cnt && rc == QAP_PT_OK; ++k) {
            rc = qap_pt_pool_state(pool, k, nullptr, E.data(), nullptr, nullptr, nullptr, nullptr);
            const int i = (k0 + k) / trials;
            for (int r = 0; r < RT && rc == QAP_PT_OK; ++r)
                avg[i] += (double)E[r] / ((double)trials * T * R);
        }
        qap_pt_pool_destroy(pool);
        if (rc) return rc;
    }
    if (avg_energy) std::memcpy(avg_energy, avg.data(), sizeof(double) * steps);
    return qap_pt_auto_select(avg.data(), steps, lower_prop, higher_prop, R, beta0_ln, ratio);
}

}  // namespace

extern "C" {

int qap_pt_auto_select(const double *avg_energy, int32_t steps, double lower_prop,
                       double higher_prop, int32_t R, double *beta0_ln, double *ratio) {
    if (!avg_energy || !beta0_ln || !ratio)
        return qap::fail(QAP_PT_E_INVAL, "auto_select: NULL argument");
    if (steps < 2 || steps > QAP_PT_AUTO_STEPS)
        return qap::fail(QAP_PT_E_INVAL, "auto_select: steps must be in [2, %d]", QAP_PT_AUTO_STEPS);
    if (R < 1) return qap::fail(QAP_PT_E_INVAL, "auto_select: R must be >= 1");
    auto beta_at = [](int j) { return std::exp2(QAP_PT_AUTO_MIN_EXP + (double)j); };
    // Only the exponents before the last one enter the rule (:434-446, j < i).
    const int last = steps - 1;
    double minE = 1e30, maxE = -1e30;
    int minIdx = 0;
    for (int j = 0; j < last; ++j) {
        if (avg_energy[j] < minE) {
            minE = avg_energy[j];
            minIdx = j;
        }
        if (j == 0) maxE = avg_energy[j];  // the hottest trial stands for random noise
    }
    double minBeta = 0.0, maxBeta = 0.0;
    for (int j = 0; j < last; ++j) {
        if (minBeta == 0.0 && ((j == minIdx - 1 && j >= 1) ||
                               avg_energy[j] < (maxE - minE) * higher_prop + minE))
            minBeta = beta_at(j);
        if (j == last - 1 || (minBeta != 0.0 && (j == minIdx ||
                                                 avg_energy[j] < (maxE - minE) * lower_prop + minE))) {
            if (minBeta == 0.0) minBeta = beta_at(0);
            maxBeta = beta_at(j + 2);
            break;
        }
    }
    // beta0 is kept in exp2 units by the reference; the ABI's ladders are in ln units.
    *beta0_ln = maxBeta / kLog2e;
    *ratio = std::pow(maxBeta / minBeta, 1.0 / R);
    return QAP_PT_OK;
}

int qap_pt_auto_beta(const qap_problem *prob, const qap_pt_params *base, double lower_prop,
                     double higher_prop, double *avg_energy, double *beta0_ln, double *ratio) {
    try {
        return auto_beta_impl(prob, base, lower_prop, higher_prop, avg_energy, beta0_ln, ratio);
    } catch (const std::bad_alloc &) {
        return qap::fail(QAP_PT_E_NOMEM, "host allocation failed");
    } catch (...) {
        return qap::fail(QAP_PT_E_INVAL, "unexpected host exception");
    }
}

}  // extern "C"
