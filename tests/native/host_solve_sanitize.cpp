// Host solve under the host sanitizers (built by `make sanitize` in the package's csrc):
// qap_pt_solve_cpu on six seeded instances, symmetric and general, n = 12 .. 57, with one to
// three ladders of 8 replicas spread over the host threads.  Each result's best cost must
// equal the objective recomputed for its best permutation.  Exit status 0 only when every
// solve succeeds and agrees; ThreadSanitizer / AddressSanitizer report on stderr.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "../../include/qap_pt.h"

static int64_t cost(int n, const std::vector<int32_t> &F, const std::vector<int32_t> &D,
                    const std::vector<int32_t> &p) {
    int64_t e = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) e += (int64_t)D[i * n + j] * F[p[i] * n + p[j]];
    return e;
}

int main() {
    int fails = 0;
    std::mt19937 rng(2024);
    for (int t = 0; t < 6; ++t) {
        const int n = 12 + 9 * t, R = 8, T = 1 + (t % 3);
        std::vector<int32_t> F(n * n), D(n * n), best(n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                F[i * n + j] = (i == j) ? 0 : (int32_t)(rng() % 50);
                D[i * n + j] = (i == j) ? 0 : (int32_t)(rng() % 40);
            }
        if (t % 2 == 0)  // symmetric instances take the symmetric update path
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < i; ++j) {
                    F[i * n + j] = F[j * n + i];
                    D[i * n + j] = D[j * n + i];
                }
        std::vector<double> beta(R);
        for (int r = 0; r < R; ++r) beta[r] = 0.05 * std::pow(10.0, r / (double)(R - 1));
        qap_problem pb{n, F.data(), D.data()};
        qap_pt_params pr{};
        pr.R = R;
        pr.T = T;
        pr.beta_ln = beta.data();
        pr.sweeps = 30;
        pr.sweeps_per_exchange = 1;
        pr.seed = 77 + t;
        pr.kick_after = 100;
        qap_pt_result res{};
        res.best_perm = best.data();
        const int rc = qap_pt_solve_cpu(&pb, &pr, &res);
        const bool ok = rc == 0 && res.best_cost == cost(n, F, D, best);
        printf("n=%d ladders=%d rc=%d best=%lld accepted=%llu %s\n", n, T, rc,
               (long long)res.best_cost, (unsigned long long)res.accepted, ok ? "ok" : "MISMATCH");
        fails += !ok;
    }
    return fails;
}
