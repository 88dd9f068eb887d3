// Checks the host's flow-structure helpers (csrc/qap_pt_flow.cpp) against brute force:
// twin classes (rows equal outside the two facilities' own columns) on random symmetric
// zero-diagonal matrices with planted twins, facility classes (identical row and
// column) on random general matrices, and the grey pattern.  Prints "ok" or the first
// disagreement; exit status 0 only when everything agrees.
#include <cstdio>
#include <random>
#include <vector>

#include "qap_pt_flow.h"

using qap::facility_classes;
using qap::facility_twins;

static bool twins_brute(int n, const std::vector<int32_t> &F, int a, int b) {
    for (int x = 0; x < n; ++x)
        if (x != a && x != b && F[a * n + x] != F[b * n + x]) return false;
    return true;
}

static bool same_class_brute(int n, const std::vector<int32_t> &F, int a, int b) {
    for (int x = 0; x < n; ++x)
        if (F[a * n + x] != F[b * n + x] || F[x * n + a] != F[x * n + b]) return false;
    return true;
}

int main() {
    std::mt19937 rng(12345);
    long pairs = 0;
    for (int it = 0; it < 3000; ++it) {
        const int n = 2 + (int)(rng() % 14), vmax = 1 + (int)(rng() % 3);
        std::vector<int32_t> F(n * n, 0);
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) F[i * n + j] = F[j * n + i] = (int32_t)(rng() % (vmax + 1));
        const int plant = (int)(rng() % 4);
        for (int k = 0; k < plant; ++k) {  // copy row/column a onto b outside {a, b}
            const int a = (int)(rng() % n), b = (int)(rng() % n);
            if (a == b) continue;
            for (int x = 0; x < n; ++x)
                if (x != a && x != b) F[b * n + x] = F[x * n + b] = F[a * n + x];
        }
        const std::vector<int32_t> tw = facility_twins(n, F.data(), true);
        const std::vector<int32_t> alone = facility_twins(n, F.data(), false);
        for (int a = 0; a < n; ++a) {
            if (alone[a] != a) {
                std::printf("non-symmetric call grouped facility %d\n", a);
                return 1;
            }
            for (int b = 0; b < n; ++b) {
                if (a == b) continue;
                const bool r = twins_brute(n, F, a, b);
                // sound always; complete from n = 3 on (smaller instances are left alone)
                if ((tw[a] == tw[b] && !r) || (n >= 3 && r && tw[a] != tw[b])) {
                    std::printf("twins n=%d a=%d b=%d brute=%d ids %d %d\n", n, a, b, r, tw[a], tw[b]);
                    return 1;
                }
                ++pairs;
            }
        }
        // facility classes on a general (asymmetric, nonzero diagonal) matrix
        std::vector<int32_t> G(n * n);
        for (auto &g : G) g = (int32_t)(rng() % 2);
        for (int k = 0; k < plant; ++k) {
            const int a = (int)(rng() % n), b = (int)(rng() % n);
            for (int x = 0; x < n; ++x) {
                G[b * n + x] = G[a * n + x];
                G[x * n + b] = G[x * n + a];
            }
            G[b * n + b] = G[a * n + a];
            G[a * n + b] = G[b * n + a] = G[a * n + a];
        }
        const std::vector<int32_t> cl = facility_classes(n, G.data());
        for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b)
                if ((cl[a] == cl[b]) != same_class_brute(n, G, a, b)) {
                    std::printf("classes n=%d a=%d b=%d\n", n, a, b);
                    return 1;
                }
    }
    // grey pattern: one twin class for the 92-facility block, one for the rest
    const int n = 256;
    std::vector<int32_t> F(n * n, 0);
    for (int i = 0; i < 92; ++i)
        for (int j = 0; j < 92; ++j)
            if (i != j) F[i * n + j] = 1;
    const std::vector<int32_t> tw = facility_twins(n, F.data(), true);
    for (int a = 0; a < n; ++a)
        if ((tw[a] == tw[0]) != (a < 92) || (tw[a] == tw[255]) != (a >= 92)) {
            std::printf("grey facility %d in the wrong twin class\n", a);
            return 1;
        }
    std::printf("ok %ld twin pairs\n", pairs);
    return 0;
}
