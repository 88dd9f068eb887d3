// qap_pt_flow.cpp -- per-instance flow structure for the sweep kernels: facility classes
// (swaps that change nothing), twin classes (swaps that change two force columns), and
// the nonzero column span of each facility.
#include "qap_pt_flow.h"

#include <algorithm>
#include <map>

namespace qap {

// Facility classes: equal id <=> identical flow column and row, i.e. swapping the two
// facilities leaves every flow difference v1, v2 of the rank-2 update zero.
std::vector<int32_t> facility_classes(int n, const int32_t *F) {
    std::map<std::vector<int32_t>, int32_t> ids;
    std::vector<int32_t> cls(n);
    std::vector<int32_t> key(2 * (size_t)n);
    for (int a = 0; a < n; ++a) {
        for (int c = 0; c < n; ++c) {
            key[c] = F[(int64_t)c * n + a];
            key[n + c] = F[(int64_t)a * n + c];
        }
        auto it = ids.find(key);
        if (it == ids.end()) it = ids.emplace(key, (int32_t)ids.size()).first;
        cls[a] = it->second;
    }
    return cls;
}

// Twin classes of a symmetric zero-diagonal flow matrix: a ~ b iff F[a][x] == F[b][x] for
// every x outside {a, b}.  For such F this is an equivalence relation (a ~ b and b ~ c
// give F[a][b] = F[b][a] = F[c][a] = F[a][c] = F[b][c] = F[c][b], so rows a and c also
// agree at b), and a swap of twins changes the flow differences v = 2 (F[b] - F[a]) at
// columns a and b only.  Candidates come from a 64-bit column hash (twins share
// H(a) + f h(a), f = F[a][b]); every merge is proved with an exact row comparison, so
// a hash collision costs time, never correctness.  Other matrices: every facility alone.
std::vector<int32_t> facility_twins(int n, const int32_t *F, bool sym) {
    std::vector<int32_t> root(n);
    for (int a = 0; a < n; ++a) root[a] = a;
    if (!sym || n < 3) return root;
    auto find = [&](int x) {
        while (root[x] != x) x = root[x] = root[root[x]];
        return x;
    };
    auto mix = [](uint64_t z) {  // splitmix64 finaliser
        z += 0x9e3779b97f4a7c15ull;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    std::vector<uint64_t> h(n), H(n, 0);
    for (int x = 0; x < n; ++x) h[x] = mix((uint64_t)x);
    for (int a = 0; a < n; ++a)
        for (int x = 0; x < n; ++x) H[a] += (uint64_t)(int64_t)F[(int64_t)a * n + x] * h[x];
    struct Cand {
        int32_t f;
        uint64_t key;
        int32_t a;
    };
    std::vector<Cand> cand;
    std::vector<int32_t> vals;
    for (int a = 0; a < n; ++a) {
        vals.assign(F + (int64_t)a * n, F + (int64_t)(a + 1) * n);
        vals.erase(vals.begin() + a);  // the pair value F[a][b] over b != a
        std::sort(vals.begin(), vals.end());
        vals.erase(std::unique(vals.begin(), vals.end()), vals.end());
        for (int32_t f : vals) cand.push_back(Cand{f, H[a] + (uint64_t)(int64_t)f * h[a], a});
    }
    std::sort(cand.begin(), cand.end(), [](const Cand &x, const Cand &y) {
        return x.f != y.f ? x.f < y.f : (x.key != y.key ? x.key < y.key : x.a < y.a);
    });
    auto twins = [&](int a, int b) {
        const int32_t *ra = F + (int64_t)a * n, *rb = F + (int64_t)b * n;
        for (int x = 0; x < n; ++x)
            if (x != a && x != b && ra[x] != rb[x]) return false;
        return true;
    };
    for (size_t g = 0; g < cand.size();) {
        size_t e = g + 1;
        while (e < cand.size() && cand[e].f == cand[g].f && cand[e].key == cand[g].key) ++e;
        for (size_t m = g + 1; m < e; ++m) {
            const int r = find(cand[g].a), q = find(cand[m].a);
            if (r != q && twins(cand[g].a, cand[m].a)) root[std::max(r, q)] = std::min(r, q);
        }
        g = e;
    }
    for (int a = 0; a < n; ++a) root[a] = find(a);
    return root;
}

// Nonzero columns of facility f's flow row and column, packed lo | hi << 10 | count << 20
// (n <= 1024).  The flow differences v of a swap of a and b are zero outside
// [min(lo_a, lo_b), max(hi_a, hi_b)] and have at most count_a + count_b nonzeros.
int32_t facility_span(int n, const int32_t *F, int f) {
    int lo = 1023, hi = 0, cnt = 0;
    for (int c = 0; c < n; ++c)
        if (F[(int64_t)f * n + c] != 0 || F[(int64_t)c * n + f] != 0) {
            lo = std::min(lo, c);
            hi = std::max(hi, c);
            ++cnt;
        }
    return lo | (hi << 10) | (cnt << 20);
}

double offdiag_density(int n, const int32_t *F) {
    if (n < 2) return 1.0;
    int64_t nz = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) nz += (i != j) && F[(int64_t)i * n + j] != 0;
    return (double)nz / ((double)n * (n - 1));
}

}  // namespace qap
