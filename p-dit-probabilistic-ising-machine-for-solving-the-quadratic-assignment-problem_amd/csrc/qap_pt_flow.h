// qap_pt_flow.h -- structure of a flow matrix that the sweep kernels exploit, computed
// once per instance on the host (plain C++, no device code).
#pragma once

#include <cstdint>
#include <vector>

namespace qap {

// Equal id <=> identical flow column and row: a swap of the two facilities leaves every
// flow difference of the rank-2 update zero.
std::vector<int32_t> facility_classes(int n, const int32_t *F);
// Twin classes of a symmetric zero-diagonal flow matrix (a ~ b iff the rows agree outside
// columns a and b); with sym false, or n < 3, every facility is alone.
std::vector<int32_t> facility_twins(int n, const int32_t *F, bool sym);
// Nonzero columns of a facility's flow row and column, lo | hi << 10 | count << 20.
int32_t facility_span(int n, const int32_t *F, int f);
// Fraction of nonzero off-diagonal flow entries.
double offdiag_density(int n, const int32_t *F);

}  // namespace qap
