// qap_pt_internal.h -- pool entry points shared by the C-ABI front ends inside the
// library (the solve/pool calls and the serving loop).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/qap_pt.h"

namespace qap {

// Sets qap_pt_last_error() for this thread and returns code.
int fail(int code, const char *fmt, ...);

int pool_create_internal(int32_t count, const qap_problem *problems, const qap_pt_params *params,
                         qap_pt_pool **out);
// Enqueues every instance's remaining sweeps (params.sweeps) on the pool stream as
// graph launches and records `done` after them; returns without waiting.
int pool_launch_async(qap_pt_pool *p, hipEvent_t done);
int pool_result_internal(qap_pt_pool *p, int inst, qap_pt_result *res);
void pool_destroy_internal(qap_pt_pool *p);
// Device bytes a pool of these instances allocates.
int64_t pool_device_bytes(int32_t count, const qap_problem *problems, const qap_pt_params *params);

// Instances whose off-diagonal flow density is below this use the sparse-flow update and
// the "sparse" rows of the tuned shape table.
constexpr double kSparseFlow = 0.35;

// One row of the tuned shape table: instances with n <= n_max (and above the previous
// row's n_max) of this flow class whose force matrix sits in placement im (IMode) use
// nt threads x cpl candidates per thread; rate is the tuned swap-attempts/s (0 = none).
struct TunedShape {
    int n_max, sparse, im, nt, cpl;
    double rate;
};
// The tuned row for an instance of size n in placement im, or nullptr.  The table is read once per
// process from $QAP_PT_TUNE_TABLE (empty: no table) or else from
// <directory of libqap_pt.so>/../tune/shapes_gfx950.txt; a missing file means no table.
const TunedShape *tuned_shape(int n, bool sparse, int im);

}  // namespace qap
