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

}  // namespace qap
