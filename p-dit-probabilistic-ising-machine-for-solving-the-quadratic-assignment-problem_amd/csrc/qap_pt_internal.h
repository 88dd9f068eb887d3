// qap_pt_internal.h -- pool entry points shared by the C-ABI front ends inside the
// library (the solve/pool calls and the serving loop).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/qap_pt.h"
#include "qap_pt_device.h"

namespace qap {

// Sets qap_pt_last_error() for this thread and returns code.
int fail(int code, const char *fmt, ...);

// Checks one instance's arguments (index k in messages): QAP_PT_OK, or the error code with
// qap_pt_last_error() set.  Shared by every solve entry point.
int validate(const qap_problem &pb, const qap_pt_params &pr, int k);

int pool_create_internal(int32_t count, const qap_problem *problems, const qap_pt_params *params,
                         qap_pt_pool **out);
// Enqueues every instance's remaining sweeps (params.sweeps) on the pool stream as
// graph launches and records `done` after them; returns without waiting.
int pool_launch_async(qap_pt_pool *p, hipEvent_t done);
int pool_result_internal(qap_pt_pool *p, int inst, qap_pt_result *res);
void pool_destroy_internal(qap_pt_pool *p);
// Whether the instance's bounds need int64 force matrices (the int64 in-HBM placement).
bool wide_instance(const qap_problem &pb);
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
// The tuned row for an instance of size n in placement im, or nullptr: from the full-chip
// table, or with `latency` from the table of latency-bound pools (at most one replica per
// CU).  The tables are read once per process from $QAP_PT_TUNE_TABLE and
// $QAP_PT_TUNE_LATENCY_TABLE (empty: no table) or else from <directory of
// libqap_pt.so>/../tune/shapes_gfx950.txt and shapes_latency_gfx950.txt; a missing file
// means no table.
const TunedShape *tuned_shape(int n, bool sparse, int im, bool latency = false);

// Slot mode (qap_pt_slots.cpp): a pool of `slots` fixed-capacity instance slots that the
// serving loop loads and releases between launches.  Every call except slots_launch
// expects the previous launch on the pool stream to have finished.
int64_t slots_device_bytes(int32_t slots, int32_t n_cap, int32_t r_cap, int32_t t_cap,
                           int32_t rt_cap);
int slots_create(int32_t slots, int32_t n_cap, int32_t r_cap, int32_t t_cap, int32_t rt_cap,
                 qap_pt_pool **out);
bool slots_fit(const qap_pt_pool *p, const qap_problem &pb, const qap_pt_params &pr);
int slots_free(const qap_pt_pool *p);    // a free slot, or -1
int slots_in_use(const qap_pt_pool *p);
// Enqueues the instance's uploads, force matrices and costs; it joins the next launch.
int slots_load(qap_pt_pool *p, int slot, const qap_problem &pb, const qap_pt_params &pr);
// Enqueues `intervals` exchange intervals of every loaded slot and records `done`.
int slots_launch(qap_pt_pool *p, int intervals, hipEvent_t done);
// Slots whose run has ended (limit reached, goal, halted); state receives every slot's.
int slots_finished(qap_pt_pool *p, std::vector<int> &done, std::vector<InstState> &state);
int slots_halt(qap_pt_pool *p, int slot);
int slots_release(qap_pt_pool *p, int slot);

}  // namespace qap
