// qap_pt_noalloc.hip -- qap_pt_solve_hip_noalloc: the solve of qap_pt_solve_hip for callers
// that cannot allocate device memory.
//
// Every device buffer of the solve (instance data, replica states, force matrices, bucket
// lists, counters) is a 256-byte-aligned piece of one fixed region, `scratch`, instead of
// a hipMalloc'ed buffer; the kernels are the same ones the allocating path runs, launched
// one by one on the caller's stream (no graph, event or side stream is created), so the
// results are bit-identical to qap_pt_solve_hip's.
//
// In host code `scratch` names a host-side stand-in, not device memory: the device address
// of this GPU's copy is looked up once per device (hipGetSymbolAddress, with that device
// current) and only that address is handed to copies and kernels.
//
// The region keeps the last call's contents.  A call writes every piece it reads before the
// kernels run (uploads, the zeroed force matrices, counters and bests set from the host),
// so nothing depends on what a previous call left.  Two calls on one GPU at the same time
// would race on the region; a per-device lock serialises them within the process.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/qap_pt.h"
#include "qap_pt_internal.h"
#include "qap_pt_pool_impl.h"

static_assert(QAP_PT_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size must be 256-byte aligned");

// The fixed device region, one copy per GPU, held until the process ends.
static __device__ __align__(256) unsigned char scratch[QAP_PT_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_dev_mu[kMaxDevices];
void *g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each GPU

}  // namespace

extern "C" int qap_pt_solve_hip_noalloc(const qap_problem *prob, const qap_pt_params *par,
                                        qap_pt_result *res, void *stream) {
    using qap::fail;
    if (!prob || !par || !res) return fail(QAP_PT_E_INVAL, "NULL argument");
    int ndev = 0, dev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(QAP_PT_E_DEVICE, "no HIP device visible");
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return fail(QAP_PT_E_DEVICE, "no usable current HIP device");
    std::lock_guard<std::mutex> lock(g_dev_mu[dev]);
    if (!g_scratch_ptr[dev]) {
        // first call on this device (it is current): the address of its copy of the region
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(QAP_PT_E_DEVICE, "hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    qap_pt_pool *p = nullptr;
    int rc = qap::pool_create_in_arena(1, prob, par, g_scratch_ptr[dev],
                                       (size_t)QAP_PT_NOALLOC_SCRATCH_BYTES,
                                       static_cast<hipStream_t>(stream), &p);
    if (rc) return rc;
    rc = qap_pt_pool_run(p, -1, nullptr, nullptr);
    if (rc == QAP_PT_OK) rc = qap_pt_pool_result(p, 0, res);
    qap_pt_pool_destroy(p);
    return rc;
}
