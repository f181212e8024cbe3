// fm_signal_execute_noalloc: the signal evaluation with its workspace in a
// file-scope device array, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  The kernels and their launches are those of
// fm_signal_execute; only the workspace comes from here.
//
// The array is FM_SIGNAL_SCRATCH_BYTES long and every device the library runs
// on holds its own copy until the process ends.  Its device address is looked
// up once per device (hipGetSymbolAddress, with that device current) on the
// first call made on the device, and kept; the array's name is never used as
// an address on the host.  No kernel here assumes the workspace starts at
// zero (fm_signal_execute's callers pass uninitialised buffers too), so no
// clearing is needed.  All calls on one device share the array: calls that
// overlap on one device (several streams, several host threads) race on it.
#include <mutex>

#include "fm_common.h"

namespace fm {

static __device__ __align__(256) unsigned char scratch[FM_SIGNAL_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array, looked up on first use.
static int scratch_address(int dev, void **out)
{
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d is beyond the %d devices the scratch table covers", dev, kMaxDevices);
        return FM_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *ptr = nullptr;
        FM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (ptr == nullptr) {
            set_error("hipGetSymbolAddress returned a NULL scratch address on device %d", dev);
            return FM_ERR_HIP;
        }
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return FM_OK;
}

}  // namespace fm

using namespace fm;

extern "C" int fm_signal_execute_noalloc(const fm_signal_plan *p, int precision, int64_t n_nodes, int n_enc,
                                         const double *packed, const float *modes, int n_modes,
                                         const double *motion_coef, double *out, void *stream)
{
    FM_CHECK_ARG(p != nullptr, "plan is NULL");
    FM_CHECK_ARG(n_nodes >= 0 && n_enc >= 1, "bad sizes (n_nodes=%lld, n_enc=%d)", (long long)n_nodes, n_enc);
    const int moving = (modes != nullptr && n_modes > 0) ? 1 : 0;
    size_t need = 0;
    int rc = fm_signal_workspace_size(p, precision, n_nodes, n_enc, moving, &need);
    if (rc != FM_OK) return rc;
    if (need > (size_t)FM_SIGNAL_SCRATCH_BYTES) {
        set_error("fm_signal_execute_noalloc: this evaluation needs %zu bytes of workspace, more than the %zu-byte "
                  "scratch array; use fm_signal_execute with a caller workspace",
                  need, (size_t)FM_SIGNAL_SCRATCH_BYTES);
        return FM_ERR_WORKSPACE;
    }
    int dev = 0;
    FM_HIP(hipGetDevice(&dev));
    void *ws = nullptr;
    if ((rc = scratch_address(dev, &ws)) != FM_OK) return rc;
    return fm_signal_execute(p, precision, n_nodes, n_enc, packed, modes, n_modes, motion_coef, ws,
                             (size_t)FM_SIGNAL_SCRATCH_BYTES, out, stream);
}
