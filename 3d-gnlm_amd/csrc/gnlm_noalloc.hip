// gnlm_noalloc.hip -- gnlm3d_forward_noalloc: the 3D forward for callers that
// cannot allocate device memory (code under stream capture, tight serving
// loops, harnesses that forbid hipMalloc).  It runs the same kernels as
// gnlm3d_forward and gives the same bits; the one workspace those kernels need
// (the rows variants' padded copy of the input, and for the large-patch tiles
// the running Z total when the caller passes no Z buffer) comes from a fixed
// device array instead of hipMallocAsync.
//
// The array is `scratch` below: every device the library runs on holds its
// own copy, at its own address, until the process ends.  Host code must never
// pass `scratch` itself to a kernel -- on the host that name is a stand-in,
// not a device address.  Its device address is looked up once per device
// (hipGetSymbolAddress, with that device current) and only that address is
// used.  All calls on one device share its array: calls that overlap on one
// device (concurrent streams, or host threads) race on it.  Nothing in it
// needs clearing between calls: the padding pass writes every float of the
// padded copy, and the kernel writes the Z total before it reads it.
#include <hip/hip_runtime.h>
#include <cmath>
#include <mutex>
#include "gnlm_internal.h"

namespace gnlm {
namespace {

static __device__ __align__(256) unsigned char scratch[GNLM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per device (NULL until looked up)

// Device address of the current device's `scratch` (looked up on the first
// call per device, with that device current); NULL if the lookup failed.
void *scratch_on_current_device(hipError_t *err)
{
    int dev = -1;
    *err = hipGetDevice(&dev);
    if (*err != hipSuccess) return nullptr;
    if (dev < 0 || dev >= kMaxDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (*err != hipSuccess || !p) {
            if (*err == hipSuccess) *err = hipErrorInvalidSymbol;
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace gnlm

using namespace gnlm;

extern "C" int gnlm3d_forward_noalloc(const float *d_in, float *d_out, float *d_Z, int nx, int ny, int nz, int r_p,
                                      int r_s, float h, void *stream)
{
    if (!d_in || !d_out) return GNLM_EINVAL;
    if (!(h > 0.0f) || !std::isfinite(h)) return GNLM_EINVAL;
    int kernel = 0;
    int st = gnlm3d_forward_route(nx, ny, nz, r_p, r_s, GNLM_PATH_AUTO, &kernel);
    if (st != GNLM_OK) return st;
    // the per-offset generic passes allocate their own buffers
    if (kernel == GNLM_KERNEL_GENERIC) return GNLM_EUNSUPPORTED;
    hipError_t e = hipSuccess;
    void *p = scratch_on_current_device(&e);
    if (!p) return GNLM_EHIP;
    const Workspace ws{p, (size_t)GNLM_NOALLOC_SCRATCH_BYTES};
    InDesc in{d_in, nx, ny, nz, 0, 0, 0};
    const int P = (2 * r_p + 1) * (2 * r_p + 1) * (2 * r_p + 1);
    const float k = (float)weight_scale(P, h);
    // the shape the forward recorded for this volume class, never a tuning run
    const int shape = fwd3d_pick_shape(in, d_out, d_Z, r_p, r_s, k, kernel, (hipStream_t)stream, false);
    e = fwd3d_tiled_launch_variant(in, d_out, d_Z, r_p, r_s, k, kernel, (hipStream_t)stream, &ws, shape);
    if (e == hipErrorOutOfMemory) return GNLM_ENOMEM;  // the volume needs more than the fixed scratch
    if (e == hipErrorInvalidValue) return GNLM_EINVAL;
    return e == hipSuccess ? GNLM_OK : GNLM_EHIP;
}
