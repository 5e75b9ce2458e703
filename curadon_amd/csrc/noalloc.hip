// cmi_forward_2d_noalloc: the 2D fan-beam forward projection without any allocation, for
// callers that cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc). Its staging scratch is the file-scope device array
// below instead of a stream-ordered allocation.
//
// The array is a module global: each GPU the library runs on holds its own copy at its own
// address for the life of the process. In host code its name is a host-side stand-in, not a
// device address, so host code never passes it to a kernel or a copy; the device address is
// taken once per device with hipGetSymbolAddress (with that device current) and kept.
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU race. The
// staging kernel writes every element a call reads (zero borders included), so nothing needs
// clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "cmi_internal.hpp"

using namespace cmi;

namespace {

constexpr size_t kNoAllocScratchBytes = (size_t)CMI_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current.
cmi_status scratch_on(int device, void **out) {
    if (device < 0 || device >= kMaxDevices)
        return fail(CMI_ERR_DEVICE, "forward_2d_noalloc: device %d beyond %d", device,
                    kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(CMI_ERR_HIP, "forward_2d_noalloc: hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(e));
        }
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return CMI_OK;
}

} // namespace

extern "C" {

CMI_API cmi_status cmi_forward_2d_noalloc(const cmi_plan2d *plan, const void *volume,
                                          int vol_dtype, void *sinogram, int sino_dtype,
                                          uint64_t batch, void *stream) {
    clear_error();
    CMI_TRY(fan2d_forward_check(plan, volume, vol_dtype, sinogram, sino_dtype, batch));
    const size_t need = fan2d_forward_scratch_bytes(plan, vol_dtype, batch);
    if (need > kNoAllocScratchBytes)
        return fail(CMI_ERR_INVALID_ARGUMENT,
                    "forward_2d_noalloc: the call needs %zu bytes of staging scratch (two "
                    "zero-bordered copies per image), above the library's static maximum of "
                    "%zu bytes; use cmi_forward_2d",
                    need, kNoAllocScratchBytes);
    const int device = fan2d_plan_device(plan);
    DeviceGuard guard;
    CMI_TRY(guard.enter(device));
    void *ptr = nullptr;
    CMI_TRY(scratch_on(device, &ptr));
    return fan2d_forward_enqueue(plan, volume, vol_dtype, sinogram, sino_dtype, batch, ptr,
                                 (hipStream_t)stream);
}

} // extern "C"
