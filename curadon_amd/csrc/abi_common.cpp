// Status / error plumbing and device checks shared by every entry point of libcuradon_mi.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "cmi_internal.hpp"

namespace {
thread_local std::string g_last_error;
}

namespace cmi {

cmi_status fail(cmi_status s, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return s;
}

void clear_error() { g_last_error.clear(); }

cmi_status check_device_ptr(const void *p, int device, const char *what) {
    hipPointerAttribute_t attr{};
    const hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(CMI_ERR_DEVICE, "%s is not device memory known to HIP (%s)", what,
                    hipGetErrorString(e));
    }
    if (attr.type == hipMemoryTypeDevice && attr.device != device)
        return fail(CMI_ERR_DEVICE, "%s lives on device %d but the plan is on device %d", what,
                    attr.device, device);
    if (attr.type == hipMemoryTypeUnregistered)
        return fail(CMI_ERR_DEVICE, "%s is pageable host memory", what);
    return CMI_OK;
}

cmi_status DeviceGuard::enter(int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return fail(CMI_ERR_DEVICE, "no HIP device available (%s)",
                    e == hipSuccess ? "device count is 0" : hipGetErrorString(e));
    }
    if (device < 0 || device >= n)
        return fail(CMI_ERR_DEVICE, "device %d out of range (%d devices)", device, n);
    e = hipGetDevice(&prev);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(CMI_ERR_HIP, "hipGetDevice failed: %s", hipGetErrorString(e));
    }
    if (prev != device) {
        e = hipSetDevice(device);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(CMI_ERR_HIP, "hipSetDevice(%d) failed: %s", device, hipGetErrorString(e));
        }
        changed = true;
    }
    return CMI_OK;
}

DeviceGuard::~DeviceGuard() {
    if (changed) (void)hipSetDevice(prev);
}

} // namespace cmi

extern "C" {

CMI_API const char *cmi_last_error(void) { return g_last_error.c_str(); }

CMI_API int cmi_version(void) { return 100; }

} // extern "C"
