// 3D cone-beam projectors for gfx950 (placeholder entry points until the kernels land).
#include <hip/hip_runtime.h>

#include "cmi_internal.hpp"

using namespace cmi;

struct cmi_plan3d {
    int device = 0;
};

extern "C" {

CMI_API cmi_status cmi_plan3d_create(int, int, int, const cmi_cone3d_desc *, cmi_plan3d **plan) {
    clear_error();
    if (plan) *plan = nullptr;
    return fail(CMI_ERR_NOT_SUPPORTED, "plan3d_create: 3D projectors are not built yet");
}

CMI_API cmi_status cmi_plan3d_destroy(cmi_plan3d *plan) {
    clear_error();
    delete plan;
    return CMI_OK;
}

CMI_API cmi_status cmi_forward_3d(const cmi_plan3d *, const void *, int, void *, int, uint64_t,
                                  void *) {
    clear_error();
    return fail(CMI_ERR_NOT_SUPPORTED, "forward_3d: not built yet");
}

CMI_API cmi_status cmi_backward_3d(const cmi_plan3d *, void *, int, const void *, int, uint64_t,
                                   int, void *) {
    clear_error();
    return fail(CMI_ERR_NOT_SUPPORTED, "backward_3d: not built yet");
}

CMI_API cmi_status cmi_adjoint_3d(const cmi_plan3d *, void *, int, const void *, int, uint64_t,
                                  void *) {
    clear_error();
    return fail(CMI_ERR_NOT_SUPPORTED, "adjoint_3d: not built yet");
}

} // extern "C"
