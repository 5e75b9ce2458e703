// Flat-index decomposition of a zero-bordered stack of rows [count][Vp][Up] (x fastest), shared
// by the staging kernels (panels of the cone-beam backprojector, served projections) and by a
// host-side unit test that compiles this header with a plain C++ compiler.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define CMI_HD __host__ __device__
#else
#define CMI_HD
#endif

namespace cmi {

struct StageIndex {
    uint64_t a; // index over [count] (row of rows: angle, or batch * A + angle)
    int vp, up; // row within the slice and column within the row
};

// Element i of [count][Vp][Up]. `narrow` selects 32-bit divisions, valid when every index the
// caller decomposes is below 2^32; otherwise 64-bit divisions.
CMI_HD inline StageIndex stage_index(uint64_t i, int Up, int Vp, bool narrow) {
    StageIndex r;
    if (narrow) {
        const uint32_t j = (uint32_t)i, r0 = j / (uint32_t)Up; // a * Vp + vp
        r.up = (int)(j - r0 * (uint32_t)Up);
        r.vp = (int)(r0 % (uint32_t)Vp);
        r.a = r0 / (uint32_t)Vp;
    } else {
        const uint64_t r0 = i / (uint64_t)Up;
        r.up = (int)(i - r0 * (uint64_t)Up);
        r.vp = (int)(r0 % (uint64_t)Vp);
        r.a = r0 / (uint64_t)Vp;
    }
    return r;
}

// The index of the element after `s`, stepping to the next row and slice at their ends.
CMI_HD inline void stage_next(StageIndex &s, int Up, int Vp) {
    if (++s.up == Up) {
        s.up = 0;
        if (++s.vp == Vp) {
            s.vp = 0;
            ++s.a;
        }
    }
}

} // namespace cmi
