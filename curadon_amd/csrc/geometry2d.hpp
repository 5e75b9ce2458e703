// FP64 2D fan-beam geometry: per-angle plan and per-ray march set-up.
//
// The per-ray set-up decides discrete quantities (miss or hit, sample count N) that must be
// identical for the forward projector and its adjoint, so it is evaluated once per plan in
// FP64 and stored in the ray table; the kernels only march. Contraction into FMA is turned
// off so the FP64 sequence is the plain IEEE one.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#pragma float_control(push)
#pragma clang fp contract(off)

namespace cmi::fan2d {

// Per-angle geometry in texel coordinates of the image (texel i centred at i + 0.5):
// src = source position, org = vector from the source to detector cell 0's centre,
// del = step from cell k to cell k + 1.
struct AngleGeom {
    double src[2];
    double org[2];
    double del[2];
};

struct Scalars {
    int64_t nx, ny, D;
    double sx, sy, ox, oy;
    double du, det_offset;
};

// Clockwise rotation by the angle with cosine c and sine s, then translation by -t
// (include/curadon/detail/rotation.hpp:51-61).
__host__ __device__ inline void rot_cw(const double v[2], double c, double s, const double t[2],
                                       double out[2]) {
    out[0] = v[0] * c + v[1] * s - t[0];
    out[1] = -v[0] * s + v[1] * c - t[1];
}

// plan_2d::precompute (include/curadon/detail/plan/plan_2d.hpp:17-80, 189-241) in FP64.
inline AngleGeom plan_angle(const Scalars &g, double angle, double dso, double dsd) {
    AngleGeom r{};
    const double c = std::cos(angle), s = std::sin(angle);
    const double det_extent = (double)g.D * g.du;
    const double ext[2] = {(double)g.nx * g.sx, (double)g.ny * g.sy};
    const double sp[2] = {g.sx, g.sy}, off[2] = {g.ox, g.oy}, zero[2] = {0.0, 0.0};
    const double dod = dsd - dso;
    const double s0[2] = {g.det_offset, dso};
    double sw[2];
    rot_cw(s0, c, s, zero, sw);
    const double p0[2] = {-det_extent / 2.0 + g.du * 0.5, -dod};
    const double p1[2] = {-det_extent / 2.0 + g.du * 1.5, -dod};
    double r0[2], r1[2];
    rot_cw(p0, c, s, sw, r0);
    rot_cw(p1, c, s, sw, r1);
    for (int d = 0; d < 2; ++d) {
        const double o = r0[d] / sp[d];
        r.org[d] = o;
        r.del[d] = r1[d] / sp[d] - o;
        r.src[d] = (sw[d] - off[d] + 0.5 * ext[d]) / sp[d];
    }
    return r;
}

// ASTRA-like view (no counterpart in the reference, whose from_vec keeps only DSO / DSD /
// angle, python/standalone/src/curadon/geometry.py:84-133): source s, detector centre c and cell
// step u in world coordinates, v = (s.x, s.y, c.x, c.y, u.x, u.y). Cell 0's centre is
// c - (D - 1)/2 u; the texel-frame quantities follow as in plan_angle.
inline AngleGeom plan_angle_vec(const Scalars &g, const double v[6]) {
    AngleGeom r{};
    const double ext[2] = {(double)g.nx * g.sx, (double)g.ny * g.sy};
    const double sp[2] = {g.sx, g.sy}, off[2] = {g.ox, g.oy};
    const double h = ((double)g.D - 1.0) / 2.0;
    for (int d = 0; d < 2; ++d) {
        const double r0 = v[2 + d] - h * v[4 + d] - v[d]; // source to cell 0
        const double r1 = r0 + v[4 + d];                   // source to cell 1
        const double o = r0 / sp[d];
        r.org[d] = o;
        r.del[d] = r1 / sp[d] - o;
        r.src[d] = (v[d] - off[d] + 0.5 * ext[d]) / sp[d];
    }
    return r;
}

// The view of a circular plan with exact geometry (cmi_plan2d_create_ex, CMI_PLAN2D_EXACT_GEOMETRY;
// no counterpart in the reference, which moves the source by det_offset in its forward only and
// stores det_pitch and COR unused, plan_2d.hpp:76-80, 121-122): before the rotation by the
// angle, the source sits at (-COR, DSO), the detector centre at (det_offset - COR,
// -(DSD - DSO)) and the cell step is du (cos pitch, sin pitch); all three are then rotated
// clockwise by the angle like the reference's points (rotation.hpp:51-61). So det_offset moves
// the detector along its own axis, pitch turns it about its centre, and the rotation centre lies
// COR along the detector axis from the line through the source and the detector centre.
inline void exact_view(const Scalars &g, double angle, double dso, double dsd, double pitch,
                       double cor, double v[6]) {
    const double c = std::cos(angle), s = std::sin(angle), zero[2] = {0.0, 0.0};
    const double s0[2] = {-cor, dso};
    const double d0[2] = {g.det_offset - cor, -(dsd - dso)};
    const double u0[2] = {g.du * std::cos(pitch), g.du * std::sin(pitch)};
    rot_cw(s0, c, s, zero, v);
    rot_cw(d0, c, s, zero, v + 2);
    rot_cw(u0, c, s, zero, v + 4);
}

// Backprojection constants of a view (the reference's formula, backward_2d.hpp:58-99, for an
// arbitrary view): with n the unit normal of the detector line pointing from the source towards
// it and e = u / |u|, a pixel at world x has depth n . x + n0 (n0 = -n . s), magnification
// w = L / depth (L = n . (c - s)), detector cell coordinate w (e . x + e0) + c0 with
// e0 = -e . s / |u| and c0 = e . (s - c) / |u| + D/2 (the e terms scaled by 1 / |u|), and
// weight w^p / |u|. For a centred circular view these are the reference's quantities:
// n . x + n0 = xi + DSO, L = DSD, e . x / |u| = yi, c0 = D/2.
struct BpView {
    double n[2], n0, L;    // depth = n . x + n0; w = L / depth
    double e[2], e0, c0;   // cell = w (e . x + e0) + c0
    double inv_du;         // 1 / |u|
};

inline BpView bp_view(const double v[6], int64_t D) {
    BpView b{};
    const double ul = std::sqrt(v[4] * v[4] + v[5] * v[5]);
    const double e0 = v[4] / ul, e1 = v[5] / ul;
    double n0 = -e1, n1 = e0;
    const double cs0 = v[2] - v[0], cs1 = v[3] - v[1]; // source to detector centre
    if (n0 * cs0 + n1 * cs1 < 0.0) {
        n0 = -n0;
        n1 = -n1;
    }
    b.n[0] = n0;
    b.n[1] = n1;
    b.n0 = -(n0 * v[0] + n1 * v[1]);
    b.L = n0 * cs0 + n1 * cs1;
    b.e[0] = e0 / ul;
    b.e[1] = e1 / ul;
    b.e0 = -(e0 * v[0] + e1 * v[1]) / ul;
    b.c0 = -(e0 * cs0 + e1 * cs1) / ul + (double)D / 2.0;
    b.inv_du = 1.0 / ul;
    return b;
}

// One ray after clipping, in the form the kernels consume.
struct Ray {
    double e[2];    // entry point, texel coordinates
    double v[2];    // step; the dominant component is exactly +-1
    double scale;   // world length of one step
    int64_t n;      // number of samples; 0 for a missed ray
};

// Slab clip against [-1, n] per axis (include/curadon/detail/intersection.h:30-46), the miss
// test tmin > tmax - 1e-6 (forward_2d.hpp:58) and the march set-up (forward_2d.hpp:63-77).
__host__ __device__ inline Ray setup_ray(const AngleGeom &ag, int64_t nx, int64_t ny, double sx,
                                         double sy, int64_t k) {
    Ray r{};
    const double rd[2] = {ag.org[0] + (double)k * ag.del[0], ag.org[1] + (double)k * ag.del[1]};
    const double hi[2] = {(double)nx, (double)ny};
    double tmin = FLT_MIN, tmax = FLT_MAX;
    for (int d = 0; d < 2; ++d) {
        const double t1 = (-1.0 - ag.src[d]) / rd[d];
        const double t2 = (hi[d] - ag.src[d]) / rd[d];
        tmin = fmin(fmax(t1, tmin), fmax(t2, tmin));
        tmax = fmax(fmin(t1, tmax), fmin(t2, tmax));
    }
    if (tmin > tmax - 1e-6) return r; // n = 0
    const double q[2] = {rd[0] * (tmax - tmin), rd[1] * (tmax - tmin)};
    const double m = fmax(fabs(q[0]), fabs(q[1]));
    r.e[0] = ag.src[0] + tmin * rd[0];
    r.e[1] = ag.src[1] + tmin * rd[1];
    r.v[0] = q[0] / m;
    r.v[1] = q[1] / m;
    r.scale = hypot(r.v[0] * sx, r.v[1] * sy);
    r.n = (int64_t)rint(m); // round half to even (forward_2d.hpp:66)
    return r;
}

} // namespace cmi::fan2d

#pragma float_control(pop)
