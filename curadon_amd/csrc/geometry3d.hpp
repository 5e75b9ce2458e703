// FP64 3D cone-beam geometry: per-angle plan (forward and backprojector frames) and the
// per-ray march set-up. Contraction into FMA is off so the FP64 sequence is the plain IEEE one
// and the sample counts match an FP64 evaluation of the same formulas on the host.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#pragma float_control(push)
#pragma clang fp contract(off)

namespace cmi::cone3d {

struct Desc {
    int64_t nx, ny, nz;       // volume (x, y, z), x fastest
    double sx, sy, sz, ox, oy, oz;
    int64_t nu, nv;           // panel (u, v), u fastest
    double du, dv, off_u, off_v;
    double pitch, roll, yaw;  // detector rotation
    double dso, dsd, cor;
};

// Per angle, forward frame in voxel coordinates (voxel i centred at i): detector pixel (0,0),
// the u- and v-steps and the source.
struct FwdGeom {
    double org[3], du[3], dv[3], src[3];
};

// rotate_yzy (include/curadon/detail/rotation.hpp:7-28): the matrix written in its comment.
inline void rot_yzy(const double v[3], double phi, double theta, double psi, double out[3]) {
    const double cf = std::cos(phi), sf = std::sin(phi), ct = std::cos(theta),
                 st = std::sin(theta), cp = std::cos(psi), sp = std::sin(psi);
    const double m[3][3] = {
        {cf * cp * ct - sf * sp, -st * cf, sf * cp + sp * cf * ct},
        {st * cp, ct, sp * st},
        {-sf * cp * ct - sp * cf, sf * st, -sf * sp * ct + cp * cf},
    };
    for (int r = 0; r < 3; ++r) out[r] = m[r][0] * v[0] + m[r][1] * v[1] + m[r][2] * v[2];
}

// rotate_roll_pitch_yaw (rotation.hpp:31-48).
inline void rot_rpy(const double v[3], double roll, double pitch, double yaw, double out[3]) {
    const double cr = std::cos(roll), sr = std::sin(roll), cp = std::cos(pitch),
                 sp = std::sin(pitch), cy = std::cos(yaw), sy = std::sin(yaw);
    const double m[3][3] = {
        {cr * cp, cr * sp * sy - sr * cy, cr * sp * cy + sr * sy},
        {sr * cp, sr * sp * sy + cr * cy, sr * sp * cy - cr * sy},
        {-sp, cp * sy, cp * cy},
    };
    for (int r = 0; r < 3; ++r) out[r] = m[r][0] * v[0] + m[r][1] * v[1] + m[r][2] * v[2];
}

// forward_3d.hpp:118-190: pixel (0,0), (1,0), (0,1) and the source, detector roll/pitch/yaw,
// +DOD along z, trajectory rotation, translation into voxel coordinates, COR shift.
inline FwdGeom fwd_angle(const Desc &g, double phi, double theta, double psi) {
    const double dod = g.dsd - g.dso;
    const double sp[3] = {g.sx, g.sy, g.sz}, off[3] = {g.ox, g.oy, g.oz};
    const double n[3] = {(double)g.nx, (double)g.ny, (double)g.nz};
    double pts[4][3] = {
        {-g.du * ((double)g.nu / 2.0) + g.du * 0.5, -g.dv * ((double)g.nv / 2.0) + g.dv * 0.5,
         0.0},
        {0, 0, 0},
        {0, 0, 0},
        {0.0, 0.0, -g.dso},
    };
    for (int d = 0; d < 3; ++d) pts[1][d] = pts[2][d] = pts[0][d];
    pts[1][0] += g.du;
    pts[2][1] += g.dv;
    for (int q = 0; q < 3; ++q) {
        double t[3];
        rot_rpy(pts[q], g.roll, g.pitch, g.yaw, t);
        t[2] += dod;
        for (int d = 0; d < 3; ++d) pts[q][d] = t[d];
    }
    const double corx = g.cor * std::sin(phi) / g.sx, corz = g.cor * std::cos(phi) / g.sz;
    for (int q = 0; q < 4; ++q) {
        double t[3];
        rot_yzy(pts[q], phi, theta, psi, t);
        for (int d = 0; d < 3; ++d) {
            const double tr = n[d] * sp[d] / 2.0 - sp[d] / 2.0;
            pts[q][d] = (t[d] - off[d] + tr) / sp[d];
        }
        pts[q][0] += corx;
        pts[q][2] += corz;
    }
    FwdGeom r{};
    for (int d = 0; d < 3; ++d) {
        r.org[d] = pts[0][d];
        r.du[d] = pts[1][d] - pts[0][d];
        r.dv[d] = pts[2][d] - pts[0][d];
        r.src[d] = pts[3][d];
    }
    return r;
}

// ASTRA-like cone view: v = (source x, y, z, detector centre x, y, z, u-step x, y, z, v-step x,
// y, z) in world coordinates (the volume's centre at the volume offset), the u-step going from
// pixel (iu, iv) to (iu + 1, iv) and the v-step to (iu, iv + 1). No counterpart in the
// reference, whose cone geometry takes Euler angles only and lists projection matrices as
// future work (README.md:58-59). Pixel (0, 0)'s centre is c - (U - 1)/2 u - (V - 1)/2 v, the
// reference's detector origin (forward_3d.hpp:123-125) for a centred panel; the voxel frame
// follows as in fwd_angle (translation by vol_size/2 - spacing/2 - offset, then / spacing).
inline FwdGeom fwd_view(const Desc &g, const double v[12]) {
    const double sp[3] = {g.sx, g.sy, g.sz}, off[3] = {g.ox, g.oy, g.oz};
    const double n[3] = {(double)g.nx, (double)g.ny, (double)g.nz};
    const double hu = ((double)g.nu - 1.0) / 2.0, hv = ((double)g.nv - 1.0) / 2.0;
    FwdGeom r{};
    for (int d = 0; d < 3; ++d) {
        const double tr = n[d] * sp[d] / 2.0 - sp[d] / 2.0;
        const double o = v[3 + d] - hu * v[6 + d] - hv * v[9 + d];
        r.org[d] = (o - off[d] + tr) / sp[d];
        r.du[d] = v[6 + d] / sp[d];
        r.dv[d] = v[9 + d] / sp[d];
        r.src[d] = (v[d] - off[d] + tr) / sp[d];
    }
    return r;
}

// Backprojector frame of the reference (backward_3d.hpp:110-178, 218-227): rotated position of
// voxel (0,0,0), the three voxel steps, and the source.
struct BpGeom {
    double org[3], dx[3], dy[3], dz[3], src[3];
};

inline BpGeom bp_angle(const Desc &g, double phi, double theta, double psi) {
    const double dod = g.dsd - g.dso;
    const double sp[3] = {g.sx, g.sy, g.sz}, off[3] = {g.ox, g.oy, g.oz};
    const double n[3] = {(double)g.nx, (double)g.ny, (double)g.nz};
    double base[3], pts[4][3];
    for (int d = 0; d < 3; ++d) base[d] = -n[d] * sp[d] / 2.0 + sp[d] / 2.0 + off[d];
    for (int q = 0; q < 4; ++q) {
        for (int d = 0; d < 3; ++d) pts[q][d] = base[d];
        if (q > 0) pts[q][q - 1] += sp[q - 1];
        double t[3], r[3];
        rot_yzy(pts[q], phi, theta, psi, t);
        t[0] += g.off_u;
        t[1] += g.off_v;
        t[2] -= dod;
        rot_rpy(t, g.roll, g.pitch, g.yaw, r);
        r[2] += dod;
        for (int d = 0; d < 3; ++d) pts[q][d] = r[d];
    }
    BpGeom b{};
    for (int d = 0; d < 3; ++d) {
        b.org[d] = pts[0][d];
        b.dx[d] = pts[1][d] - pts[0][d];
        b.dy[d] = pts[2][d] - pts[0][d];
        b.dz[d] = pts[3][d] - pts[0][d];
    }
    const double s0[3] = {-g.off_u, -g.off_v, -g.dso - dod};
    double r[3];
    rot_rpy(s0, g.roll, g.pitch, g.yaw, r);
    r[2] += dod;
    for (int d = 0; d < 3; ++d) b.src[d] = r[d];
    return b;
}

// Ray (iu, iv) of one angle (forward_3d.hpp:57-90): direction from the source to the pixel,
// sample budget ceil(|dir|), slab clip against [-1, n+1]^3 (intersection.h:11-28), nst =
// ceil((tmax - tmin) * budget) steps of dt; nst + 1 samples t_i = tmin + i dt. hit = false for
// a ray that misses the clip box.
struct Ray3 {
    double s[3], d[3];
    double tmin, dt;
    int64_t nst;
    bool hit;
};

__host__ __device__ inline Ray3 setup_ray3(const FwdGeom &a, int64_t nx, int64_t ny, int64_t nz,
                                           int64_t iu, int64_t iv) {
    Ray3 r{};
    double nrm2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double q = a.org[d] + (double)iu * a.du[d] + (double)iv * a.dv[d];
        r.s[d] = a.src[d];
        r.d[d] = q - a.src[d];
        nrm2 += r.d[d] * r.d[d];
    }
    const double budget = ceil(sqrt(nrm2));
    const double hi[3] = {(double)nx + 1.0, (double)ny + 1.0, (double)nz + 1.0};
    double tmin = FLT_MIN, tmax = FLT_MAX;
    for (int d = 0; d < 3; ++d) {
        const double t1 = (-1.0 - r.s[d]) / r.d[d];
        const double t2 = (hi[d] - r.s[d]) / r.d[d];
        tmin = fmin(fmax(t1, tmin), fmax(t2, tmin));
        tmax = fmax(fmin(t1, tmax), fmin(t2, tmax));
    }
    r.hit = tmin <= tmax;
    if (!r.hit) return r;
    const double nst = ceil((tmax - tmin) * budget);
    r.tmin = tmin;
    r.dt = nst > 0.0 ? (tmax - tmin) / nst : 0.0;
    r.nst = (int64_t)nst;
    return r;
}

} // namespace cmi::cone3d

#pragma float_control(pop)
