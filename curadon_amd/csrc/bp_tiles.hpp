// Per-pixel / per-column bodies of the voxel-driven backprojectors, shared by the plain
// kernels (fan2d.hip, cone3d.hip) and the persistent multi-job kernel (serve.hip).
//
// Both read sinograms staged with a zero border of kPadDet cells on every detector edge
// (2D: [A][D + 4]; 3D: [A][V + 4][U + 4]); the lower-tap detector coordinate is clamped into
// [-2, D] (resp. [-2, U], [-2, V]), so off-detector samples read zeros with no per-tap test and
// every pair load stays inside the staged array.
#pragma once

#include "cmi_internal.hpp"

namespace cmi {

constexpr int kPadDet = 2; // zero border of staged sinograms and panels, in cells
constexpr int kZB = 8;     // z voxels per thread in the 3D backprojector (general geometry)
constexpr int kYB = 8;     // y voxels per thread in the 3D backprojector (y-invariant U, W)

// Linear interpolation of a pair at weight f, FP32 arithmetic.
__device__ __forceinline__ float lerp_pair(const Pair &p, float f) {
    return fmaf(f, p.hi - p.lo, p.lo);
}

// Sum over angles for one pixel at world position (x, y) (backward_2d.hpp:58-96):
//   sum_a w_a^WPOW lerp(sino[a], w_a eta_a + D/2),  w_a = DSD / (xi_a + DSO)
// meta[a] = (-sin, cos, DSD, DSO); px, py = x / du, y / du. `R` covers [A][D + 4].
template <class SS, int WPOW>
__device__ __forceinline__ float bp2d_pixel(rsrc_t R, const float4 *__restrict__ meta, int A,
                                           int D, float x, float y, float px, float py,
                                           float half_d) {
    const int Dp = D + 2 * kPadDet;
    const float hi = (float)D;
    float acc = 0.0f;
    int row = kPadDet;
#pragma unroll 4
    for (int a = 0; a < A; ++a, row += Dp) {
        const float4 m = meta[a]; // wave-uniform: scalar loads
        const float xi = m.x * x - m.y * y;
        const float eta = m.y * px + m.x * py;
        const float w = m.z * __builtin_amdgcn_rcpf(xi + m.w);
        const float u = __builtin_amdgcn_fmed3f(fmaf(w, eta, half_d) - 0.5f, -2.0f, hi);
        const float al = __builtin_amdgcn_fractf(u);
        const Pair p = buf_pair<SS>(R, row + (int)floorf(u), 0);
        acc = fmaf(fmaf(al, p.hi - p.lo, p.lo), WPOW == 2 ? w * w : w, acc);
    }
    return acc;
}

// Sum over angles for kZB consecutive voxels of one (x, y) column. M[a] is a 3x4 matrix over
// the centred voxel index h = (x - cx, y - cy, z - cz, 1) giving (U, V, W); (U/W, V/W) is the
// lower-tap panel coordinate. MATCHED drops voxels behind the source (W <= 0); FDK weights by
// 1 / W^2. (hx, hy, hz) is the centred index of the column's first voxel. `R` covers
// [A][V + 4][U + 4].
template <class SS, bool MATCHED, bool FDK>
__device__ __forceinline__ void bp3d_column(rsrc_t R, const float *__restrict__ M, int A, int U,
                                           int V, float hx, float hy, float hz,
                                           float (&acc)[kZB]) {
    const int Up = U + 2 * kPadDet, Vp = V + 2 * kPadDet;
    const int panel = Up * Vp;
    const int sRow = Up * (int)sizeof(typename SS::raw);
    const float ulim = (float)U, vlim = (float)V;
    int base = kPadDet * Up + kPadDet;
    for (int a = 0; a < A; ++a, base += panel) {
        const float *m = M + 12 * a; // wave-uniform: scalar loads
        float Uh = fmaf(m[0], hx, fmaf(m[1], hy, fmaf(m[2], hz, m[3])));
        float Vh = fmaf(m[4], hx, fmaf(m[5], hy, fmaf(m[6], hz, m[7])));
        float Wh = fmaf(m[8], hx, fmaf(m[9], hy, fmaf(m[10], hz, m[11])));
#pragma unroll
        for (int z = 0; z < kZB; ++z) {
            const float rw = __builtin_amdgcn_rcpf(Wh);
            const float u = __builtin_amdgcn_fmed3f(Uh * rw, -2.0f, ulim);
            const float v = __builtin_amdgcn_fmed3f(Vh * rw, -2.0f, vlim);
            const float fu = __builtin_amdgcn_fractf(u), fv = __builtin_amdgcn_fractf(v);
            const int o = base + __mul24((int)floorf(v), Up) + (int)floorf(u);
            const Pair p0 = buf_pair<SS>(R, o, 0);
            const Pair p1 = buf_pair<SS>(R, o, sRow);
            const float r0 = fmaf(fu, p0.hi - p0.lo, p0.lo);
            const float r1 = fmaf(fu, p1.hi - p1.lo, p1.lo);
            float s = fmaf(fv, r1 - r0, r0);
            if (FDK) s *= rw * rw;
            if (MATCHED) s = Wh > 0.0f ? s : 0.0f;
            acc[z] += s;
            Uh += m[2];
            Vh += m[6];
            Wh += m[10];
        }
    }
}

// Sum over angles for kYB consecutive voxels of one (x, z) row along y, for matrices whose U and
// W rows do not depend on y (M[a][1] == M[a][9] == 0: circular source orbits about the y axis
// with an untilted panel). The perspective divide, the u tap and weight and the distance
// weight are then shared by the whole run; only V advances along y. (hx, hy, hz) is the centred
// index of the run's first voxel. Semantics as bp3d_column.
template <class SS, bool MATCHED, bool FDK>
__device__ __forceinline__ void bp3d_yrun(rsrc_t R, const float *__restrict__ M, int A, int U,
                                         int V, float hx, float hy, float hz,
                                         float (&acc)[kYB]) {
    const int Up = U + 2 * kPadDet, Vp = V + 2 * kPadDet;
    const int panel = Up * Vp;
    const int sRow = Up * (int)sizeof(typename SS::raw);
    const float ulim = (float)U, vlim = (float)V;
    int base = kPadDet * Up + kPadDet;
    for (int a = 0; a < A; ++a, base += panel) {
        const float *m = M + 12 * a; // wave-uniform: scalar loads
        const float Uh = fmaf(m[0], hx, fmaf(m[2], hz, m[3]));
        const float Wh = fmaf(m[8], hx, fmaf(m[10], hz, m[11]));
        float Vh = fmaf(m[4], hx, fmaf(m[5], hy, fmaf(m[6], hz, m[7])));
        const float rw = __builtin_amdgcn_rcpf(Wh);
        const float u = __builtin_amdgcn_fmed3f(Uh * rw, -2.0f, ulim);
        const float fu = __builtin_amdgcn_fractf(u);
        const int col = base + (int)floorf(u);
        float w = FDK ? rw * rw : 1.0f;
        if (MATCHED) w = Wh > 0.0f ? w : 0.0f;
#pragma unroll
        for (int y = 0; y < kYB; ++y) {
            const float v = __builtin_amdgcn_fmed3f(Vh * rw, -2.0f, vlim);
            const float fv = __builtin_amdgcn_fractf(v);
            const int o = col + __mul24((int)floorf(v), Up);
            const float r0 = lerp_pair(buf_pair<SS>(R, o, 0), fu);
            const float r1 = lerp_pair(buf_pair<SS>(R, o, sRow), fu);
            const float s = fmaf(fv, r1 - r0, r0);
            acc[y] = (MATCHED || FDK) ? fmaf(s, w, acc[y]) : acc[y] + s;
            Vh += m[5];
        }
    }
}

} // namespace cmi
