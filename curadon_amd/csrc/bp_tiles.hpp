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
constexpr int kYB = 16;    // y voxels per thread in the 3D backprojector (y-invariant U, W)

// Per-angle tables: the bodies take the table pointer type as a template parameter. The plain
// kernels pass their `const __restrict__` arguments (scalar loads already). The serving kernels
// pass constant-address-space pointers: their tables come from job records in memory, and
// with generic pointers the compiler loads every entry per lane, since the volume stores of
// earlier tiles might alias; tables do not change during a launch.
using ctab = const __attribute__((address_space(4))) float *;
__device__ __forceinline__ ctab const_table(const void *p) { return (ctab)p; }
// Entry i of a table of float4.
__device__ __forceinline__ float4 table4(const float4 *__restrict__ t, int i) { return t[i]; }
__device__ __forceinline__ float4 table4(ctab t, int i) {
    return make_float4(t[4 * i], t[4 * i + 1], t[4 * i + 2], t[4 * i + 3]);
}

// Linear interpolation of a pair at weight f, FP32 arithmetic.
__device__ __forceinline__ float lerp_pair(const Pair &p, float f) {
    return fmaf(f, p.hi - p.lo, p.lo);
}

// Sum over angles for one pixel at world position (x, y) (backward_2d.hpp:58-96):
//   sum_a w_a^WPOW lerp(sino[a], w_a eta_a + D/2),  w_a = DSD / (xi_a + DSO)
// meta[a] = (-sin, cos, DSD, DSO); px, py = x / du, y / du. `R` covers [A][D + 4].
template <class SS, int WPOW, class TAB>
__device__ __forceinline__ float bp2d_pixel(rsrc_t R, TAB meta, int A,
                                           int D, float x, float y, float px, float py,
                                           float half_d) {
    const int Dp = D + 2 * kPadDet;
    const float hi = (float)D;
    const float c0 = half_d - 0.5f; // D/2 and the texel-centre shift of the lerp, exact
    float acc = 0.0f;
    int row = kPadDet;
    // one angle's sample: weight and lerp fraction, with the tap pair loaded
    auto sample = [&](const float4 m, int r, float &wt, float &al, Pair &p) {
        const float depth = fmaf(m.x, x, fmaf(-m.y, y, m.w)); // xi + DSO
        const float eta = fmaf(m.y, px, m.x * py);
        const float w = m.z * __builtin_amdgcn_rcpf(depth);
        const float u = __builtin_amdgcn_fmed3f(fmaf(w, eta, c0), -2.0f, hi);
        al = __builtin_amdgcn_fractf(u);
        wt = WPOW == 2 ? w * w : w;
        p = buf_pair<SS>(R, r + floor_i32(u), 0);
    };
    int a = 0;
    for (; a + 3 < A; a += 4, row += 4 * Dp) { // four loads in flight
        float wt[4], al[4];
        Pair p[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) sample(table4(meta, a + k), row + k * Dp, wt[k], al[k], p[k]);
#pragma unroll
        for (int k = 0; k < 4; ++k) acc = fmaf(fmaf(al[k], p[k].hi - p[k].lo, p[k].lo), wt[k], acc);
    }
    for (; a < A; ++a, row += Dp) {
        float wt, al;
        Pair p;
        sample(table4(meta, a), row, wt, al, p);
        acc = fmaf(fmaf(al, p.hi - p.lo, p.lo), wt, acc);
    }
    return acc;
}

// Sum over angles for kZB consecutive voxels of one (x, y) column. M[a] is a 3x4 matrix over
// the centred voxel index h = (x - cx, y - cy, z - cz, 1) giving (U, V, W); (U/W, V/W) is the
// lower-tap panel coordinate. MATCHED drops voxels behind the source (W <= 0); FDK weights by
// 1 / W^2. (hx, hy, hz) is the centred index of the column's first voxel. `R` covers
// [A][V + 4][U + 4].
template <class SS, bool MATCHED, bool FDK, class TAB>
__device__ __forceinline__ void bp3d_column(rsrc_t R, TAB M, int A, int U,
                                           int V, float hx, float hy, float hz,
                                           float (&acc)[kZB]) {
    const int Up = U + 2 * kPadDet, Vp = V + 2 * kPadDet;
    const int panel = Up * Vp;
    const int sRow = Up * (int)sizeof(typename SS::raw);
    const float ulim = (float)U, vlim = (float)V;
    int base = kPadDet * Up + kPadDet;
    for (int a = 0; a < A; ++a, base += panel) {
        const TAB m = M + 12 * a; // wave-uniform: scalar loads
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

// ------------------------------------------------------------------------------------------
// LDS-staged y runs, for matrices whose U and W rows do not depend on y (M[a][1] == M[a][9] == 0:
// circular orbits about the y axis, panel turned at most about y). Then a run of voxels along y
// shares its perspective divide, u tap, u weight and distance weight, and v advances by a
// constant step.
//
// A workgroup of 256 threads owns a tile of kTileX x kYB x kTileZ voxels (x, y, z); thread t
// takes the run of kYB y voxels at (x0 + t % kTileX, z0 + t / kTileX). Angles go in groups of
// kGroup: every wavefront computes the group's panel footprints from the tile's 8 corners (8
// lanes per angle; U/W and V/W are linear-fractional, so over a box where W > 0 their extremes
// are at corners). Each group goes in chunks of kChunk angles: the workgroup stages the
// chunk's footprints (rows of at most kFootW columns, at most kFootRows rows, from the
// zero-bordered panel) into LDS and the runs sample from there. A chunk with a footprint that
// does not fit (or a corner with W <= 0) is sampled from global memory instead, with the same
// arithmetic.
//
// Per voxel: u = clamp(U/W), v = clamp(fma(y, dv, V0/W)) with dv = M[a][5]/W, lower taps floor(u),
// floor(v) in the staged panel, bilinear weights fract(u), fract(v); MATCHED drops W <= 0, FDK
// weights by 1/W^2.
constexpr int kTileX = 16, kTileZ = 16;
constexpr int kFootW = 64;                        // staged columns per row
constexpr int kFootCols = 68;                     // LDS row pitch: 16-byte rows, banks shifted
constexpr int kFootRows = 60;
constexpr int kFootSlot = kFootCols * kFootRows;  // floats per staged angle (16 KB)
// Angles per chunk. One: the 16 KB of LDS and the body's 80 VGPRs keep 6 workgroups (24
// wavefronts) per CU; four angles per chunk (64 KB) kept two and ran the 384^3 FDK
// backprojection at half the speed.
constexpr int kChunk = 1;
constexpr int kFootFloats = kChunk * kFootSlot;

// Per-angle set-up of a run: u tap column (element offset of the lower tap's column relative to
// `base`), u weight, first v, v step and the sample weight.
struct RunAngle {
    int col;
    float fu, v, dv, w;
};

template <bool MATCHED, bool FDK, class TAB>
__device__ __forceinline__ RunAngle run_angle(TAB m, float hx, float hy, float hz,
                                              float ulim) {
    const float Uh = fmaf(m[0], hx, fmaf(m[2], hz, m[3]));
    const float Wh = fmaf(m[8], hx, fmaf(m[10], hz, m[11]));
    const float Vh = fmaf(m[4], hx, fmaf(m[5], hy, fmaf(m[6], hz, m[7])));
    const float rw = __builtin_amdgcn_rcpf(Wh);
    const float u = __builtin_amdgcn_fmed3f(Uh * rw, -2.0f, ulim);
    RunAngle r;
    r.col = floor_i32(u);
    r.fu = __builtin_amdgcn_fractf(u);
    r.v = Vh * rw;
    r.dv = m[5] * rw;
    r.w = FDK ? rw * rw : 1.0f;
    if (MATCHED) r.w = Wh > 0.0f ? r.w : 0.0f;
    return r;
}

// Staged footprints: one workgroup-shared array (16 KB) for every instantiation a kernel
// uses, so a kernel that dispatches between modes allocates it once.
static __shared__ __align__(16) float foot_lds[kFootFloats];

// Angles whose footprints one pass of a wavefront evaluates: 8 lanes (the tile's corners) each.
constexpr int kGroup = 8;

// The kYB samples of one run at one angle from its staged footprint. `bk`: byte offset in
// foot_lds of the lower tap's column in the slot's row 0; `vbase`: the padded-panel row of that
// slot row minus kPadDet, so the lower tap's slot row is floor(v) - vbase.
template <bool WEIGHTED>
__device__ __forceinline__ void yrun_samples(const RunAngle &r, unsigned bk, int vbase,
                                             float vlim, float (&acc)[kYB]) {
#pragma unroll
    for (int y = 0; y < kYB; ++y) {
        const float v = __builtin_amdgcn_fmed3f(fmaf((float)y, r.dv, r.v), -2.0f, vlim);
        const unsigned addr = __umul24((unsigned)(floor_i32(v) - vbase), 4u * kFootCols) + bk;
        const float *p =
            reinterpret_cast<const float *>(reinterpret_cast<const char *>(foot_lds) + addr);
        const float r0 = fmaf(r.fu, p[1] - p[0], p[0]);
        const float r1 = fmaf(r.fu, p[kFootCols + 1] - p[kFootCols], p[kFootCols]);
        const float s = fmaf(__builtin_amdgcn_fractf(v), r1 - r0, r0);
        acc[y] = WEIGHTED ? fmaf(s, r.w, acc[y]) : acc[y] + s;
    }
}

// (hx, hy, hz): centred index of this thread's run start; (tx, ty, tz): of the tile origin.
// All 256 threads of the workgroup call it together (lanes without voxels too).
template <class SS, bool MATCHED, bool FDK, class TAB>
__device__ __forceinline__ void bp3d_yrun_lds(rsrc_t R, TAB M, int A, int U,
                                             int V, float hx, float hy, float hz, float tx,
                                             float ty, float tz, float (&acc)[kYB]) {
    const int Up = U + 2 * kPadDet, Vp = V + 2 * kPadDet;
    const int panel = Up * Vp;
    const int sRow = Up * (int)sizeof(typename SS::raw);
    const float ulim = (float)U, vlim = (float)V;
    const int tid = threadIdx.x;
    const int lane = tid & 63, k8 = lane >> 3, c8 = lane & 7;
    // the tile's corner this lane evaluates
    const float cx = (c8 & 1) ? tx + (float)(kTileX - 1) : tx;
    const float cy = (c8 & 2) ? ty + (float)(kYB - 1) : ty;
    const float cz = (c8 & 4) ? tz + (float)(kTileZ - 1) : tz;
    // staging: thread tid copies quad tid % 16 of footprint row tid / 16 (and rows + 16 ...)
    const int qrow = tid >> 4, qcol = (tid & 15) << 2;
    for (int g0 = 0; g0 < A; g0 += kGroup) {
        const int ng = min(kGroup, A - g0);
        // footprints: lanes 8k .. 8k+7 take angle g0 + k
        float umin = 0.0f, umax = 0.0f, vmin = 0.0f, vmax = 0.0f;
        int ok = 1;
        if (k8 < ng) {
            const TAB m = M + 12 * (g0 + k8);
            const float Uh = fmaf(m[0], cx, fmaf(m[2], cz, m[3]));
            const float Wh = fmaf(m[8], cx, fmaf(m[10], cz, m[11]));
            const float Vh = fmaf(m[4], cx, fmaf(m[5], cy, fmaf(m[6], cz, m[7])));
            const float rw = __builtin_amdgcn_rcpf(Wh);
            umin = umax = __builtin_amdgcn_fmed3f(Uh * rw, -2.0f, ulim);
            vmin = vmax = __builtin_amdgcn_fmed3f(Vh * rw, -2.0f, vlim);
            ok = Wh > 0.0f ? 1 : 0;
        }
#pragma unroll
        for (int off = 1; off < 8; off <<= 1) {
            umin = fminf(umin, __shfl_xor(umin, off, 64));
            umax = fmaxf(umax, __shfl_xor(umax, off, 64));
            vmin = fminf(vmin, __shfl_xor(vmin, off, 64));
            vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
            ok = min(ok, __shfl_xor(ok, off, 64));
        }
        // every lane now holds its angle's bounds: the slot's padded-panel origin (su0, sv0) and
        // row count, evaluated once for the group's angles. Lower taps lie within
        // [floor(min) - 1, floor(max) + 1] (one-texel margin for the rounding of the samples'
        // u and v against the corners') and the upper tap adds one more row / column.
        const int ulo = floor_i32(umin) - 1 + kPadDet, uhi = floor_i32(umax) + 2 + kPadDet;
        const int vlo = floor_i32(vmin) - 1 + kPadDet, vhi = floor_i32(vmax) + 2 + kPadDet;
        const int su0 = max(ulo, 0), sv0 = max(vlo, 0);
        const int srows = min(vhi, Vp - 1) - sv0 + 1;
        const uint64_t fitm =
            __ballot(ok && uhi - ulo + 1 <= kFootW && vhi - vlo + 1 <= kFootRows);
        // not unrolled: eight copies of the sampling code per mode cost registers and
        // instruction cache (the readlanes take the uniform lane index from an SGPR)
#pragma unroll 1
        for (int c = 0; c < kGroup; c += kChunk) {
            if (c >= ng) break;
            const int na = min(kChunk, ng - c);
            const int a0 = g0 + c;
            int ssu[kChunk], ssv[kChunk], srw[kChunk];
            bool fits = true;
#pragma unroll
            for (int k = 0; k < kChunk; ++k) {
                ssu[k] = __builtin_amdgcn_readlane(su0, 8 * (c + k));
                ssv[k] = __builtin_amdgcn_readlane(sv0, 8 * (c + k));
                srw[k] = __builtin_amdgcn_readlane(srows, 8 * (c + k));
                if (k < na) fits = fits && ((fitm >> (8 * (c + k))) & 1u);
            }
            if (!fits) { // from global memory
                for (int k = 0; k < na; ++k) {
                    const TAB m = M + 12 * (a0 + k); // wave-uniform: scalar loads
                    const RunAngle r = run_angle<MATCHED, FDK>(m, hx, hy, hz, ulim);
                    const int col = (a0 + k) * panel + kPadDet * Up + kPadDet + r.col;
#pragma unroll
                    for (int y = 0; y < kYB; ++y) {
                        const float v =
                            __builtin_amdgcn_fmed3f(fmaf((float)y, r.dv, r.v), -2.0f, vlim);
                        const int o = col + __mul24(floor_i32(v), Up);
                        const float r0 = lerp_pair(buf_pair<SS>(R, o, 0), r.fu);
                        const float r1 = lerp_pair(buf_pair<SS>(R, o, sRow), r.fu);
                        const float s = fmaf(__builtin_amdgcn_fractf(v), r1 - r0, r0);
                        acc[y] = (MATCHED || FDK) ? fmaf(s, r.w, acc[y]) : acc[y] + s;
                    }
                }
                continue;
            }
            __syncthreads(); // the previous chunk's samples are read
            // staging: thread tid copies quad tid % 16 of rows tid / 16 + 16 j of every slot;
            // the first rows of the chunk's slots with their loads in flight together, then the
            // rest (more loads in flight raised the body to 120 VGPRs, four wavefronts a SIMD)
            {
                Quad first[kChunk];
#pragma unroll
                for (int k = 0; k < kChunk; ++k)
                    if (k < na && qrow < srw[k])
                        first[k] = buf_quad<SS>(R, (a0 + k) * panel + (ssv[k] + qrow) * Up +
                                                       ssu[k] + qcol);
#pragma unroll
                for (int k = 0; k < kChunk; ++k)
                    if (k < na && qrow < srw[k])
                        *reinterpret_cast<float4 *>(
                            &foot_lds[k * kFootSlot + qrow * kFootCols + qcol]) =
                            make_float4(first[k].v[0], first[k].v[1], first[k].v[2],
                                        first[k].v[3]);
#pragma unroll
                for (int k = 0; k < kChunk; ++k) {
                    if (k >= na) break;
                    const int g = (a0 + k) * panel + ssv[k] * Up + ssu[k] + qcol;
                    for (int row = qrow + 16; row < srw[k]; row += 16) {
                        const Quad v = buf_quad<SS>(R, g + row * Up);
                        *reinterpret_cast<float4 *>(
                            &foot_lds[k * kFootSlot + row * kFootCols + qcol]) =
                            make_float4(v.v[0], v.v[1], v.v[2], v.v[3]);
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kChunk; ++k) {
                if (k >= na) break;
                const TAB m = M + 12 * (a0 + k); // wave-uniform: scalar loads
                const RunAngle r = run_angle<MATCHED, FDK>(m, hx, hy, hz, ulim);
                const unsigned bk = 4u * (unsigned)(k * kFootSlot + r.col + kPadDet - ssu[k]);
                yrun_samples<MATCHED || FDK>(r, bk, ssv[k] - kPadDet, vlim, acc);
            }
        }
    }
}

} // namespace cmi
