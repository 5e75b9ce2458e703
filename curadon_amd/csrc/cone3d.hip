// 3D cone-beam projectors for gfx950: plan, forward ray march, voxel-driven backprojection
// (reference and geometry-matched projection) and the transposed ray march.
//
// Data layout in HBM
//   plan tables  fgeo [A] FwdGeom (FP64)     detector pixel (0,0), u/v steps, source (voxel frame)
//                mref [A][12] float          reference-mode projection matrix (see bp_matrices)
//                mmat [A][12] float          geometry-matched projection matrix
//                mfdk [A][12] float          matched matrix scaled so that W = (depth / DSO)
//                fast [A] int                axis (0 x, 1 y, 2 z) the detector's v step follows most
//                                            closely: the forward reads the copy in which it is fastest
//   per call     forward:  [B][copies] volume with a 2-voxel zero border, one copy per fast axis
//                          in use: x fastest [z][y][x], y fastest [z][x][y], z fastest [y][x][z]
//                backward: [B][A][V+4][U+4]      panels with a 2-pixel zero border
//                adjoint:  per-ray march table [2][A][V][U] float4 (kernel k_ray_table3d)
//
// Forward rays take their discrete quantities (hit, step count) from an FP64 set-up in the
// kernel, identical to the host oracle's, then march in FP32 from the ray's midpoint in
// volume-centred padded coordinates, summing in blocks of 32 samples.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <type_traits>
#include <vector>

#include "bp_tiles.hpp"
#include "cmi_internal.hpp"
#include "geometry3d.hpp"
#include "stage_index.hpp"

using namespace cmi;
using cmi::cone3d::FwdGeom;

struct cmi_plan3d {
    int device = 0;
    int vol_prec = CMI_F32, det_prec = CMI_F32;
    cone3d::Desc g{};
    int64_t A = 0;
    uint64_t total_samples = 0;
    uint64_t table_bytes = 0;
    FwdGeom *d_fgeo = nullptr;
    int *d_fast = nullptr;      // [A] fast axis of each angle's volume copy
    int copy_slot[3] = {-1, -1, -1}; // per fast axis: its copy's index, -1 if no angle uses it
    int ncopies = 0;
    float *d_mref = nullptr;
    float *d_mmat = nullptr;
    float *d_mfdk = nullptr;
    bool ycols[3] = {false, false, false}; // per cmi_bp3d_mode: U and W rows free of y
    bool views = false; // built from ASTRA-like views: reference mode = matched, no FDK mode
    // per cmi_bp3d_mode: smallest W of the mode's matrices over the volume's corner voxels and
    // all angles (W is affine in the voxel, so this bounds it over the volume); the transpose
    // of the backprojection needs it positive
    double wmin[3] = {0.0, 0.0, 0.0};
};

namespace {

constexpr int kPad = 2;        // zero border of staged volumes and panels
constexpr int kMaxBlock = 256; // threads per workgroup every kernel here is compiled for
constexpr int kSumBlock = 32;  // samples per partial sum in the forward march
static_assert(kPad == kPadDet, "panels are staged with the backprojector's border");

// ---------------------------------------------------------------- ray parameters ----------

// FP32 march parameters of one ray in centred padded coordinates (voxel coordinate + kPad
// minus the padded volume's integer centre c0): midpoint m, step st, first index j0 = -nst/2,
// sample count n = nst + 1. Shared by the forward march and its transpose so both see the same
// FP32 positions.
struct March3 {
    float m[3], st[3];
    float j0;
    int n;
};

__device__ __forceinline__ March3 march_params(const FwdGeom &g, int nx, int ny, int nz, int iu,
                                               int iv, int Wp, int Hp, int Dp) {
    March3 p{};
    const cone3d::Ray3 r = cone3d::setup_ray3(g, nx, ny, nz, iu, iv);
    if (!r.hit) return p; // n = 0
    const int64_t im = r.nst / 2;
    const double tm = r.tmin + (double)im * r.dt;
    const double c0[3] = {(double)(Wp / 2), (double)(Hp / 2), (double)(Dp / 2)};
    for (int d = 0; d < 3; ++d) {
        p.m[d] = (float)(r.s[d] + tm * r.d[d] + (double)kPad - c0[d]);
        p.st[d] = (float)(r.dt * r.d[d]);
    }
    p.j0 = (float)(-im);
    p.n = (int)(r.nst + 1);
    return p;
}

__global__ __launch_bounds__(kMaxBlock) void k_count3d(const FwdGeom *__restrict__ geo, int U,
                                                       int V, int nx, int ny, int nz,
                                                       unsigned long long *__restrict__ total) {
    const int iu = blockIdx.x * blockDim.x + threadIdx.x;
    const int iv = blockIdx.y;
    const int a = blockIdx.z;
    unsigned long long n = 0;
    if (iu < U) {
        const cone3d::Ray3 r = cone3d::setup_ray3(geo[a], nx, ny, nz, iu, iv);
        n = r.hit ? (unsigned long long)(r.nst + 1) : 0ull;
    }
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(total, n);
}

// ---------------------------------------------------------------- staging -----------------

// pad[b][zp][yp][xp] = vol[b][zp-2][yp-2][xp-2], 0 in the border.
template <class S>
__global__ __launch_bounds__(kMaxBlock) void k_stage3d(const typename S::raw *__restrict__ vol,
                                                       typename S::raw *__restrict__ pad, int nx,
                                                       int ny, int nz, int Wp, int Hp, int Dp) {
    using raw = typename S::raw;
    const int xp = blockIdx.x * blockDim.x + threadIdx.x;
    if (xp >= Wp) return;
    const int yp = blockIdx.y;
    const int zp = blockIdx.z % Dp;
    const size_t b = blockIdx.z / Dp;
    const int x = xp - kPad, y = yp - kPad, z = zp - kPad;
    raw v = (raw)0.0f;
    if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz)
        v = vol[((b * nz + z) * ny + y) * (size_t)nx + x];
    pad[((b * Dp + zp) * Hp + yp) * (size_t)Wp + xp] = v;
}

// Copy with fast axis F = 1 (y) or 2 (z) and x as the middle axis: for each slow coordinate
// (z for F = 1, y for F = 2), dst[slow][xp][fp] = pad(xp, yp, zp), transposed through LDS in
// 32 x 32 tiles (reads along x, writes along the fast axis).
template <class S, int F>
__global__ __launch_bounds__(kMaxBlock) void k_stage3d_t(const typename S::raw *__restrict__ vol,
                                                         typename S::raw *__restrict__ dst,
                                                         int nx, int ny, int nz, int Wp, int Hp,
                                                         int Dp) {
    static_assert(F == 1 || F == 2, "transposed copies have y or z fastest");
    using raw = typename S::raw;
    __shared__ raw tile[32][33];
    const int Ef = F == 1 ? Hp : Dp, Es = F == 1 ? Dp : Hp;
    const int sp = blockIdx.z % Es;
    const size_t b = blockIdx.z / Es;
    const int xp0 = blockIdx.x * 32, fp0 = blockIdx.y * 32;
    const raw *src = vol + b * (size_t)nx * ny * nz;
    raw *out = dst + (size_t)sp * Wp * Ef;
    for (int r = threadIdx.y; r < 32; r += 8) { // r: fast-axis offset, threadIdx.x: x offset
        const int xp = xp0 + threadIdx.x, fp = fp0 + r;
        const int yp = F == 1 ? fp : sp, zp = F == 1 ? sp : fp;
        const int x = xp - kPad, y = yp - kPad, z = zp - kPad;
        raw v = (raw)0.0f;
        if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz)
            v = src[((size_t)z * ny + y) * nx + x];
        tile[r][threadIdx.x] = v;
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) { // r: x offset, threadIdx.x: fast-axis offset
        const int xp = xp0 + r, fp = fp0 + threadIdx.x;
        if (xp < Wp && fp < Ef) out[(size_t)xp * Ef + fp] = tile[threadIdx.x][r];
    }
}

// ppad[b][a][vp][up] = sino[b][a][vp-2][up-2], 0 in the border. Flat over the n destination
// elements, four consecutive ones per thread with one vector store (its offset is a multiple
// of four elements; the scratch stack is allocated with at least that alignment); the source
// row is followed element by element.
template <class S>
__global__ __launch_bounds__(kMaxBlock) void k_stage_panel(const typename S::raw *__restrict__ sino,
                                                           typename S::raw *__restrict__ ppad,
                                                           int U, int V, uint64_t n) {
    using raw = typename S::raw;
    struct alignas(4 * sizeof(raw)) Four {
        raw v[4];
    };
    const uint64_t i0 = ((uint64_t)blockIdx.x * kMaxBlock + threadIdx.x) * 4;
    if (i0 >= n) return;
    const int Up = U + 2 * kPad, Vp = V + 2 * kPad;
    StageIndex at = stage_index(i0, Up, Vp, n <= 0xffffffffull); // at.a = b * A + a
    Four f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int u = at.up - kPad, v = at.vp - kPad;
        raw val = (raw)0.0f;
        if (i0 + k < n && u >= 0 && u < U && v >= 0 && v < V)
            val = sino[(at.a * V + v) * (size_t)U + u];
        f.v[k] = val;
        stage_next(at, Up, Vp);
    }
    if (i0 + 4 <= n) {
        *reinterpret_cast<Four *>(ppad + i0) = f;
    } else {
        for (uint64_t k = 0; i0 + k < n; ++k) ppad[i0 + k] = f.v[k];
    }
}

// ---------------------------------------------------------------- forward -----------------

// Trilinear sample at centred padded coordinates (f, m, l) along the copy's fast, middle and
// slow axes. `cofs` is the element offset of the copy's centre; sM, sL, sML are the byte strides
// of one step along the middle axis, the slow axis, and both.
template <class SI>
__device__ __forceinline__ float sample3d(rsrc_t R, float f, float m, float l, int Ef, int Em,
                                         int cofs, int sM, int sL, int sML) {
    const int i_f = floor_i32(f), i_m = floor_i32(m), i_l = floor_i32(l);
    const float ff = __builtin_amdgcn_fractf(f), fm = __builtin_amdgcn_fractf(m),
                fl = __builtin_amdgcn_fractf(l);
    const int o = __mul24(__mul24(i_l, Em) + i_m, Ef) + i_f + cofs;
    const Pair p00 = buf_pair<SI>(R, o, 0);   // (m,     l)
    const Pair p10 = buf_pair<SI>(R, o, sM);  // (m + 1, l)
    const Pair p01 = buf_pair<SI>(R, o, sL);  // (m,     l + 1)
    const Pair p11 = buf_pair<SI>(R, o, sML); // (m + 1, l + 1)
    const float a00 = fmaf(ff, p00.hi - p00.lo, p00.lo);
    const float a10 = fmaf(ff, p10.hi - p10.lo, p10.lo);
    const float a01 = fmaf(ff, p01.hi - p01.lo, p01.lo);
    const float a11 = fmaf(ff, p11.hi - p11.lo, p11.lo);
    const float b0 = fmaf(fm, a10 - a00, a00);
    const float b1 = fmaf(fm, a11 - a01, a01);
    return fmaf(fl, b1 - b0, b0);
}

// LDS-staged march. The rays of a workgroup (16 x 16 detector pixels of one projection) lie
// inside the pyramid spanned by its four corner rays. The workgroup marches along the axis K
// that its rays follow most closely, in slabs of kSlab slices of lower taps: per slab it stages
// into LDS the box of voxels that the pyramid's samples there can tap (the corner rays'
// crossings of the slab's two bounding planes, one voxel of margin for FP32 positions, and the
// upper taps), read in 16-byte row pieces from the volume copy whose fastest axis P is not K,
// and every ray takes its samples whose lower tap lies in the slab from there, in its own order
// j = 0 .. n - 1, slab after slab: positions and trilinear weights are those of a plain per-ray
// march, and the ray's sum adds one partial sum per slab (at most about 2 kSlab samples each).
// A sample outside the box (not expected; a safety net), and workgroups whose rays do not all
// advance the same way along K or whose K is the copy's fast axis, read global memory instead,
// with the same arithmetic (the latter in blocks of kSumBlock samples).
// A wavefront covers 4 x 16 detector pixels, the 16 along v, and the copy read is the one
// whose fastest axis v follows (fast[a]), so that a wavefront's rows of a staged box are runs
// of consecutive voxels.
// sino[b][a][iv][iu] = sum_{i=0..nst} trilinear(vol, s + (tmin + i dt) d)  (forward_3d.hpp:87-93)
constexpr int kSlab = 8;          // slices of lower taps per slab
// floats of one staged box (16 KiB). A smaller box keeps more workgroups per CU and sends more
// slabs to the global path: at 256^3 the staged-box capacity in floats gave 2560: 18.3 ms,
// 3072: 16.7, 3584: 16.3, 4096: 16.3, 4608: 16.5, 5120: 17.5, 5632: 17.6, 6144: 19.1
// (bit-identical: both paths add the same values in the same order).
constexpr int kBoxCells = 4096;
constexpr int kBoxQuads = 8;      // at most 8 x 4 cells per staged row
constexpr unsigned kFwdGroup = 32; // projections per group in the workgroup order

template <class SI, class SS>
__global__ __launch_bounds__(kMaxBlock) void k_fwd3d(const typename SI::raw *__restrict__ pad,
                                                          const FwdGeom *__restrict__ geo,
                                                          const int *__restrict__ fast, int slot0,
                                                          int slot1, int slot2, int ncopies,
                                                          int nx, int ny, int nz, int U, int V,
                                                          int A,
                                                          typename SS::raw *__restrict__ sino) {
    __shared__ float4 box4[kBoxCells / 4];
    float *box = reinterpret_cast<float *>(box4);
    __shared__ float cl[4][6]; // corner rays in padded index coordinates: point, direction
    __shared__ int red[2][kMaxBlock / 64];
    const int tid = threadIdx.x;
    // XCD-aware order. Workgroups are dealt to the 8 XCDs in turn, so dispatch slot hw runs
    // on XCD hw % 8; logical index li gives each XCD a contiguous share of the order below.
    // In that order kFwdGroup consecutive projections of one tile come first, then the tiles
    // (u fastest), then the next group of projections: the 256 or so workgroups resident on an
    // XCD cover about 8 tiles over 32 neighbouring projections, whose pyramids overlap, so
    // their staged boxes share L2 (21.2 -> 18.8 ms at 256^3 against tiles first; projections
    // first across all 360 was 19.8 ms).
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned nwg = gx * gy * gridDim.z;
    const unsigned hw = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned n8 = nwg & ~7u;
    const unsigned li = hw < n8 ? (hw & 7u) * (n8 >> 3) + (hw >> 3) : hw;
    unsigned tx, ty, tz;
    {
        const unsigned T = gx * gy, AT = (unsigned)A * T, G = kFwdGroup;
        const unsigned bb = li / AT, l = li % AT;
        const unsigned full = ((unsigned)A / G) * G * T; // workgroups in whole groups
        unsigned t, ai;
        if (l < full) {
            const unsigned r = l % (G * T);
            ai = l / (G * T) * G + r % G;
            t = r / G;
        } else { // the last, partial group
            const unsigned r = l - full, Gl = (unsigned)A - (unsigned)A / G * G;
            ai = (unsigned)A / G * G + r % Gl;
            t = r / Gl;
        }
        tx = t % gx;
        ty = t / gx;
        tz = ai + (unsigned)A * bb;
    }
    const int iu0 = (int)tx * 16, iv0 = (int)ty * 16;
    // a wavefront covers 4 x 16 pixels, consecutive lanes along v
    const int iu = iu0 + (tid >> 4), iv = iv0 + (tid & 15);
    const bool live = iu < U && iv < V;
    const int a = (int)tz % A;
    const size_t b = tz / A;
    const int Wp = nx + 2 * kPad, Hp = ny + 2 * kPad, Dp = nz + 2 * kPad;
    if (tid < 4) {
        const FwdGeom &g = geo[a];
        const int cu = (tid & 1) ? min(iu0 + 15, U - 1) : iu0;
        const int cv = (tid & 2) ? min(iv0 + 15, V - 1) : iv0;
        for (int d = 0; d < 3; ++d) {
            const double q = g.org[d] + (double)cu * g.du[d] + (double)cv * g.dv[d];
            cl[tid][d] = (float)(g.src[d] + (double)kPad);
            cl[tid][3 + d] = (float)(q - g.src[d]);
        }
    }
    March3 p{};
    if (live) p = march_params(geo[a], nx, ny, nz, iu, iv, Wp, Hp, Dp);
    // copy axes (fast, middle, slow): x fastest (x, y, z), y fastest (y, x, z), z fastest (z, x, y)
    // wave-uniform (so that the buffer resource below is a scalar one)
    const int F = __builtin_amdgcn_readfirstlane(fast[a]);
    const int M = F == 0 ? 1 : 0, L = F == 2 ? 1 : 2;
    const int slot = F == 0 ? slot0 : F == 1 ? slot1 : slot2;
    // component d of a three-vector without indexing a local array (which would live in
    // scratch memory)
    auto pick = [](auto v0, auto v1, auto v2, int d) { return d == 0 ? v0 : d == 1 ? v1 : v2; };
    const int Ef = pick(Wp, Hp, Dp, F), Em = pick(Wp, Hp, Dp, M), El = pick(Wp, Hp, Dp, L);
    const uint32_t one = (uint32_t)Wp * Hp * Dp;
    const rsrc_t R = make_rsrc(pad + (b * ncopies + slot) * one, bytes_of<SI>(one));
    constexpr int es = (int)sizeof(typename SI::raw);
    const int sM = Ef * es, sL = Ef * Em * es, sML = sL + sM;
    const int cofs = ((El / 2) * Em + Em / 2) * Ef + Ef / 2;
    const float mf = pick(p.m[0], p.m[1], p.m[2], F), mm = pick(p.m[0], p.m[1], p.m[2], M),
                ml = pick(p.m[0], p.m[1], p.m[2], L);
    const float tf = pick(p.st[0], p.st[1], p.st[2], F), tm = pick(p.st[0], p.st[1], p.st[2], M),
                tl = pick(p.st[0], p.st[1], p.st[2], L);
    __syncthreads(); // corner rays
    // march axis K: the largest component of the corner rays' summed direction
    float ds[3];
    for (int d = 0; d < 3; ++d) ds[d] = cl[0][3 + d] + cl[1][3 + d] + cl[2][3 + d] + cl[3][3 + d];
    const int K = fabsf(ds[0]) >= fabsf(ds[1]) ? (fabsf(ds[0]) >= fabsf(ds[2]) ? 0 : 2)
                                               : (fabsf(ds[1]) >= fabsf(ds[2]) ? 1 : 2);
    const float sgn = ds[K] > 0.0f ? 1.0f : -1.0f;
    bool slab_ok = K != F;
    for (int c = 0; c < 4; ++c) slab_ok = slab_ok && cl[c][3 + K] * sgn > 0.0f;
    const int Q = 3 - K - F; // the third axis (when K != F)
    // padded index = centred coordinate + E / 2 on each axis
    const int cK = pick(Wp, Hp, Dp, K) / 2, cP = Ef / 2, cQ = pick(Wp, Hp, Dp, Q) / 2;
    const float mK = pick(p.m[0], p.m[1], p.m[2], K), tK = pick(p.st[0], p.st[1], p.st[2], K);
    const float mP = mf, tP = tf;
    const float mQ = pick(p.m[0], p.m[1], p.m[2], Q), tQ = pick(p.st[0], p.st[1], p.st[2], Q);
    // element strides of K and Q in the copy; and whether K is the copy's middle axis (then a
    // step in K is the trilinear m step, else the l step)
    const bool kmid = K == M;
    const int sK = kmid ? Ef : Ef * Em, sQ = kmid ? Ef * Em : Ef;

    float acc = 0.0f;
    int j = 0;
    const int n = p.n;
    float jf = p.j0;
    // sample j at the forward's position, from global memory
    auto global_sample = [&]() {
        return sample3d<SI>(R, fmaf(jf, tf, mf), fmaf(jf, tm, mm), fmaf(jf, tl, ml), Ef, Em,
                            cofs, sM, sL, sML);
    };
    if (!__builtin_amdgcn_readfirstlane((int)slab_ok)) {
        while (j < n) {
            const int jend = min(j + kSumBlock, n);
            float s0 = 0.0f;
            for (; j < jend; ++j, jf += 1.0f) s0 += global_sample();
            acc += s0;
        }
    } else {
        // lower-tap K range of the workgroup's samples
        int klo = 0x7fffffff, khi = -0x7fffffff - 1;
        if (n > 0) {
            const int k0 = floor_i32(fmaf(p.j0, tK, mK)) + cK;
            const int k1 = floor_i32(fmaf(p.j0 + (float)(n - 1), tK, mK)) + cK;
            klo = min(k0, k1);
            khi = max(k0, k1);
        }
        for (int off = 32; off > 0; off >>= 1) {
            klo = min(klo, __shfl_xor(klo, off, 64));
            khi = max(khi, __shfl_xor(khi, off, 64));
        }
        if ((tid & 63) == 0) {
            red[0][tid >> 6] = klo;
            red[1][tid >> 6] = khi;
        }
        __syncthreads();
        klo = min(min(red[0][0], red[0][1]), min(red[0][2], red[0][3]));
        khi = max(max(red[1][0], red[1][1]), max(red[1][2], red[1][3]));
        const int nslab = khi >= klo ? (khi - klo) / kSlab + 1 : 0;
        const float rtK = n > 0 ? 1.0f / tK : 0.0f;
        // lanes 8 g + c + 4 e of every wavefront project corner ray c onto plane e of a slab
        const int cc = tid & 3, ce = (tid >> 2) & 1;
        const float cKo = cl[cc][K], cKr = __builtin_amdgcn_rcpf(cl[cc][3 + K]);
        const float cPo = cl[cc][F], cPd = cl[cc][3 + F], cQo = cl[cc][Q], cQd = cl[cc][3 + Q];
        for (int si = 0; si < nslab; ++si) {
            // lower taps K in [D0, D0 + kSlab - 1]; slabs follow the rays' direction along K
            const int D0 = sgn > 0.0f ? klo + si * kSlab : khi - kSlab + 1 - si * kSlab;
            // the box: the corner rays where they cross K = D0 and K = D0 + kSlab
            const float t = ((float)(D0 + ce * kSlab) - cKo) * cKr;
            float pmin = fmaf(t, cPd, cPo), qmin = fmaf(t, cQd, cQo);
            float pmax = pmin, qmax = qmin;
#pragma unroll
            for (int off = 1; off < 8; off <<= 1) {
                pmin = fminf(pmin, __shfl_xor(pmin, off, 64));
                pmax = fmaxf(pmax, __shfl_xor(pmax, off, 64));
                qmin = fminf(qmin, __shfl_xor(qmin, off, 64));
                qmax = fmaxf(qmax, __shfl_xor(qmax, off, 64));
            }
            const int P0 = __builtin_amdgcn_readfirstlane((int)floorf(pmin) - 1);
            const int Q0 = __builtin_amdgcn_readfirstlane((int)floorf(qmin) - 1);
            const int WP = __builtin_amdgcn_readfirstlane((int)floorf(pmax) + 2 - P0 + 1);
            const int WQ = __builtin_amdgcn_readfirstlane((int)floorf(qmax) + 2 - Q0 + 1);
            const int WPp = (WP + 3) & ~3;
            const int S = WQ * WPp; // floats per staged slice
            const bool staged = WPp <= 4 * kBoxQuads && WQ <= kBoxCells / ((kSlab + 1) * WPp);
            if (staged) {
                // rows (slice kk, row qq) of WP cells, 16 bytes per thread: 32 rows a pass
                const int rows = (kSlab + 1) * WQ, quad = tid & (kBoxQuads - 1);
                int r = tid >> 3, kk = r / WQ, qq = r - kk * WQ;
                for (; r < rows; r += kMaxBlock / kBoxQuads) {
                    if (4 * quad < WPp) {
                        const int e = (D0 + kk) * sK + (Q0 + qq) * sQ + P0 + 4 * quad;
                        const Quad v = buf_quad_at<SI>(R, e * es);
                        *reinterpret_cast<float4 *>(&box[(kk * WQ + qq) * WPp + 4 * quad]) =
                            make_float4(v.v[0], v.v[1], v.v[2], v.v[3]);
                    }
                    qq += kMaxBlock / kBoxQuads;
                    while (qq >= WQ) {
                        qq -= WQ;
                        ++kk;
                    }
                }
            }
            __syncthreads(); // the box is staged
            // this ray's samples in the slab: j .. je - 1, where its position along K crosses
            // the slab's far plane. A sample that the rounding of this estimate puts in the
            // wrong slab lies outside the box and takes the global path.
            const float bK = (float)((sgn > 0.0f ? D0 + kSlab : D0) - cK);
            const int je = j < n ? min(max((int)ceilf(fmaf(bK - mK, rtK, -p.j0)), j), n) : j;
            const int kofs = cK - D0, pofs = cP - P0, qofs = cQ - Q0;
            // KM: K is the copy's middle axis (a step in K is the trilinear m step)
            auto sample_slab = [&](auto km) {
                constexpr bool KM = decltype(km)::value;
                const int mstep = KM ? S : WPp, lstep = KM ? WPp : S;
                auto in_box = [&](int kk, int ip, int iq) {
                    return staged && (unsigned)kk < (unsigned)kSlab &&
                           (unsigned)ip <= (unsigned)(WP - 2) && (unsigned)iq <= (unsigned)(WQ - 2);
                };
                // strides in bytes, so that a sample's address is two 24-bit products and one
                // shifted add
                const int S4 = 4 * S, W4 = 4 * WPp, m4 = 4 * mstep, l4 = 4 * lstep;
                // the sample at byte offset `ofs` of the box (its lower tap pair)
                auto box_at = [&](float pK, float pf, float pq, int ofs) {
                    const char *cb = reinterpret_cast<const char *>(box) + ofs;
                    const float *c = reinterpret_cast<const float *>(cb);
                    const float *cm = reinterpret_cast<const float *>(cb + m4);
                    const float *cl = reinterpret_cast<const float *>(cb + l4);
                    const float *cml = reinterpret_cast<const float *>(cb + m4 + l4);
                    const float ff = __builtin_amdgcn_fractf(pf);
                    const float fm = __builtin_amdgcn_fractf(KM ? pK : pq);
                    const float fl = __builtin_amdgcn_fractf(KM ? pq : pK);
                    const float a00 = fmaf(ff, c[1] - c[0], c[0]);
                    const float a10 = fmaf(ff, cm[1] - cm[0], cm[0]);
                    const float a01 = fmaf(ff, cl[1] - cl[0], cl[0]);
                    const float a11 = fmaf(ff, cml[1] - cml[0], cml[0]);
                    const float b0 = fmaf(fm, a10 - a00, a00);
                    const float b1 = fmaf(fm, a11 - a01, a01);
                    return fmaf(fl, b1 - b0, b0);
                };
                auto box_sample = [&](float pK, float pf, float pq, int kk, int ip, int iq) {
                    return box_at(pK, pf, pq, __mul24(kk, S4) + __mul24(iq, W4) + 4 * ip);
                };
                // Along a ray every box coordinate is monotonic in jf (jf takes exact integer
                // values), so a lane whose first and last samples of the slab lie in the box has
                // all of them there. When every lane of the wavefront is such a lane, it runs the
                // loop without the per-sample test: the same samples from the same cells, added
                // in the same order.
                bool inside = true;
                if (j < je) {
                    const float jl = jf + (float)(je - 1 - j);
                    inside = in_box(floor_i32(fmaf(jf, tK, mK)) + kofs,
                                    floor_i32(fmaf(jf, tP, mP)) + pofs,
                                    floor_i32(fmaf(jf, tQ, mQ)) + qofs) &&
                             in_box(floor_i32(fmaf(jl, tK, mK)) + kofs,
                                    floor_i32(fmaf(jl, tP, mP)) + pofs,
                                    floor_i32(fmaf(jl, tQ, mQ)) + qofs);
                }
                float part = 0.0f;
                if (__builtin_amdgcn_ballot_w64(!inside) == 0) {
                    // counted by jf alone (exact integers); the box offsets of the copy's
                    // coordinates folded into one wave-uniform byte offset
                    const float jfe = jf + (float)max(je - j, 0);
                    const int ofs0 = kofs * S4 + qofs * W4 + 4 * pofs;
                    for (; jf < jfe; jf += 1.0f) {
                        const float pK = fmaf(jf, tK, mK);
                        const float pf = fmaf(jf, tP, mP), pq = fmaf(jf, tQ, mQ);
                        part += box_at(pK, pf, pq,
                                       __mul24(floor_i32(pK), S4) + __mul24(floor_i32(pq), W4) +
                                           4 * floor_i32(pf) + ofs0);
                    }
                    j = max(j, je);
                } else {
                    for (; j < je; ++j, jf += 1.0f) {
                        const float pK = fmaf(jf, tK, mK);
                        const float pf = fmaf(jf, tP, mP), pq = fmaf(jf, tQ, mQ);
                        const int kk = floor_i32(pK) + kofs;
                        const int ip = floor_i32(pf) + pofs, iq = floor_i32(pq) + qofs;
                        if (in_box(kk, ip, iq)) part += box_sample(pK, pf, pq, kk, ip, iq);
                        else part += global_sample();
                    }
                }
                acc += part;
            };
            if (kmid) sample_slab(std::true_type{});
            else sample_slab(std::false_type{});
            __syncthreads(); // the box is read before the next slab is staged
        }
    }
    if (live) sino[((b * A + a) * V + iv) * (size_t)U + iu] = from_f32<SS>(acc);
}

// ---------------------------------------------------------------- adjoint -----------------

// Per-ray march table for the adjoint: rm = (m.xyz, j0), rs = (st.xyz, n as int bits), the same
// FP32 parameters k_fwd3d derives with march_params; n = 0 for rays that miss the volume.
__global__ __launch_bounds__(kMaxBlock) void k_ray_table3d(const FwdGeom *__restrict__ geo, int nx,
                                                          int ny, int nz, int U, int V,
                                                          float4 *__restrict__ rm,
                                                          float4 *__restrict__ rs) {
    const int iu = blockIdx.x * blockDim.x + threadIdx.x;
    const int iv = blockIdx.y;
    const int a = blockIdx.z;
    if (iu >= U) return;
    const int Wp = nx + 2 * kPad, Hp = ny + 2 * kPad, Dp = nz + 2 * kPad;
    const March3 p = march_params(geo[a], nx, ny, nz, iu, iv, Wp, Hp, Dp);
    const size_t r = ((size_t)a * V + iv) * U + iu;
    rm[r] = make_float4(p.m[0], p.m[1], p.m[2], p.j0);
    rs[r] = make_float4(p.st[0], p.st[1], p.st[2], __int_as_float(p.n));
}

// Largest |y| of each projection stack of a batch, in two steps: per stack b, kMaxParts
// workgroups each write the maximum of their grid-stride share to parts[b][*]; the adjoint
// reduces the kMaxParts values. Each volume of a batch gets its own fixed-point unit, so that
// its result does not depend on the other volumes.
constexpr int kMaxParts = 256;

template <class S>
__global__ __launch_bounds__(kMaxBlock) void k_absmax_parts(const typename S::raw *__restrict__ y,
                                                            size_t n, float *__restrict__ parts) {
    __shared__ float red[kMaxBlock / 64];
    y += (size_t)blockIdx.y * n;
    parts += (size_t)blockIdx.y * kMaxParts;
    float m = 0.0f;
    for (size_t i = (size_t)blockIdx.x * kMaxBlock + threadIdx.x; i < n;
         i += (size_t)kMaxParts * kMaxBlock)
        m = fmaxf(m, fabsf(to_f32<S>(y[i])));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kMaxBlock / 64; ++w) m = fmaxf(m, red[w]);
        parts[blockIdx.x] = m;
    }
}

// Matched adjoint, tile-driven: a workgroup owns a kAT^3 voxel tile with accumulators in LDS.
// For each angle it projects the tile's box, grown by one voxel on each side (the samples
// whose trilinear footprint reaches the tile), through the geometry-matched matrices onto the
// panel; the rays through the covered pixels (plus one pixel of margin) are the candidates. Each
// candidate walks only its samples inside the grown box, at the forward's FP32 positions, and
// adds y * w to the tile voxels its footprint covers. The accumulators are fixed point in units
// of max|y| / 2^19: a contribution y * w is rounded to that grid (|term| <= 2^19) and added to a
// 32-bit LDS cell with an integer LDS atomic, which gfx950 executes about 45 times faster than an
// FP32 LDS atomic add (and 32-bit cells halve the LDS atomic traffic of 64-bit ones). Integer
// sums are exact, so the result does not depend on the order of the additions; the rounding of
// each term is at most max|y| 2^-20.
// Overflow: a ray adds at most two terms to a voxel (its samples are one apart along the
// dominant axis, so at most two have their lower corner within one voxel of it), and only rays
// through the projection of the voxel's 2 x 2 x 2 neighbourhood reach it. Per angle that
// projection is at most Du x Dv pixels, with Du <= 2 sum_i |du/dx_i| bounded over the grown box
// by (sum |m0..2| + max|u| sum |m8..10|) / min w from its corners (and Dv alike), so one angle
// adds at most 2 (Du + 2)(Dv + 2) terms to a voxel (2 per candidate where that is smaller).
// The workgroup keeps a running bound of these and, before it could pass 2^31 / 2^19, moves
// the 32-bit sums of its 16^3 voxels into 64-bit sums and clears the tile. The 64-bit sums live
// in a global array of which each workgroup owns one tile's worth (written by the first drain,
// read and written only by the thread that owns the voxel), not in registers: sixteen 64-bit
// registers per thread held the kernel at four wavefronts per SIMD, and without them it keeps
// six (48.0 -> 43.3 ms at 256^3, bit-identical).
// The tile is added to the volume once, in its storage type.
constexpr int kAT = 16;
// candidate rays of one wavefront are kPh pixels apart in u and v (at 256^3, six wavefronts per
// SIMD, tiles in XCD order: kPh 1: 47.1 ms, 2: 38.2; tiles in launch order: 2: 38.7, 3: 42.5,
// 4: 43.7, 5: 45.5, 6: 46.1)
constexpr int kPh = 2;
// LDS tile: padded by one voxel before and after on each axis, so that every tap of a sample
// whose lower corner lies in [-1, kAT) on each axis lands in the array (taps outside the tile
// are added to the border and dropped): the eight taps are then immediate offsets from one LDS
// address. Odd row and slice pitches (in words) spread voxels that differ only in y or z over
// different banks, as the lanes of one atomic instruction hit voxels at various depths.
constexpr int kATP = kAT + 2, kATRow = kATP + 1, kATSlice = kATP * kATRow + 1;
constexpr int kATCells = kATP * kATSlice;
constexpr int kATOwn = kAT * kAT * kAT / kMaxBlock; // tile voxels per thread
static_assert(kAT * kAT * kAT % kMaxBlock == 0, "whole voxels per thread");
constexpr int kAdj3Bits = 19;                             // a term's bound: 2^kAdj3Bits units
constexpr int kAdj3Budget = (1 << (31 - kAdj3Bits)) - 16; // terms per voxel between drains

template <class SV, class SS>
__global__ __launch_bounds__(kMaxBlock) void k_adj3d(typename SV::raw *__restrict__ vol,
                                                     const typename SS::raw *__restrict__ sino,
                                                     const float4 *__restrict__ rm,
                                                     const float4 *__restrict__ rs,
                                                     const float *__restrict__ M,
                                                     const float *__restrict__ parts, int nx,
                                                     int ny, int nz, int U, int V, int A,
                                                     long long *__restrict__ totals) {
    __shared__ unsigned int tile[kATCells];
    const int tid = threadIdx.x;
    // this workgroup's 64-bit voxel sums so far, in global memory it alone uses; none before
    // the first drain
    long long *own = totals + ((size_t)blockIdx.x +
                               gridDim.x * ((size_t)blockIdx.y + gridDim.y * (size_t)blockIdx.z)) *
                                  (kAT * kAT * kAT);
    bool first = true;
    auto cell_of = [](int j) {
        return (j / (kAT * kAT) + 1) * kATSlice + ((j / kAT) % kAT + 1) * kATRow + j % kAT + 1;
    };
    // the tile's voxel sums -> the workgroup's 64-bit sums, then the tile (border included)
    // cleared
    auto drain = [&]() {
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < kATOwn; ++k) {
            const int i = tid + k * kMaxBlock;
            const long long v = (int)tile[cell_of(i)];
            own[i] = first ? v : own[i] + v;
        }
        first = false;
        __syncthreads();
        for (int i = tid; i < kATCells; i += kMaxBlock) tile[i] = 0u;
        __syncthreads();
    };
    int bound = 0; // terms a voxel may have received since the last drain
    const int ntz = (nz + kAT - 1) / kAT;
    // XCD-aware order: workgroups are dealt to the 8 XCDs in turn, and logical index li gives
    // each XCD a contiguous run of tiles (x fastest), whose candidate rays share L2
    int x0, y0, z0;
    size_t b;
    {
        const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
        const unsigned nwg = gx * gy * gz;
        const unsigned hw = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
        const unsigned n8 = nwg & ~7u;
        const unsigned li = hw < n8 ? (hw & 7u) * (n8 >> 3) + (hw >> 3) : hw;
        const unsigned per = gx * gy * (unsigned)ntz;
        b = li / per;
        const unsigned l = li % per;
        x0 = (int)(l % gx) * kAT;
        y0 = (int)((l / gx) % gy) * kAT;
        z0 = (int)(l / (gx * gy)) * kAT;
    }
    // fixed-point scale: max|y| of this volume's projections maps to 2^30
    float ymax = 0.0f;
    for (int i = tid & 63; i < kMaxParts; i += 64) ymax = fmaxf(ymax, parts[b * kMaxParts + i]);
    for (int off = 32; off > 0; off >>= 1) ymax = fmaxf(ymax, __shfl_xor(ymax, off, 64));
    const float yscale = ymax > 0.0f ? (float)(1 << kAdj3Bits) / ymax : 0.0f;
    for (int i = tid; i < kATCells; i += kMaxBlock) tile[i] = 0u;
    const int Wp = nx + 2 * kPad, Hp = ny + 2 * kPad, Dp = nz + 2 * kPad;
    // centred padded coordinate p -> voxel index floor(p) + c
    const int cx = Wp / 2 - kPad, cy = Hp / 2 - kPad, cz = Dp / 2 - kPad;
    // centred padded coordinate p -> padded tile index floor(p) + o
    const int ox = cx - x0 + 1, oy = cy - y0 + 1, oz = cz - z0 + 1;
    // samples reaching the tile: floor(p) + c in [t0 - 1, t0 + kAT - 1] on each axis
    const float lox = (float)(x0 - 1 - cx), loy = (float)(y0 - 1 - cy), loz = (float)(z0 - 1 - cz);
    const float hix = lox + (float)(kAT + 1), hiy = loy + (float)(kAT + 1),
                hiz = loz + (float)(kAT + 1);
    // the grown box's corner this lane projects (centred voxel index of the matrices)
    const int c8 = tid & 7;
    const float qx = (float)((c8 & 1) ? x0 + kAT : x0 - 1) - (float)((nx - 1) / 2.0);
    const float qy = (float)((c8 & 2) ? y0 + kAT : y0 - 1) - (float)((ny - 1) / 2.0);
    const float qz = (float)((c8 & 4) ? z0 + kAT : z0 - 1) - (float)((nz - 1) / 2.0);
    const typename SS::raw *sb = sino + b * (size_t)A * V * U;
    __syncthreads();
    for (int a = 0; a < A; ++a) {
        const float *m = M + 12 * a; // wave-uniform: scalar loads
        const float Uh = fmaf(m[0], qx, fmaf(m[1], qy, fmaf(m[2], qz, m[3])));
        const float Vh = fmaf(m[4], qx, fmaf(m[5], qy, fmaf(m[6], qz, m[7])));
        const float Wh = fmaf(m[8], qx, fmaf(m[9], qy, fmaf(m[10], qz, m[11])));
        const float rw = __builtin_amdgcn_rcpf(Wh);
        float umin = Uh * rw, umax = umin, vmin = Vh * rw, vmax = vmin;
        float uabs = fabsf(umin), vabs = fabsf(vmin), wmin = Wh;
        int ok = Wh > 0.0f ? 1 : 0;
#pragma unroll
        for (int off = 1; off < 8; off <<= 1) {
            umin = fminf(umin, __shfl_xor(umin, off, 64));
            umax = fmaxf(umax, __shfl_xor(umax, off, 64));
            vmin = fminf(vmin, __shfl_xor(vmin, off, 64));
            vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
            ok = min(ok, __shfl_xor(ok, off, 64));
            uabs = fmaxf(uabs, __shfl_xor(uabs, off, 64));
            vabs = fmaxf(vabs, __shfl_xor(vabs, off, 64));
            wmin = fminf(wmin, __shfl_xor(wmin, off, 64));
        }
        // every lane group holds the same values; take lane 0's
        int iu0 = 0, iu1 = U - 1, iv0 = 0, iv1 = V - 1;
        // this angle's terms per voxel, at most (no bound where a corner is not in front of
        // the source); 2^30 stands for unbounded
        int per_angle = 1 << 30;
        if (__builtin_amdgcn_readfirstlane(ok)) {
            const float ua = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(uabs)));
            const float va = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(vabs)));
            const float wm = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(wmin)));
            const float su = fabsf(m[0]) + fabsf(m[1]) + fabsf(m[2]);
            const float sv = fabsf(m[4]) + fabsf(m[5]) + fabsf(m[6]);
            const float sw = fabsf(m[8]) + fabsf(m[9]) + fabsf(m[10]);
            // 1 % margin for the approximate reciprocals; the +2 covers the pixel count of an
            // interval (floor + 1) and rays that FP32 positions put one pixel further
            const float du = 2.02f * (su + ua * sw) / wm, dv = 2.02f * (sv + va * sw) / wm;
            if (du < 1024.0f && dv < 1024.0f)
                per_angle = 2 * ((int)du + 2) * ((int)dv + 2);
            const float u0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(umin)));
            const float u1 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(umax)));
            const float v0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(vmin)));
            const float v1 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(vmax)));
            if (u1 < -2.0f || u0 > (float)(U + 1) || v1 < -2.0f || v0 > (float)(V + 1)) continue;
            iu0 = max(0, (int)ceilf(u0) - 1);
            iu1 = min(U - 1, (int)floorf(u1) + 1);
            iv0 = max(0, (int)ceilf(v0) - 1);
            iv1 = min(V - 1, (int)floorf(v1) + 1);
        }
        const int niu = iu1 - iu0 + 1, niv = iv1 - iv0 + 1;
        if (niu <= 0 || niv <= 0) continue;
        // Candidates in phase order: phase (pu, pv) in [0, kPh)^2 holds the rays
        // (pu + kPh qu, pv + kPh qv), so the consecutive candidates of a wavefront are kPh
        // pixels apart in u and v and their samples' taps rarely meet in one atomic instruction
        // (LDS serialises lanes that add to the same address). Padding the phases to equal
        // size leaves holes past the rectangle's edge, which are skipped.
        const int nqu = (niu + kPh - 1) / kPh, nqv = (niv + kPh - 1) / kPh;
        const int nq = nqu * nqv, nc = kPh * kPh * nq;
        // passes of kMaxBlock candidates, the same for the whole workgroup; a pass adds at most
        // two terms per candidate to a voxel, and the angle at most per_angle in all
        int allowance = per_angle;
        for (int cb = 0; cb < nc; cb += kMaxBlock) {
            const int inc = min(2 * min(kMaxBlock, nc - cb), allowance);
            allowance -= inc;
            if (bound + inc > kAdj3Budget) {
                drain();
                bound = 0;
            }
            bound += inc;
            const int c = cb + tid;
            if (c >= nc) continue;
            const int ph = c / nq, rq = c - ph * nq;
            const int qv = rq / nqu, qu = rq - qv * nqu;
            const int lu = (ph % kPh) + kPh * qu, lv = (ph / kPh) + kPh * qv;
            if (lu >= niu || lv >= niv) continue;
            const int iu = iu0 + lu, iv = iv0 + lv;
            const size_t r = ((size_t)a * V + iv) * U + iu;
            const float4 s4 = rs[r];
            const int n = __float_as_int(s4.w);
            if (n <= 0) continue;
            const float y = to_f32<SS>(sb[r]) * yscale; // |y| <= 2^30
            if (y == 0.0f) continue;
            const float4 m4 = rm[r];
            // jf range with p in [lo, hi) on every axis, one sample of margin each side
            float jlo = m4.w, jhi = m4.w + (float)(n - 1);
            const float st[3] = {s4.x, s4.y, s4.z}, mm[3] = {m4.x, m4.y, m4.z};
            const float lo[3] = {lox, loy, loz}, hi[3] = {hix, hiy, hiz};
            bool empty = false;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (st[d] != 0.0f) {
                    const float r0 = __builtin_amdgcn_rcpf(st[d]);
                    const float t1 = (lo[d] - mm[d]) * r0, t2 = (hi[d] - mm[d]) * r0;
                    jlo = fmaxf(jlo, fminf(t1, t2) - 1.0f);
                    jhi = fminf(jhi, fmaxf(t1, t2) + 1.0f);
                } else if (mm[d] < lo[d] || mm[d] >= hi[d]) {
                    empty = true;
                }
            }
            if (empty) continue;
            const float jend = fminf(floorf(jhi), m4.w + (float)(n - 1));
            for (float jf = fmaxf(ceilf(jlo), m4.w); jf <= jend; jf += 1.0f) {
                // the forward's own positions (k_fwd3d)
                const float px = fmaf(jf, s4.x, m4.x), py = fmaf(jf, s4.y, m4.y),
                            pz = fmaf(jf, s4.z, m4.z);
                // padded tile indices of the lower corner, kept if in [0, kAT] on each axis
                const int lx = floor_i32(px) + ox, ly = floor_i32(py) + oy,
                          lz = floor_i32(pz) + oz;
                if ((unsigned)lx > (unsigned)kAT || (unsigned)ly > (unsigned)kAT ||
                    (unsigned)lz > (unsigned)kAT)
                    continue;
                const float fx = __builtin_amdgcn_fractf(px), fy = __builtin_amdgcn_fractf(py),
                            fz = __builtin_amdgcn_fractf(pz);
                // the eight trilinear weights times y, as differences of products
                const float x1 = y * fx, x0w = y - x1;
                const float y01 = x0w * fy, y00 = x0w - y01, y11 = x1 * fy, y10 = x1 - y11;
                const float w[8] = {y00 - y00 * fz, y10 - y10 * fz, y01 - y01 * fz,
                                    y11 - y11 * fz, y00 * fz, y10 * fz, y01 * fz, y11 * fz};
                unsigned int *t = &tile[__mul24(lz, kATSlice) + __mul24(ly, kATRow) + lx];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    atomicAdd(t + (k >> 2) * kATSlice + ((k >> 1) & 1) * kATRow + (k & 1),
                              (unsigned int)round_i32(w[k]));
            }
        }
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < kATOwn; ++k) {
        const int i = tid + k * kMaxBlock;
        const long long total = (first ? 0ll : own[i]) + (int)tile[cell_of(i)];
        const int lx = i % kAT, ly = (i / kAT) % kAT, lz = i / (kAT * kAT);
        const int X = x0 + lx, Y = y0 + ly, Z = z0 + lz;
        if (X < nx && Y < ny && Z < nz) {
            const size_t o = ((b * nz + Z) * ny + Y) * (size_t)nx + X;
            const float add = (float)((double)total * ((double)ymax * 0x1p-19));
            vol[o] = from_f32<SV>(to_f32<SV>(vol[o]) + add);
        }
    }
}

// ---------------------------------------------------------------- backprojection ----------

// One thread per voxel column (ix, iy) and kZB consecutive z; a wavefront covers a 16 x 4 x 8
// block (bp3d_column). vol += sum over angles.
template <class SV, class SS, bool MATCHED, bool FDK>
__global__ __launch_bounds__(kMaxBlock) void k_bp3d(typename SV::raw *__restrict__ vol,
                                                    const typename SS::raw *__restrict__ ppad,
                                                    const float *__restrict__ M, int nx, int ny,
                                                    int nz, int U, int V, int A, float cx,
                                                    float cy, float cz) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    const int nzb = (nz + kZB - 1) / kZB;
    const int iz0 = (blockIdx.z % nzb) * kZB;
    const size_t b = blockIdx.z / nzb;
    if (ix >= nx || iy >= ny) return;
    const uint32_t stack = (uint32_t)(U + 2 * kPad) * (V + 2 * kPad) * A;
    const rsrc_t R = make_rsrc(ppad + b * (size_t)stack, bytes_of<SS>(stack));
    float acc[kZB];
#pragma unroll
    for (int z = 0; z < kZB; ++z) acc[z] = 0.0f;
    bp3d_column<SS, MATCHED, FDK>(R, M, A, U, V, (float)ix - cx, (float)iy - cy,
                                  (float)iz0 - cz, acc);
#pragma unroll
    for (int z = 0; z < kZB; ++z) {
        const int iz = iz0 + z;
        if (iz < nz) {
            const size_t i = ((b * nz + iz) * ny + iy) * (size_t)nx + ix;
            vol[i] = from_f32<SV>(to_f32<SV>(vol[i]) + acc[z]);
        }
    }
}

// For matrices whose U and W rows do not depend on y: a workgroup of 256 threads owns a tile of
// kTileX x kYB x kTileZ voxels, each thread a run of kYB y voxels, and the angles' panel
// footprints are staged in LDS (bp3d_yrun_lds). vol += sum over angles.
template <class SV, class SS, bool MATCHED, bool FDK>
__global__ __launch_bounds__(kMaxBlock) void k_bp3d_y(typename SV::raw *__restrict__ vol,
                                                      const typename SS::raw *__restrict__ ppad,
                                                      const float *__restrict__ M, int nx, int ny,
                                                      int nz, int U, int V, int A, float cx,
                                                      float cy, float cz) {
    const int nzt = (nz + kTileZ - 1) / kTileZ;
    // XCD-aware order: workgroups are dealt to the 8 XCDs in turn, and logical index li gives
    // each XCD a contiguous run of tiles (x fastest), whose panel footprints share L2
    // (2.51 -> 2.37 ms at 256^3)
    int ix0, iy0, iz0;
    size_t b;
    {
        const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
        const unsigned nwg = gx * gy * gz;
        const unsigned hw = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
        const unsigned n8 = nwg & ~7u;
        const unsigned li = hw < n8 ? (hw & 7u) * (n8 >> 3) + (hw >> 3) : hw;
        const unsigned per = gx * gy * (unsigned)nzt;
        b = li / per;
        const unsigned l = li % per;
        unsigned tx, ty, tz;
        tx = l % gx;
        ty = (l / gx) % gy;
        tz = l / (gx * gy);
        ix0 = (int)tx * kTileX;
        iy0 = (int)ty * kYB;
        iz0 = (int)tz * kTileZ;
    }
    const int ix = ix0 + (int)(threadIdx.x % kTileX), iz = iz0 + (int)(threadIdx.x / kTileX);
    const uint32_t stack = (uint32_t)(U + 2 * kPad) * (V + 2 * kPad) * A;
    const rsrc_t R = make_rsrc(ppad + b * (size_t)stack, bytes_of<SS>(stack));
    float acc[kYB];
#pragma unroll
    for (int y = 0; y < kYB; ++y) acc[y] = 0.0f;
    bp3d_yrun_lds<SS, MATCHED, FDK>(R, M, A, U, V, (float)ix - cx, (float)iy0 - cy,
                                    (float)iz - cz, (float)ix0 - cx, (float)iy0 - cy,
                                    (float)iz0 - cz, acc);
    if (ix >= nx || iz >= nz) return;
#pragma unroll
    for (int y = 0; y < kYB; ++y) {
        const int iy = iy0 + y;
        if (iy < ny) {
            const size_t i = ((b * nz + iz) * ny + iy) * (size_t)nx + ix;
            vol[i] = from_f32<SV>(to_f32<SV>(vol[i]) + acc[y]);
        }
    }
}

// ------------------------------------------------ transpose of the backprojection -------

// Bᵀ: the transpose of the voxel-driven backprojection (k_bp3d, k_bp3d_y), the gradient of
// TorchBackprojection. The backprojection adds to voxel h, per projection a, the bilinear
// interpolation of panel a at (u, v) = (U/W, V/W) from the mode's matrix rows (cells floor(u)
// + {0, 1} and floor(v) + {0, 1}, weights 1 - fu, fu and 1 - fv, fv; cells outside the panel
// read zero), times 1/W^2 for FDK. So sino[a][j][i] = sum_h (those weights) vol[h].
// A workgroup owns a 16 x 16 pixel tile of one projection and gathers it. The voxels that reach
// the tile are those with (u, v) in [iu0 - 1, iu0 + 16) x [iv0 - 1, iv0 + 16): a pyramid from
// the source, bounded by the four lines through the source and that rectangle's corners. The
// workgroup marches through the volume along the axis K that the tile's central ray follows
// most closely, in chunks of kBTSlab slices. In each slice the pyramid's section is the
// quadrilateral where the corner lines cross it; its bounding box (one voxel of margin on each
// side, clipped to the volume) is walked by the workgroup's threads, x fastest (y when K is x,
// from a y-fastest copy of the volume), so that a wavefront reads runs of consecutive voxels.
// Every voxel computes its own (u, v) and adds its taps inside the tile to the tile's cells in
// LDS. The corner lines need the rays of the pyramid to cross the slices (each corner line's
// direction has a K component of a quarter of its length, all of one sign); otherwise the
// whole slice is walked.
// Fixed point: terms are rounded to units of max|vol| wbound / 2^kBTBits, where wbound >= the
// weight (1, or 1 / min W^2 for FDK, W > 0 over the volume), so |term| <= 2^kBTBits and a
// 32-bit LDS cell holds the sum of up to 2^(31 - kBTBits) - 1 terms. A voxel adds at most one
// term to a cell, so the workgroup moves the cells into 64-bit registers (one cell per thread)
// and clears them before its voxel visits since the last such drain could pass that count.
// Integer sums are exact: the result does not depend on the order of the additions. The tile is
// written once, in the sinogram's storage type.
constexpr int kBT = 16;                                        // panel tile edge, pixels
constexpr int kBTP = kBT + 2;                                  // LDS cells per row and rows
constexpr int kBTSlab = 32;                                    // slices per chunk
constexpr int kBTBits = 18;                                    // a term's bound: 2^kBTBits
constexpr int kBTBudget = (1 << (31 - kBTBits)) - 1;           // visits between drains
// voxels per thread at a time (at 256^3: 1: 12.7 ms, 2: 11.4, 4: 11.7, 8: 13.8)
constexpr int kBTBatch = 2;
constexpr int kBTPass = kBTBudget / (kBTBatch * kMaxBlock) * (kBTBatch * kMaxBlock); // per pass
static_assert(kBT * kBT == kMaxBlock, "one tile cell per thread");
static_assert(kBTSlab <= 64, "one lane per slice of a chunk");

template <class SV, class SS, bool FDK>
__global__ __launch_bounds__(kMaxBlock) void k_bpT3d(typename SS::raw *__restrict__ sino,
                                                     const typename SV::raw *__restrict__ vol,
                                                     const typename SV::raw *__restrict__ ycopy,
                                                     const float *__restrict__ M,
                                                     const float *__restrict__ parts,
                                                     float wbound, int nx, int ny, int nz, int U,
                                                     int V, int A) {
    __shared__ unsigned int cells[kBTP * kBTP]; // pixel iu0 - 1 + i, iv0 - 1 + j at j kBTP + i
    __shared__ int4 slab[2][kBTSlab];           // per slice: F range, G range (inclusive)
    __shared__ int dims[2][2];                  // per chunk: box extents along F and G
    __shared__ float coef[4][4];                // per corner line: F = a + b hK, G = c + d hK
    __shared__ int whole;                       // 1: walk whole slices
    const int tid = threadIdx.x, lane = tid & 63;
    const int a = (int)(blockIdx.x % (unsigned)A), ti = (int)(blockIdx.x / (unsigned)A);
    const int tv = (V + kBT - 1) / kBT;
    const int tj = (int)(blockIdx.y % (unsigned)tv);
    const size_t b = blockIdx.y / (unsigned)tv;
    const int iu0 = ti * kBT, iv0 = tj * kBT;
    const float *m = M + 12 * a;
    const float cx = (float)((nx - 1) / 2.0), cy = (float)((ny - 1) / 2.0),
                cz = (float)((nz - 1) / 2.0);
    // fixed-point scale (FP64: max|vol| wbound may lie far from 1)
    float vmax = 0.0f;
    for (int i = lane; i < kMaxParts; i += 64) vmax = fmaxf(vmax, parts[b * kMaxParts + i]);
    for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
    // a voxel value maps to vol ldexp(smant, sexp) units, the FP64 scale 2^18 / (max|vol| wbound)
    // split into an FP32 mantissa and an exponent (the scale may lie outside the FP32 range)
    static_assert(kBTBits == 18, "the scale uses 2^18");
    const double unit_inv = vmax > 0.0f ? 0x1p18 / ((double)vmax * (double)wbound) : 0.0;
    const int sexp = unit_inv > 0.0 ? ilogb(unit_inv) : 0;
    const float smant = (float)ldexp(unit_inv, -sexp);
    // march axis: the largest component of the central ray's direction, the cross product of
    // the (U - uc W) and (V - vc W) rows' voxel parts
    int K;
    {
        const float uc = (float)iu0 + 0.5f * (kBT - 1), vc = (float)iv0 + 0.5f * (kBT - 1);
        float e1[3], e2[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = fmaf(-uc, m[8 + k], m[k]);
            e2[k] = fmaf(-vc, m[8 + k], m[4 + k]);
        }
        const float d0 = fabsf(e1[1] * e2[2] - e1[2] * e2[1]);
        const float d1 = fabsf(e1[2] * e2[0] - e1[0] * e2[2]);
        const float d2 = fabsf(e1[0] * e2[1] - e1[1] * e2[0]);
        K = d0 >= d1 && d0 >= d2 ? 0 : (d1 >= d2 ? 1 : 2);
        K = __builtin_amdgcn_readfirstlane(K);
    }
    const int F = K == 0 ? 1 : 0, G = K == 2 ? 1 : 2;
    // corner lines (FP64), lanes 0..3 of the first wavefront
    if (tid < 64) {
        const int c = lane & 3;
        const double uc = (double)(iu0 - 1 + ((c & 1) ? kBT + 1 : 0));
        const double vc = (double)(iv0 - 1 + ((c & 2) ? kBT + 1 : 0));
        const double a0 = (double)m[0] - uc * (double)m[8], a1 = (double)m[1] - uc * (double)m[9],
                     a2 = (double)m[2] - uc * (double)m[10], a3 = (double)m[3] - uc * (double)m[11];
        const double q0 = (double)m[4] - vc * (double)m[8], q1 = (double)m[5] - vc * (double)m[9],
                     q2 = (double)m[6] - vc * (double)m[10], q3 = (double)m[7] - vc * (double)m[11];
        const double cr0 = a1 * q2 - a2 * q1, cr1 = a2 * q0 - a0 * q2, cr2 = a0 * q1 - a1 * q0;
        // components along K, F and G
        auto pick = [](double x, double y, double z, int i) {
            return i == 0 ? x : (i == 1 ? y : z);
        };
        const double pF = pick(a0, a1, a2, F), pG = pick(a0, a1, a2, G), pK = pick(a0, a1, a2, K);
        const double qF = pick(q0, q1, q2, F), qG = pick(q0, q1, q2, G), qK = pick(q0, q1, q2, K);
        const double det = pF * qG - pG * qF; // +-cr[K]
        const double len = sqrt(cr0 * cr0 + cr1 * cr1 + cr2 * cr2);
        const bool ok = fabs(det) >= 0.25 * len && len > 0.0;
        const unsigned long long okm = __ballot(ok), posm = __ballot(det > 0.0);
        if (lane < 4) {
            // (a . (h, 1) = 0, q . (h, 1) = 0) solved for (hF, hG) given hK
            const double id = ok ? 1.0 / det : 0.0;
            coef[c][0] = (float)((pG * q3 - a3 * qG) * id);
            coef[c][1] = (float)((pG * qK - pK * qG) * id);
            coef[c][2] = (float)((a3 * qF - pF * q3) * id);
            coef[c][3] = (float)((pK * qF - pF * qK) * id);
        }
        if (lane == 0)
            whole = !(okm == ~0ull && (posm == 0ull || posm == ~0ull));
    }
    for (int i = tid; i < kBTP * kBTP; i += kMaxBlock) cells[i] = 0u;
    __syncthreads();
    const int nK = K == 0 ? nx : (K == 1 ? ny : nz);
    const int nF = F == 0 ? nx : ny, nG = G == 1 ? ny : nz;
    const float cK = K == 0 ? cx : (K == 1 ? cy : cz);
    const float cF = F == 0 ? cx : cy, cG = G == 1 ? cy : cz;
    // voxel (f, g, k) along (F, G, K) lies at src[f + g strG + k strK]: the volume, or for K = x
    // the y-fastest copy [z + 2][x + 2][y + 2]
    const int Wp = nx + 2 * kPad, Hp = ny + 2 * kPad;
    const typename SV::raw *src =
        K == 0 ? ycopy + b * (size_t)Wp * Hp * (nz + 2 * kPad) + (kPad * Wp + kPad) * Hp + kPad
               : vol + b * (size_t)nx * ny * nz;
    const int strG = K == 0 ? Wp * Hp : (K == 1 ? nx * ny : nx);
    const int strK = K == 0 ? Hp : (K == 1 ? nx : nx * ny);
    // matrix rows over (hF, hG, hK, 1)
    auto row = [&](int r, float &rF, float &rG, float &rK, float &r3) {
        const float x = m[4 * r], y = m[4 * r + 1], z = m[4 * r + 2];
        rF = F == 0 ? x : y;
        rG = G == 1 ? y : z;
        rK = K == 0 ? x : (K == 1 ? y : z);
        r3 = m[4 * r + 3];
    };
    float uF, uG, uK, u3, vF, vG, vK, v3, wF, wG, wK, w3;
    row(0, uF, uG, uK, u3);
    row(1, vF, vG, vK, v3);
    row(2, wF, wG, wK, w3);
    unsigned int *mine = &cells[(tid / kBT + 1) * kBTP + tid % kBT + 1];
    long long own = 0;
    // cells -> registers (border cells, which belong to other tiles, are never read)
    auto drain = [&]() {
        __syncthreads();
        own += (long long)(int)*mine;
        *mine = 0u;
        __syncthreads();
    };
    int run = 0; // voxel visits since the last drain
    int buf = 0;
    for (int k0 = 0; k0 < nK; k0 += kBTSlab, buf ^= 1) {
        const int S = min(kBTSlab, nK - k0);
        if (tid < 64) {
            int flo = 0, fhi = -1, glo = 0, ghi = -1;
            if (lane < S) {
                if (whole) {
                    fhi = nF - 1;
                    ghi = nG - 1;
                } else {
                    const float hK = (float)(k0 + lane) - cK;
                    float f0 = INFINITY, f1 = -INFINITY, g0 = INFINITY, g1 = -INFINITY;
                    for (int c = 0; c < 4; ++c) {
                        const float f = fmaf(coef[c][1], hK, coef[c][0]);
                        const float g = fmaf(coef[c][3], hK, coef[c][2]);
                        f0 = fminf(f0, f);
                        f1 = fmaxf(f1, f);
                        g0 = fminf(g0, g);
                        g1 = fmaxf(g1, g);
                    }
                    constexpr float big = 0x1p24f;
                    flo = max(0, floor_i32(__builtin_amdgcn_fmed3f(f0 + cF, -big, big)) - 1);
                    fhi = min(nF - 1, floor_i32(__builtin_amdgcn_fmed3f(f1 + cF, -big, big)) + 1);
                    glo = max(0, floor_i32(__builtin_amdgcn_fmed3f(g0 + cG, -big, big)) - 1);
                    ghi = min(nG - 1, floor_i32(__builtin_amdgcn_fmed3f(g1 + cG, -big, big)) + 1);
                }
                slab[buf][lane] = make_int4(flo, fhi, glo, ghi);
            }
            int ef = max(fhi - flo + 1, 0), eg = max(ghi - glo + 1, 0);
            for (int off = 32; off > 0; off >>= 1) {
                ef = max(ef, __shfl_xor(ef, off, 64));
                eg = max(eg, __shfl_xor(eg, off, 64));
            }
            if (lane == 0) {
                dims[buf][0] = ef;
                dims[buf][1] = eg;
            }
        }
        __syncthreads();
        const int NF = dims[buf][0], NG = dims[buf][1];
        const int total = NF * NG * S;
        if (total == 0) continue;
        // this thread's element e = tid + n kMaxBlock of the chunk's box, as (dF, dG, dk) with
        // dF fastest; a step of kMaxBlock elements is (sF, sG, sK)
        int dF = tid % NF, dG = (tid / NF) % NG, dk = tid / (NF * NG);
        const int sF = kMaxBlock % NF, sG = (kMaxBlock / NF) % NG, sK = kMaxBlock / (NF * NG);
        int e = tid;
        auto advance = [&]() {
            dF += sF;
            if (dF >= NF) {
                dF -= NF;
                ++dG;
            }
            dG += sG;
            if (dG >= NG) {
                dG -= NG;
                ++dk;
            }
            dk += sK;
        };
        for (int base = 0; base < total; base += kBTPass) {
            const int n = min(kBTPass, total - base);
            if (run + n > kBTBudget) {
                drain();
                run = 0;
            }
            run += n;
            // kBTBatch elements per thread at a time: their loads are issued together. Passes
            // start at multiples of kBTBatch kMaxBlock elements, so a batch never straddles two.
            for (; e < base + n; e += kBTBatch * kMaxBlock) {
                int fq[kBTBatch], gq[kBTBatch], kq[kBTBatch], addr[kBTBatch];
                bool live[kBTBatch];
#pragma unroll
                for (int q = 0; q < kBTBatch; ++q) {
                    const int4 sl = slab[buf][min(dk, S - 1)];
                    fq[q] = sl.x + dF;
                    gq[q] = sl.z + dG;
                    kq[q] = k0 + dk;
                    live[q] = e + q * kMaxBlock < base + n && fq[q] <= sl.y && gq[q] <= sl.w;
                    addr[q] = live[q] ? fq[q] + gq[q] * strG + kq[q] * strK : 0;
                    advance();
                }
                float val[kBTBatch];
#pragma unroll
                for (int q = 0; q < kBTBatch; ++q) val[q] = to_f32<SV>(src[addr[q]]);
#pragma unroll
                for (int q = 0; q < kBTBatch; ++q) {
                    const float vs = live[q] ? ldexpf(val[q] * smant, sexp) : 0.0f;
                    if (vs == 0.0f) continue;
                    const float hF = (float)fq[q] - cF, hG = (float)gq[q] - cG,
                                hK = (float)kq[q] - cK;
                    const float Uh = fmaf(uF, hF, fmaf(uG, hG, fmaf(uK, hK, u3)));
                    const float Vh = fmaf(vF, hF, fmaf(vG, hG, fmaf(vK, hK, v3)));
                    const float Wh = fmaf(wF, hF, fmaf(wG, hG, fmaf(wK, hK, w3)));
                    const float rw = __builtin_amdgcn_rcpf(Wh);
                    const float u = Uh * rw, v = Vh * rw;
                    const int i = floor_i32(u) - (iu0 - 1), j = floor_i32(v) - (iv0 - 1);
                    if ((unsigned)i > (unsigned)kBT || (unsigned)j > (unsigned)kBT) continue;
                    const float fu = __builtin_amdgcn_fractf(u), fv = __builtin_amdgcn_fractf(v);
                    const float t = FDK ? vs * (rw * rw) : vs;
                    const float t1 = t * fu, t0 = t - t1;
                    const float t01 = t0 * fv, t00 = t0 - t01;
                    const float t11 = t1 * fv, t10 = t1 - t11;
                    unsigned int *c = &cells[j * kBTP + i];
                    atomicAdd(c, (unsigned)round_i32(t00));
                    atomicAdd(c + 1, (unsigned)round_i32(t10));
                    atomicAdd(c + kBTP, (unsigned)round_i32(t01));
                    atomicAdd(c + kBTP + 1, (unsigned)round_i32(t11));
                }
            }
        }
    }
    __syncthreads();
    own += (long long)(int)*mine;
    const int iu = iu0 + tid % kBT, iv = iv0 + tid / kBT;
    if (iu < U && iv < V) {
        const double unit = (double)vmax * (double)wbound * 0x1p-18;
        sino[((b * A + a) * V + iv) * (size_t)U + iu] = from_f32<SS>((float)((double)own * unit));
    }
}

// ---------------------------------------------------------------- host helpers -----------

// True when no angle's U or W row depends on y, so k_bp3d_y applies.
bool y_invariant(const std::vector<float> &m, int64_t A) {
    for (int64_t a = 0; a < A; ++a)
        if (m[12 * a + 1] != 0.0f || m[12 * a + 9] != 0.0f) return false;
    return true;
}

// Smallest W = M[a][8..11] . (h, 1) over the volume's eight corner voxels (h = index - centre)
// and all angles, in FP64 from the FP32 matrices.
double corner_wmin(const cone3d::Desc &g, const std::vector<float> &m, int64_t A) {
    const double c[3] = {((double)g.nx - 1.0) / 2.0, ((double)g.ny - 1.0) / 2.0,
                         ((double)g.nz - 1.0) / 2.0};
    double w = INFINITY;
    for (int64_t a = 0; a < A; ++a) {
        const float *r = m.data() + 12 * a + 8;
        for (int k = 0; k < 8; ++k)
            w = std::min(w, (double)r[3] + (k & 1 ? c[0] : -c[0]) * (double)r[0] +
                                (k & 2 ? c[1] : -c[1]) * (double)r[1] +
                                (k & 4 ? c[2] : -c[2]) * (double)r[2]);
    }
    return w;
}

// Projection matrices for the backprojector, rows (U, V, W) over h = (ix-cx, iy-cy, iz-cz, 1),
// scaled so that W = 1 at the volume centre.
//   reference: the reference's frame (backward_3d.hpp:79-94): dir = O + ix dx + iy dy + iz dz
//              - s, t = (DSO - DSD - s.z) / dir.z, u = dir.x t + s.x + floor(U/2) (likewise
//              v); the lower-tap coordinate is u - 0.5.
//   matched:   the forward frame: voxel P = src + t (org + iu du + iv dv - src) solved for
//              (t, t iu, t iv); the lower-tap coordinate of pixel centre iu is iu.
// Rows are scaled by `fixed_scale` if it is positive, else so that W = 1 at the volume centre.
void finish_matrix(const cone3d::Desc &g, double rows[3][4], float *out, double fixed_scale) {
    const double c[3] = {((double)g.nx - 1.0) / 2.0, ((double)g.ny - 1.0) / 2.0,
                         ((double)g.nz - 1.0) / 2.0};
    for (int r = 0; r < 3; ++r)
        rows[r][3] += rows[r][0] * c[0] + rows[r][1] * c[1] + rows[r][2] * c[2];
    const double w = rows[2][3];
    const double sc = fixed_scale > 0.0 ? fixed_scale
                                        : ((w != 0.0 && std::isfinite(w)) ? 1.0 / w : 1.0);
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 4; ++k) out[4 * r + k] = (float)(rows[r][k] * sc);
}

void reference_matrices(const cone3d::Desc &g, const double *phi, const double *theta,
                        const double *psi, int64_t A, std::vector<float> &mref) {
    mref.assign(12 * A, 0.0f);
    const double hu = (double)(g.nu / 2), hv = (double)(g.nv / 2);
    for (int64_t a = 0; a < A; ++a) {
        const cone3d::BpGeom bg = cone3d::bp_angle(g, phi[a], theta[a], psi[a]);
        const double cc = (g.dso - g.dsd) - bg.src[2];
        double d0[3];
        for (int d = 0; d < 3; ++d) d0[d] = bg.org[d] - bg.src[d];
        double rr[3][4];
        const double rowW[4] = {bg.dx[2], bg.dy[2], bg.dz[2], d0[2]};
        const double rowX[4] = {bg.dx[0], bg.dy[0], bg.dz[0], d0[0]};
        const double rowY[4] = {bg.dx[1], bg.dy[1], bg.dz[1], d0[1]};
        for (int k = 0; k < 4; ++k) {
            rr[0][k] = cc * rowX[k] + (bg.src[0] + hu - 0.5) * rowW[k];
            rr[1][k] = cc * rowY[k] + (bg.src[1] + hv - 0.5) * rowW[k];
            rr[2][k] = rowW[k];
        }
        finish_matrix(g, rr, mref.data() + 12 * a, 0.0);
    }
}

// Matched matrices from the forward frames; with `mfdk`, also the FDK ones: W = t (the voxel's
// parameter on its ray, 1 at the panel) times DSD / DSO, so that W = depth / DSO.
void matched_matrices(const cone3d::Desc &g, const std::vector<FwdGeom> &fgeo,
                      std::vector<float> &mmat, std::vector<float> *mfdk) {
    const int64_t A = (int64_t)fgeo.size();
    mmat.assign(12 * A, 0.0f);
    if (mfdk) mfdk->assign(12 * A, 0.0f);
    for (int64_t a = 0; a < A; ++a) {
        const FwdGeom &fg = fgeo[a];
        double Mc[3][3]; // columns org - src, du, dv
        for (int d = 0; d < 3; ++d) {
            Mc[d][0] = fg.org[d] - fg.src[d];
            Mc[d][1] = fg.du[d];
            Mc[d][2] = fg.dv[d];
        }
        const double det = Mc[0][0] * (Mc[1][1] * Mc[2][2] - Mc[1][2] * Mc[2][1]) -
                           Mc[0][1] * (Mc[1][0] * Mc[2][2] - Mc[1][2] * Mc[2][0]) +
                           Mc[0][2] * (Mc[1][0] * Mc[2][1] - Mc[1][1] * Mc[2][0]);
        double inv[3][3];
        inv[0][0] = (Mc[1][1] * Mc[2][2] - Mc[1][2] * Mc[2][1]) / det;
        inv[0][1] = (Mc[0][2] * Mc[2][1] - Mc[0][1] * Mc[2][2]) / det;
        inv[0][2] = (Mc[0][1] * Mc[1][2] - Mc[0][2] * Mc[1][1]) / det;
        inv[1][0] = (Mc[1][2] * Mc[2][0] - Mc[1][0] * Mc[2][2]) / det;
        inv[1][1] = (Mc[0][0] * Mc[2][2] - Mc[0][2] * Mc[2][0]) / det;
        inv[1][2] = (Mc[0][2] * Mc[1][0] - Mc[0][0] * Mc[1][2]) / det;
        inv[2][0] = (Mc[1][0] * Mc[2][1] - Mc[1][1] * Mc[2][0]) / det;
        inv[2][1] = (Mc[0][1] * Mc[2][0] - Mc[0][0] * Mc[2][1]) / det;
        inv[2][2] = (Mc[0][0] * Mc[1][1] - Mc[0][1] * Mc[1][0]) / det;
        // rows ordered (t iu, t iv, t) = (U, V, W)
        const int order[3] = {1, 2, 0};
        double rr[3][4], rk[3][4];
        for (int r = 0; r < 3; ++r) {
            const double *ir = inv[order[r]];
            rr[r][0] = ir[0];
            rr[r][1] = ir[1];
            rr[r][2] = ir[2];
            rr[r][3] = -(ir[0] * fg.src[0] + ir[1] * fg.src[1] + ir[2] * fg.src[2]);
            for (int k = 0; k < 4; ++k) rk[r][k] = rr[r][k];
        }
        finish_matrix(g, rr, mmat.data() + 12 * a, 0.0);
        if (mfdk) finish_matrix(g, rk, mfdk->data() + 12 * a, g.dsd / g.dso);
    }
}

void bp_matrices(const cone3d::Desc &g, const double *phi, const double *theta, const double *psi,
                 int64_t A, std::vector<float> &mref, std::vector<float> &mmat,
                 std::vector<float> &mfdk) {
    reference_matrices(g, phi, theta, psi, A, mref);
    std::vector<FwdGeom> fgeo(A);
    for (int64_t a = 0; a < A; ++a) fgeo[a] = cone3d::fwd_angle(g, phi[a], theta[a], psi[a]);
    matched_matrices(g, fgeo, mmat, &mfdk);
}

// Python order (z, y, x) / (v, u) -> (x, y, z) / (u, v), as forward.cu:113-124
cone3d::Desc desc_of(const cmi_cone3d_desc *d) {
    cone3d::Desc g{};
    g.nx = (int64_t)d->vol_shape[2];
    g.ny = (int64_t)d->vol_shape[1];
    g.nz = (int64_t)d->vol_shape[0];
    g.sx = d->vol_spacing[2];
    g.sy = d->vol_spacing[1];
    g.sz = d->vol_spacing[0];
    g.ox = d->vol_offset[2];
    g.oy = d->vol_offset[1];
    g.oz = d->vol_offset[0];
    g.nu = (int64_t)d->det_shape[1];
    g.nv = (int64_t)d->det_shape[0];
    g.du = d->det_spacing[1];
    g.dv = d->det_spacing[0];
    g.off_u = d->det_offset[1];
    g.off_v = d->det_offset[0];
    g.pitch = d->det_rotation[0]; // forward.cu:136-138
    g.roll = d->det_rotation[1];
    g.yaw = d->det_rotation[2];
    g.dso = d->DSO;
    g.dsd = d->DSD;
    g.cor = d->COR;

    return g;
}

cmi_status check_desc(const cmi_cone3d_desc *d, const char *fn) {
    if (!d) return fail(CMI_ERR_INVALID_ARGUMENT, "%s: descriptor is null", fn);
    if (!d->phi || !d->theta || !d->psi)
        return fail(CMI_ERR_INVALID_ARGUMENT, "%s: null angle array", fn);
    for (int i = 0; i < 3; ++i)
        if (d->vol_shape[i] == 0 || d->vol_shape[i] > 4096 || !(d->vol_spacing[i] > 0))
            return fail(CMI_ERR_INVALID_ARGUMENT,
                        "%s: volume shape / spacing out of range (axis %d)", fn, i);
    for (int i = 0; i < 2; ++i)
        if (d->det_shape[i] == 0 || d->det_shape[i] > 16384 || !(d->det_spacing[i] > 0))
            return fail(CMI_ERR_INVALID_ARGUMENT,
                        "%s: detector shape / spacing out of range (axis %d)", fn, i);
    if (d->nangles == 0 || d->nangles > 65535)
        return fail(CMI_ERR_INVALID_ARGUMENT, "%s: nangles must be in [1, 65535]", fn);
    if (!std::isfinite(d->DSO) || !std::isfinite(d->DSD) || !std::isfinite(d->COR))
        return fail(CMI_ERR_INVALID_ARGUMENT, "%s: non-finite distance", fn);
    for (uint64_t a = 0; a < d->nangles; ++a)
        if (!std::isfinite(d->phi[a]) || !std::isfinite(d->theta[a]) || !std::isfinite(d->psi[a]))
            return fail(CMI_ERR_INVALID_ARGUMENT, "%s: non-finite angle %llu", fn,
                        (unsigned long long)a);
    return CMI_OK;
}

template <class F>
cmi_status dispatch2(int dt_a, int dt_b, F &&f) {
    auto inner = [&](auto ta) -> cmi_status {
        switch (dt_b) {
        case CMI_F32: return f(ta, F32{});
        case CMI_F16: return f(ta, F16{});
        case CMI_BF16: return f(ta, BF16{});
        default: return fail(CMI_ERR_DTYPE, "unsupported dtype %d", dt_b);
        }
    };
    switch (dt_a) {
    case CMI_F32: return inner(F32{});
    case CMI_F16: return inner(F16{});
    case CMI_BF16: return inner(BF16{});
    default: return fail(CMI_ERR_DTYPE, "unsupported dtype %d", dt_a);
    }
}

cmi_status check_call(const cmi_plan3d *p, const void *vol, int vol_dtype, const void *sino,
                      int sino_dtype, uint64_t batch, const char *fn) {
    if (!p) return fail(CMI_ERR_INVALID_ARGUMENT, "%s: plan is null", fn);
    if (!vol) return fail(CMI_ERR_INVALID_ARGUMENT, "%s: volume pointer is null", fn);
    if (!sino) return fail(CMI_ERR_INVALID_ARGUMENT, "%s: sinogram pointer is null", fn);
    if (batch == 0) return fail(CMI_ERR_INVALID_ARGUMENT, "%s: batch must be >= 1", fn);
    if (!dtype_ok(vol_dtype) || !dtype_ok(sino_dtype))
        return fail(CMI_ERR_DTYPE, "%s: unsupported dtype (volume %d, sinogram %d)", fn,
                    vol_dtype, sino_dtype);
    if (vol_dtype != p->vol_prec)
        return fail(CMI_ERR_DTYPE, "%s: plan precision does not match volume precision (%s)",
                    fn, dtype_name(vol_dtype));
    if (sino_dtype != p->det_prec)
        return fail(CMI_ERR_DTYPE,
                    "%s: plan precision does not match sinogram precision (%s)", fn,
                    dtype_name(sino_dtype));
    if (batch * (uint64_t)p->A > 65535 ||
        batch * (uint64_t)((p->g.nz + kZB - 1) / kZB) > 65535 ||
        batch * (uint64_t)(p->g.nz + 2 * kPad) > 65535)
        return fail(CMI_ERR_INVALID_ARGUMENT, "%s: batch too large for one launch", fn);
    CMI_TRY(check_device_ptr(vol, p->device, "volume"));
    CMI_TRY(check_device_ptr(sino, p->device, "sinogram"));
    return CMI_OK;
}

// Shared tail of the plan constructors: per-angle fast axes and volume copies, matrix layouts,
// device tables and the total sample count (k_count3d). `mfdk` is empty for plans from views.
cmi_status finish_plan3d(std::unique_ptr<cmi_plan3d> p, const std::vector<FwdGeom> &fgeo,
                         const std::vector<float> &mref, const std::vector<float> &mmat,
                         const std::vector<float> &mfdk, cmi_plan3d **plan) {
    const cone3d::Desc &g = p->g;
    const int64_t A = p->A;
    std::vector<int> fast(A);
    for (int64_t a = 0; a < A; ++a) {
        const double *v = fgeo[a].dv;
        const int f = std::fabs(v[1]) >= std::fabs(v[0]) && std::fabs(v[1]) >= std::fabs(v[2]) ? 1
                      : std::fabs(v[2]) > std::fabs(v[0])                                   ? 2
                                                                                            : 0;
        fast[a] = f;
        if (p->copy_slot[f] < 0) p->copy_slot[f] = p->ncopies++;
    }
    p->ycols[CMI_BP3D_REFERENCE] = y_invariant(mref, A);
    p->ycols[CMI_BP3D_MATCHED] = y_invariant(mmat, A);
    p->ycols[CMI_BP3D_FDK] = !mfdk.empty() && y_invariant(mfdk, A);
    p->wmin[CMI_BP3D_REFERENCE] = corner_wmin(g, mref, A);
    p->wmin[CMI_BP3D_MATCHED] = corner_wmin(g, mmat, A);
    p->wmin[CMI_BP3D_FDK] = mfdk.empty() ? 0.0 : corner_wmin(g, mfdk, A);

    DeviceGuard guard;
    CMI_TRY(guard.enter(p->device));
    struct Cleanup {
        cmi_plan3d *p;
        bool armed = true;
        ~Cleanup() {
            if (!armed) return;
            (void)hipFree(p->d_fgeo);
            (void)hipFree(p->d_fast);
            (void)hipFree(p->d_mref);
            (void)hipFree(p->d_mmat);
            (void)hipFree(p->d_mfdk);
        }
    } cleanup{p.get()};
    CMI_HIP_TRY(hipMalloc(&p->d_fgeo, sizeof(FwdGeom) * A));
    CMI_HIP_TRY(hipMalloc(&p->d_fast, sizeof(int) * A));
    CMI_HIP_TRY(hipMalloc(&p->d_mref, sizeof(float) * 12 * A));
    CMI_HIP_TRY(hipMalloc(&p->d_mmat, sizeof(float) * 12 * A));
    if (!mfdk.empty()) CMI_HIP_TRY(hipMalloc(&p->d_mfdk, sizeof(float) * 12 * A));
    p->table_bytes = (sizeof(FwdGeom) + sizeof(int) + (mfdk.empty() ? 2 : 3) * 12 * sizeof(float)) * A;
    unsigned long long *d_total = nullptr;
    CMI_HIP_TRY(hipMalloc(&d_total, sizeof(unsigned long long)));
    std::unique_ptr<unsigned long long, void (*)(unsigned long long *)> total_guard(
        d_total, [](unsigned long long *q) { (void)hipFree(q); });
    hipStream_t s;
    CMI_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::unique_ptr<ihipStream_t, void (*)(hipStream_t)> stream_guard(
        s, [](hipStream_t q) { (void)hipStreamDestroy(q); });
    CMI_HIP_TRY(hipMemcpyAsync(p->d_fgeo, fgeo.data(), sizeof(FwdGeom) * A,
                               hipMemcpyHostToDevice, s));
    CMI_HIP_TRY(hipMemcpyAsync(p->d_fast, fast.data(), sizeof(int) * A, hipMemcpyHostToDevice,
                               s));
    CMI_HIP_TRY(hipMemcpyAsync(p->d_mref, mref.data(), sizeof(float) * 12 * A,
                               hipMemcpyHostToDevice, s));
    CMI_HIP_TRY(hipMemcpyAsync(p->d_mmat, mmat.data(), sizeof(float) * 12 * A,
                               hipMemcpyHostToDevice, s));
    if (!mfdk.empty())
        CMI_HIP_TRY(hipMemcpyAsync(p->d_mfdk, mfdk.data(), sizeof(float) * 12 * A,
                                   hipMemcpyHostToDevice, s));
    CMI_HIP_TRY(hipMemsetAsync(d_total, 0, sizeof(unsigned long long), s));
    dim3 blk(256), grd((unsigned)((g.nu + 255) / 256), (unsigned)g.nv, (unsigned)A);
    hipLaunchKernelGGL(k_count3d, grd, blk, 0, s, p->d_fgeo, (int)g.nu, (int)g.nv, (int)g.nx,
                       (int)g.ny, (int)g.nz, d_total);
    CMI_HIP_TRY(hipGetLastError());
    unsigned long long total = 0;
    CMI_HIP_TRY(hipMemcpyAsync(&total, d_total, sizeof total, hipMemcpyDeviceToHost, s));
    CMI_HIP_TRY(hipStreamSynchronize(s));
    p->total_samples = total;
    cleanup.armed = false;
    *plan = p.release();
    return CMI_OK;
}

} // namespace

// ===================================================================== C-ABI ===============

extern "C" {

CMI_API cmi_status cmi_plan3d_create(int device, int vol_prec, int det_prec,
                                     const cmi_cone3d_desc *d, cmi_plan3d **plan) {
    clear_error();
    if (!plan) return fail(CMI_ERR_INVALID_ARGUMENT, "plan3d_create: output pointer is null");
    *plan = nullptr;
    CMI_TRY(check_desc(d, "plan3d_create"));
    if (!dtype_ok(vol_prec) || !dtype_ok(det_prec))
        return fail(CMI_ERR_DTYPE, "plan3d_create: precision must be f32, f16 or bf16");
    const uint64_t padvol = (d->vol_shape[0] + 2 * kPad) * (d->vol_shape[1] + 2 * kPad) *
                            (d->vol_shape[2] + 2 * kPad);
    const uint64_t padpan = (d->det_shape[0] + 2 * kPad) * (d->det_shape[1] + 2 * kPad) *
                            d->nangles;
    if (padvol * 4 >= (1ull << 31) || padpan * 4 >= (1ull << 31))
        return fail(CMI_ERR_INVALID_ARGUMENT,
                    "plan3d_create: volume or projection stack exceeds 2 GiB");

    auto p = std::make_unique<cmi_plan3d>();
    p->device = device;
    p->vol_prec = vol_prec;
    p->det_prec = det_prec;
    p->g = desc_of(d);
    const cone3d::Desc &g = p->g;
    p->A = (int64_t)d->nangles;
    const int64_t A = p->A;
    std::vector<FwdGeom> fgeo(A);
    for (int64_t a = 0; a < A; ++a) fgeo[a] = cone3d::fwd_angle(g, d->phi[a], d->theta[a], d->psi[a]);
    std::vector<float> mref, mmat, mfdk;
    reference_matrices(g, d->phi, d->theta, d->psi, A, mref);
    matched_matrices(g, fgeo, mmat, &mfdk);
    return finish_plan3d(std::move(p), fgeo, mref, mmat, mfdk, plan);
}

CMI_API cmi_status cmi_plan3d_create_vec(int device, int vol_prec, int det_prec,
                                         const uint64_t vol_shape[3], const double vol_spacing[3],
                                         const double vol_offset[3], const uint64_t det_shape[2],
                                         const double *vecs, uint64_t nangles,
                                         cmi_plan3d **plan) {
    clear_error();
    if (!plan) return fail(CMI_ERR_INVALID_ARGUMENT, "plan3d_create_vec: output pointer is null");
    *plan = nullptr;
    if (!vol_shape || !vol_spacing || !vol_offset || !det_shape || !vecs)
        return fail(CMI_ERR_INVALID_ARGUMENT, "plan3d_create_vec: null geometry array");
    if (!dtype_ok(vol_prec) || !dtype_ok(det_prec))
        return fail(CMI_ERR_DTYPE, "plan3d_create_vec: precision must be f32, f16 or bf16");
    // the shape, spacing and size limits of cmi_plan3d_create, through a descriptor
    const double zero3[3] = {0.0, 0.0, 0.0};
    cmi_cone3d_desc d{};
    for (int i = 0; i < 3; ++i) {
        d.vol_shape[i] = vol_shape[i];
        d.vol_spacing[i] = vol_spacing[i];
        d.vol_offset[i] = vol_offset[i];
    }
    for (int i = 0; i < 2; ++i) {
        d.det_shape[i] = det_shape[i];
        d.det_spacing[i] = 1.0;
    }
    d.DSO = d.DSD = 1.0;
    d.nangles = nangles;
    d.phi = d.theta = d.psi = zero3; // not read past index 0 below; angles come from the views
    if (nangles == 0 || nangles > 65535)
        return fail(CMI_ERR_INVALID_ARGUMENT, "plan3d_create_vec: nangles must be in [1, 65535]");
    {
        cmi_cone3d_desc d1 = d;
        d1.nangles = 1;
        CMI_TRY(check_desc(&d1, "plan3d_create_vec"));
    }
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(vol_offset[i]))
            return fail(CMI_ERR_INVALID_ARGUMENT, "plan3d_create_vec: non-finite volume offset");
    const uint64_t padvol = (vol_shape[0] + 2 * kPad) * (vol_shape[1] + 2 * kPad) *
                            (vol_shape[2] + 2 * kPad);
    const uint64_t padpan = (det_shape[0] + 2 * kPad) * (det_shape[1] + 2 * kPad) * nangles;
    if (padvol * 4 >= (1ull << 31) || padpan * 4 >= (1ull << 31))
        return fail(CMI_ERR_INVALID_ARGUMENT,
                    "plan3d_create_vec: volume or projection stack exceeds 2 GiB");
    for (uint64_t a = 0; a < nangles; ++a) {
        const double *v = vecs + 12 * a;
        for (int i = 0; i < 12; ++i)
            if (!std::isfinite(v[i]))
                return fail(CMI_ERR_INVALID_ARGUMENT, "plan3d_create_vec: non-finite view %llu",
                            (unsigned long long)a);
        // the panel normal u x v must be non-zero and the source off the panel plane
        const double n[3] = {v[7] * v[11] - v[8] * v[10], v[8] * v[9] - v[6] * v[11],
                             v[6] * v[10] - v[7] * v[9]};
        const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        const double L = (n[0] * (v[3] - v[0]) + n[1] * (v[4] - v[1]) + n[2] * (v[5] - v[2]));
        const double ul = std::sqrt(v[6] * v[6] + v[7] * v[7] + v[8] * v[8]);
        const double vl = std::sqrt(v[9] * v[9] + v[10] * v[10] + v[11] * v[11]);
        if (!(nn > 1e-12 * ul * vl) || !(std::fabs(L) > 1e-9 * nn))
            return fail(CMI_ERR_INVALID_ARGUMENT,
                        "plan3d_create_vec: view %llu has parallel or zero u / v steps or its "
                        "source on the panel plane",
                        (unsigned long long)a);
    }
    auto p = std::make_unique<cmi_plan3d>();
    p->device = device;
    p->vol_prec = vol_prec;
    p->det_prec = det_prec;
    p->views = true;
    p->g = desc_of(&d);
    p->g.du = std::sqrt(vecs[6] * vecs[6] + vecs[7] * vecs[7] + vecs[8] * vecs[8]);
    p->g.dv = std::sqrt(vecs[9] * vecs[9] + vecs[10] * vecs[10] + vecs[11] * vecs[11]);
    p->g.dso = p->g.dsd = 0.0; // no circular-orbit distances: no FDK weighting
    p->A = (int64_t)nangles;
    const int64_t A = p->A;
    std::vector<FwdGeom> fgeo(A);
    for (int64_t a = 0; a < A; ++a) fgeo[a] = cone3d::fwd_view(p->g, vecs + 12 * a);
    std::vector<float> mmat;
    matched_matrices(p->g, fgeo, mmat, nullptr);
    // reference mode of a plan from views projects like matched mode (the reference's own
    // backprojection frame is defined by Euler angles only)
    return finish_plan3d(std::move(p), fgeo, mmat, mmat, {}, plan);
}

CMI_API cmi_status cmi_cone3d_bp_matrices(const cmi_cone3d_desc *d, int mode, float *out,
                                          int *y_invariant_out) {
    clear_error();
    CMI_TRY(check_desc(d, "cone3d_bp_matrices"));
    if (!out) return fail(CMI_ERR_INVALID_ARGUMENT, "cone3d_bp_matrices: null output");
    if (mode != CMI_BP3D_REFERENCE && mode != CMI_BP3D_MATCHED && mode != CMI_BP3D_FDK)
        return fail(CMI_ERR_INVALID_ARGUMENT, "cone3d_bp_matrices: unknown mode %d", mode);
    const cone3d::Desc g = desc_of(d);
    std::vector<float> mref, mmat, mfdk;
    bp_matrices(g, d->phi, d->theta, d->psi, (int64_t)d->nangles, mref, mmat, mfdk);
    const std::vector<float> &m = mode == CMI_BP3D_MATCHED ? mmat : mode == CMI_BP3D_FDK ? mfdk : mref;
    std::copy(m.begin(), m.end(), out);
    if (y_invariant_out) *y_invariant_out = y_invariant(m, (int64_t)d->nangles) ? 1 : 0;
    return CMI_OK;
}

CMI_API cmi_status cmi_plan3d_destroy(cmi_plan3d *plan) {
    clear_error();
    if (!plan) return CMI_OK;
    DeviceGuard guard;
    const cmi_status st = guard.enter(plan->device);
    if (st == CMI_OK) {
        (void)hipDeviceSynchronize();
        (void)hipFree(plan->d_fgeo);
        (void)hipFree(plan->d_fast);
        (void)hipFree(plan->d_mref);
        (void)hipFree(plan->d_mmat);
        (void)hipFree(plan->d_mfdk);
    }
    delete plan;
    return st;
}

CMI_API cmi_status cmi_plan3d_get_info(const cmi_plan3d *plan, cmi_plan3d_info *info) {
    clear_error();
    if (!plan || !info) return fail(CMI_ERR_INVALID_ARGUMENT, "plan3d_get_info: null argument");
    info->device = plan->device;
    info->vol_prec = plan->vol_prec;
    info->det_prec = plan->det_prec;
    info->nx = (uint64_t)plan->g.nx;
    info->ny = (uint64_t)plan->g.ny;
    info->nz = (uint64_t)plan->g.nz;
    info->nu = (uint64_t)plan->g.nu;
    info->nv = (uint64_t)plan->g.nv;
    info->nangles = (uint64_t)plan->A;
    info->total_samples = plan->total_samples;
    info->table_bytes = plan->table_bytes;
    return CMI_OK;
}

CMI_API cmi_status cmi_forward_3d(const cmi_plan3d *plan, const void *volume, int vol_dtype,
                                  void *sinogram, int sino_dtype, uint64_t batch,
                                  void *stream) {
    clear_error();
    CMI_TRY(check_call(plan, volume, vol_dtype, sinogram, sino_dtype, batch, "forward_3d"));
    DeviceGuard guard;
    CMI_TRY(guard.enter(plan->device));
    const hipStream_t s = (hipStream_t)stream;
    const cone3d::Desc &g = plan->g;
    const int nx = (int)g.nx, ny = (int)g.ny, nz = (int)g.nz, U = (int)g.nu, V = (int)g.nv,
              A = (int)plan->A;
    const int Wp = nx + 2 * kPad, Hp = ny + 2 * kPad, Dp = nz + 2 * kPad;
    void *scratch = nullptr;
    const size_t one = (size_t)Wp * Hp * Dp, nc = (size_t)plan->ncopies;
    CMI_HIP_TRY(hipMallocAsync(&scratch, one * nc * dtype_size(vol_dtype) * batch, s));
    const cmi_status st = dispatch2(vol_dtype, sino_dtype, [&](auto si, auto ss) -> cmi_status {
        using SI = decltype(si);
        using SS = decltype(ss);
        using RI = typename SI::raw;
        for (uint64_t b = 0; b < batch; ++b) {
            const RI *src = (const RI *)volume + b * (size_t)nx * ny * nz;
            RI *copies = (RI *)scratch + b * nc * one;
            if (plan->copy_slot[0] >= 0) {
                dim3 pb(256), pg((unsigned)((Wp + 255) / 256), (unsigned)Hp, (unsigned)Dp);
                hipLaunchKernelGGL((k_stage3d<SI>), pg, pb, 0, s, src,
                                   copies + plan->copy_slot[0] * one, nx, ny, nz, Wp, Hp, Dp);
            }
            if (plan->copy_slot[1] >= 0) {
                dim3 tb(32, 8), tg((unsigned)((Wp + 31) / 32), (unsigned)((Hp + 31) / 32),
                                   (unsigned)Dp);
                hipLaunchKernelGGL((k_stage3d_t<SI, 1>), tg, tb, 0, s, src,
                                   copies + plan->copy_slot[1] * one, nx, ny, nz, Wp, Hp, Dp);
            }
            if (plan->copy_slot[2] >= 0) {
                dim3 tb(32, 8), tg((unsigned)((Wp + 31) / 32), (unsigned)((Dp + 31) / 32),
                                   (unsigned)Hp);
                hipLaunchKernelGGL((k_stage3d_t<SI, 2>), tg, tb, 0, s, src,
                                   copies + plan->copy_slot[2] * one, nx, ny, nz, Wp, Hp, Dp);
            }
        }
        dim3 fb(kMaxBlock), fg((unsigned)((U + 15) / 16), (unsigned)((V + 15) / 16),
                               (unsigned)(A * batch));
        hipLaunchKernelGGL((k_fwd3d<SI, SS>), fg, fb, 0, s, (const RI *)scratch, plan->d_fgeo,
                           plan->d_fast, plan->copy_slot[0], plan->copy_slot[1],
                           plan->copy_slot[2], plan->ncopies, nx, ny, nz, U, V, A,
                           (typename SS::raw *)sinogram);
        CMI_HIP_TRY(hipGetLastError());
        return CMI_OK;
    });
    const hipError_t fe = hipFreeAsync(scratch, s);
    if (st != CMI_OK) return st;
    CMI_HIP_TRY(fe);
    return CMI_OK;
}

CMI_API cmi_status cmi_backward_3d(const cmi_plan3d *plan, void *volume, int vol_dtype,
                                   const void *sinogram, int sino_dtype, uint64_t batch,
                                   int mode, void *stream) {
    clear_error();
    CMI_TRY(check_call(plan, volume, vol_dtype, sinogram, sino_dtype, batch, "backward_3d"));
    if (mode != CMI_BP3D_REFERENCE && mode != CMI_BP3D_MATCHED && mode != CMI_BP3D_FDK)
        return fail(CMI_ERR_INVALID_ARGUMENT, "backward_3d: unknown mode %d", mode);
    if (mode == CMI_BP3D_FDK && plan->views)
        return fail(CMI_ERR_NOT_SUPPORTED,
                    "backward_3d: FDK weighting needs a circular orbit (DSO, DSD); this plan "
                    "was built from views");
    DeviceGuard guard;
    CMI_TRY(guard.enter(plan->device));
    const hipStream_t s = (hipStream_t)stream;
    const cone3d::Desc &g = plan->g;
    const int nx = (int)g.nx, ny = (int)g.ny, nz = (int)g.nz, U = (int)g.nu, V = (int)g.nv,
              A = (int)plan->A;
    const int Up = U + 2 * kPad, Vp = V + 2 * kPad;
    void *scratch = nullptr;
    CMI_HIP_TRY(
        hipMallocAsync(&scratch, (size_t)Up * Vp * A * dtype_size(sino_dtype) * batch, s));
    const float cx = (float)((nx - 1) / 2.0), cy = (float)((ny - 1) / 2.0),
                cz = (float)((nz - 1) / 2.0);
    const float *M = mode == CMI_BP3D_MATCHED ? plan->d_mmat
                     : mode == CMI_BP3D_FDK   ? plan->d_mfdk
                                              : plan->d_mref;
    const cmi_status st = dispatch2(vol_dtype, sino_dtype, [&](auto sv, auto ss) -> cmi_status {
        using SV = decltype(sv);
        using SS = decltype(ss);
        using RS = typename SS::raw;
        const uint64_t n = (uint64_t)Up * Vp * A * batch;
        hipLaunchKernelGGL((k_stage_panel<SS>), dim3((unsigned)((n + 4 * 256 - 1) / (4 * 256))),
                           dim3(256), 0, s, (const RS *)sinogram, (RS *)scratch, U, V, n);
        if (plan->ycols[mode]) {
            const int nzt = (nz + kTileZ - 1) / kTileZ;
            dim3 bb(kTileX * kTileZ), bg((unsigned)((nx + kTileX - 1) / kTileX),
                                         (unsigned)((ny + kYB - 1) / kYB), (unsigned)(nzt * batch));
            auto kern = mode == CMI_BP3D_MATCHED ? k_bp3d_y<SV, SS, true, false>
                        : mode == CMI_BP3D_FDK   ? k_bp3d_y<SV, SS, true, true>
                                                 : k_bp3d_y<SV, SS, false, false>;
            hipLaunchKernelGGL(kern, bg, bb, 0, s, (typename SV::raw *)volume, (const RS *)scratch,
                               M, nx, ny, nz, U, V, A, cx, cy, cz);
        } else {
            const int nzb = (nz + kZB - 1) / kZB;
            dim3 bb(16, 16), bg((unsigned)((nx + 15) / 16), (unsigned)((ny + 15) / 16),
                                (unsigned)(nzb * batch));
            auto kern = mode == CMI_BP3D_MATCHED ? k_bp3d<SV, SS, true, false>
                        : mode == CMI_BP3D_FDK   ? k_bp3d<SV, SS, true, true>
                                                 : k_bp3d<SV, SS, false, false>;
            hipLaunchKernelGGL(kern, bg, bb, 0, s, (typename SV::raw *)volume,
                               (const RS *)scratch, M, nx, ny, nz, U, V, A, cx, cy, cz);
        }
        CMI_HIP_TRY(hipGetLastError());
        return CMI_OK;
    });
    const hipError_t fe = hipFreeAsync(scratch, s);
    if (st != CMI_OK) return st;
    CMI_HIP_TRY(fe);
    return CMI_OK;
}

CMI_API cmi_status cmi_adjoint_3d(const cmi_plan3d *plan, void *volume, int vol_dtype,
                                  const void *sinogram, int sino_dtype, uint64_t batch,
                                  void *stream) {
    clear_error();
    CMI_TRY(check_call(plan, volume, vol_dtype, sinogram, sino_dtype, batch, "adjoint_3d"));
    DeviceGuard guard;
    CMI_TRY(guard.enter(plan->device));
    const hipStream_t s = (hipStream_t)stream;
    const cone3d::Desc &g = plan->g;
    const int nx = (int)g.nx, ny = (int)g.ny, nz = (int)g.nz, U = (int)g.nu, V = (int)g.nv,
              A = (int)plan->A;
    const size_t rays = (size_t)A * U * V;
    // scratch: the march table [2][rays], [batch][kMaxParts] partial maxima of |y|, then (256-byte
    // aligned) the tiles' 64-bit sums, one tile of kAT^3 per workgroup
    const size_t ntiles = (size_t)((nx + kAT - 1) / kAT) * ((ny + kAT - 1) / kAT) *
                          ((nz + kAT - 1) / kAT) * batch;
    float4 *table = nullptr;
    CMI_HIP_TRY(hipMallocAsync((void **)&table,
                               2 * rays * sizeof(float4) + kMaxParts * batch * sizeof(float) +
                                   256 + ntiles * kAT * kAT * kAT * sizeof(long long),
                               s));
    float *parts = reinterpret_cast<float *>(table + 2 * rays);
    long long *totals = reinterpret_cast<long long *>(
        (reinterpret_cast<uintptr_t>(parts + kMaxParts * batch) + 255) & ~(uintptr_t)255);
    const cmi_status st = dispatch2(vol_dtype, sino_dtype, [&](auto sv, auto ss) -> cmi_status {
        using SV = decltype(sv);
        using SS = decltype(ss);
        dim3 tb(256), tg((unsigned)((U + 255) / 256), (unsigned)V, (unsigned)A);
        hipLaunchKernelGGL(k_ray_table3d, tg, tb, 0, s, plan->d_fgeo, nx, ny, nz, U, V, table,
                           table + rays);
        hipLaunchKernelGGL((k_absmax_parts<SS>), dim3(kMaxParts, (unsigned)batch),
                           dim3(kMaxBlock), 0, s, (const typename SS::raw *)sinogram, rays, parts);
        const int ntz = (nz + kAT - 1) / kAT;
        dim3 ab(kMaxBlock), ag((unsigned)((nx + kAT - 1) / kAT), (unsigned)((ny + kAT - 1) / kAT),
                               (unsigned)(ntz * batch));
        hipLaunchKernelGGL((k_adj3d<SV, SS>), ag, ab, 0, s, (typename SV::raw *)volume,
                           (const typename SS::raw *)sinogram, (const float4 *)table,
                           (const float4 *)(table + rays), plan->d_mmat, parts, nx, ny, nz,
                           U, V, A, totals);
        CMI_HIP_TRY(hipGetLastError());
        return CMI_OK;
    });
    const hipError_t fe = hipFreeAsync(table, s);
    if (st != CMI_OK) return st;
    CMI_HIP_TRY(fe);
    return CMI_OK;
}

CMI_API cmi_status cmi_backward_3d_transpose(const cmi_plan3d *plan, void *sinogram,
                                             int sino_dtype, const void *volume, int vol_dtype,
                                             uint64_t batch, int mode, void *stream) {
    clear_error();
    CMI_TRY(check_call(plan, volume, vol_dtype, sinogram, sino_dtype, batch,
                       "backward_3d_transpose"));
    if (mode != CMI_BP3D_REFERENCE && mode != CMI_BP3D_MATCHED && mode != CMI_BP3D_FDK)
        return fail(CMI_ERR_INVALID_ARGUMENT, "backward_3d_transpose: unknown mode %d", mode);
    if (mode == CMI_BP3D_FDK && plan->views)
        return fail(CMI_ERR_NOT_SUPPORTED,
                    "backward_3d_transpose: FDK weighting needs a circular orbit (DSO, DSD); "
                    "this plan was built from views");
    const double wmin = plan->wmin[mode];
    if (!(wmin > 0.0) || !std::isfinite(wmin))
        return fail(CMI_ERR_NOT_SUPPORTED,
                    "backward_3d_transpose: a voxel lies at or behind a source (W <= 0), so "
                    "the backprojection's projection is unbounded");
    const cone3d::Desc &g = plan->g;
    const int nx = (int)g.nx, ny = (int)g.ny, nz = (int)g.nz, U = (int)g.nu, V = (int)g.nv,
              A = (int)plan->A;
    const uint64_t tu = (uint64_t)(U + kBT - 1) / kBT, tv = (uint64_t)(V + kBT - 1) / kBT;
    if ((uint64_t)A * tu * kMaxBlock >= (1ull << 32) || tv * batch > 65535)
        return fail(CMI_ERR_INVALID_ARGUMENT,
                    "backward_3d_transpose: too many projections, tiles or volumes for one "
                    "launch");
    // the largest weight: 1 (bilinear), times 1 / W^2 for FDK; 1 % of margin for FP32
    const double wbound = 1.01 * (mode == CMI_BP3D_FDK ? 1.0 / (wmin * wmin) : 1.0);
    if (!(wbound < 1e30))
        return fail(CMI_ERR_NOT_SUPPORTED, "backward_3d_transpose: FDK weight out of range");
    DeviceGuard guard;
    CMI_TRY(guard.enter(plan->device));
    const hipStream_t s = (hipStream_t)stream;
    const int Wp = nx + 2 * kPad, Hp = ny + 2 * kPad, Dp = nz + 2 * kPad;
    const size_t one = (size_t)Wp * Hp * Dp;
    // scratch: per volume a y-fastest copy (read where the march axis is x), then kMaxParts
    // partial maxima of |vol| per volume
    const size_t copy_bytes = (one * dtype_size(vol_dtype) * batch + 255) / 256 * 256;
    void *scratch = nullptr;
    CMI_HIP_TRY(hipMallocAsync(&scratch, copy_bytes + kMaxParts * batch * sizeof(float), s));
    float *parts = (float *)((char *)scratch + copy_bytes);
    const float *M = mode == CMI_BP3D_MATCHED ? plan->d_mmat
                     : mode == CMI_BP3D_FDK   ? plan->d_mfdk
                                              : plan->d_mref;
    const cmi_status st = dispatch2(vol_dtype, sino_dtype, [&](auto sv, auto ss) -> cmi_status {
        using SV = decltype(sv);
        using SS = decltype(ss);
        using RV = typename SV::raw;
        for (uint64_t b = 0; b < batch; ++b) {
            dim3 tb(32, 8), tg((unsigned)((Wp + 31) / 32), (unsigned)((Hp + 31) / 32),
                               (unsigned)Dp);
            hipLaunchKernelGGL((k_stage3d_t<SV, 1>), tg, tb, 0, s,
                               (const RV *)volume + b * (size_t)nx * ny * nz,
                               (RV *)scratch + b * one, nx, ny, nz, Wp, Hp, Dp);
        }
        hipLaunchKernelGGL((k_absmax_parts<SV>), dim3(kMaxParts, (unsigned)batch),
                           dim3(kMaxBlock), 0, s, (const RV *)volume, (size_t)nx * ny * nz,
                           parts);
        auto kern = mode == CMI_BP3D_FDK ? k_bpT3d<SV, SS, true> : k_bpT3d<SV, SS, false>;
        hipLaunchKernelGGL(kern, dim3((unsigned)(A * tu), (unsigned)(tv * batch)),
                           dim3(kMaxBlock), 0, s, (typename SS::raw *)sinogram,
                           (const RV *)volume, (const RV *)scratch, M, parts, (float)wbound, nx,
                           ny, nz, U, V, A);
        CMI_HIP_TRY(hipGetLastError());
        return CMI_OK;
    });
    const hipError_t fe = hipFreeAsync(scratch, s);
    if (st != CMI_OK) return st;
    CMI_HIP_TRY(fe);
    return CMI_OK;
}

} // extern "C"
