// 2D fan-beam projectors for gfx950: plan, per-ray table, forward ray march, voxel-driven
// backprojection and the matched (transposed) ray march.
//
// Data layout in HBM
//   plan tables  geo  [A] AngleGeom (FP64)       source / cell-0 vector / cell step
//                meta [A] float4                 (-sin, cos, DSD, DSO) for the backprojector
//                                                (plans from angles / DSO / DSD)
//                bpv  [A][3] float4              (n.x, n.y, n0, L), (e.x, e.y, e0, c0 - .5),
//                                                (1/|u|, 0, 0, 0): the backprojector's view
//                                                constants (plans from ASTRA-like views,
//                                                geometry2d.hpp bp_view)
//                ageo [A] float[6]               FP32 copy of geo for the adjoint's ray search
//                ev   [A*D] float4               (mx, my, vx, vy) per ray: m is the sample
//                                                j = N/2 in centred padded coordinates (see
//                                                below), v the step
//                ns   [A*D] float4               (world step length, N as int bits, -N/2, 0)
// Sample j of a ray sits at m + (j - N/2) v. Coordinates are "centred padded": padded-copy
// texel coordinate minus 0.5 (so floor() is the lower tap) minus the integer centre c0 of the
// padded copy. Anchoring at the ray's midpoint near the image centre keeps |m| small. The step
// is +-1 along the ray's dominant axis, so with jf = j - N/2, md the midpoint's dominant
// coordinate and s the step's sign, sample jf's dominant lower tap is the integer
// floor(md) + s jf with weight fract(md) (exact, one row per sample); its other coordinate is
// c = fma(jf, v_other, m_other) in FP32, lower tap floor(c), weight fract(c). The forward and
// the adjoint both use exactly this, so they are each other's transpose.
//   per call     for each image b of the batch, one block of 2 * (nx+4) * (ny+4) elements:
//                pad  [ny+4][nx+4]               image with a 2-texel zero border
//                padT [nx+4][ny+4]               the same, transposed
//                (backprojection) sinogram with a 2-cell zero border: [B][A][D+4]
// The forward kernel reads the copy whose rows run along each ray's dominant axis, so the 64
// rays of a wavefront (adjacent detector cells of one angle) gather from a few cache lines per
// step whichever way the fan points, and the zero border replaces per-tap range checks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "bp_tiles.hpp"
#include "cmi_internal.hpp"
#include "geometry2d.hpp"

using namespace cmi;
using cmi::fan2d::AngleGeom;

struct cmi_plan2d {
    int device = 0;
    int vol_prec = CMI_F32, det_prec = CMI_F32;
    fan2d::Scalars g{};
    int64_t A = 0;
    double pitch = 0.0, cor = 0.0;
    uint32_t block_x = 64, block_y = 4;
    uint64_t total_samples = 0;
    uint64_t table_bytes = 0;
    AngleGeom *d_geo = nullptr;
    float4 *d_meta = nullptr;
    float4 *d_bpv = nullptr; // views (plans from vectors); d_meta is null then
    float *d_ageo = nullptr;
    float4 *d_ev = nullptr;
    float4 *d_ns = nullptr;
    bool quarter = false; // backprojection by quarter turns applies (k_bp2d_qt)
    // bounds on the backprojection's per-projection weight over the image (w / du, w^2 / du
    // for circular plans, with the view's 1 / |u| for plans from views); infinite when a
    // source lies inside the image's bounding box
    double bt_bound[2] = {INFINITY, INFINITY};
};

namespace {

constexpr int kPad = 2;        // zero border of the staged image copies, in texels
constexpr int kMaxBlock = 256; // threads per workgroup every kernel here is compiled for
static_assert(kPad == kPadDet, "sinograms are staged with the backprojector's border");

// Per-angle constants of the backprojector: (-sin, cos, DSD, DSO) (backward_2d.hpp:58-62).
float4 bp2d_meta(double angle, double dso, double dsd) {
    return make_float4((float)-std::sin(angle), (float)std::cos(angle), (float)dsd, (float)dso);
}

// ---------------------------------------------------------------- ray table ----------------

__global__ __launch_bounds__(kMaxBlock) void k_ray_table(const AngleGeom *__restrict__ geo, int A,
                                                   int D, int64_t nx, int64_t ny, double sx,
                                                   double sy, float4 *__restrict__ ev,
                                                   float4 *__restrict__ ns,
                                                   unsigned long long *__restrict__ total) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    unsigned long long n = 0;
    if (k < D && a < A) {
        const fan2d::Ray r = fan2d::setup_ray(geo[a], nx, ny, sx, sy, k);
        const size_t i = (size_t)a * D + k;
        const int64_t jm = r.n / 2;
        // texel coordinate -> centred padded coordinate
        const double cx = (double)(kPad - 0.5) - (double)((nx + 2 * kPad) / 2);
        const double cy = (double)(kPad - 0.5) - (double)((ny + 2 * kPad) / 2);
        const double mx = r.e[0] + (double)jm * r.v[0] + cx;
        const double my = r.e[1] + (double)jm * r.v[1] + cy;
        ev[i] = r.n ? make_float4((float)mx, (float)my, (float)r.v[0], (float)r.v[1])
                    : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        ns[i] = make_float4((float)r.scale, __int_as_float((int)r.n), (float)(-jm), 0.0f);
        n = (unsigned long long)r.n;
    }
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(total, n);
}

// ---------------------------------------------------------------- staging ------------------

// pad[py][px] = img[py-2][px-2] (0 outside) and padT[px][py] = pad[py][px], tile by tile.
template <class S>
__global__ __launch_bounds__(kMaxBlock) void k_stage2d(const typename S::raw *__restrict__ img,
                                                 typename S::raw *__restrict__ scratch, int nx,
                                                 int ny, int Wp, int Hp) {
    using raw = typename S::raw;
    __shared__ raw tile[32][33];
    const size_t b = blockIdx.z;
    img += b * (size_t)nx * ny;
    raw *pad = scratch + b * 2 * (size_t)Wp * Hp;
    raw *padT = pad + (size_t)Wp * Hp;
    const int px0 = blockIdx.x * 32, py0 = blockIdx.y * 32;
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int py = py0 + r, px = px0 + threadIdx.x;
        const int x = px - kPad, y = py - kPad;
        raw v = (raw)0.0f;
        if (x >= 0 && x < nx && y >= 0 && y < ny) v = img[(size_t)y * nx + x];
        if (px < Wp && py < Hp) pad[(size_t)py * Wp + px] = v;
        tile[r][threadIdx.x] = v;
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int px = px0 + r, py = py0 + threadIdx.x;
        if (px < Wp && py < Hp) padT[(size_t)px * Hp + py] = tile[threadIdx.x][r];
    }
}

// spad[b][a][c] = sino[b][a][c-2] for c in [2, D+2), 0 in the two border cells each side.
template <class S>
__global__ __launch_bounds__(kMaxBlock) void k_stage_sino2d(const typename S::raw *__restrict__ sino,
                                                      typename S::raw *__restrict__ spad, int D,
                                                      int A) {
    using raw = typename S::raw;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int Dp = D + 2 * kPad;
    if (c >= Dp) return;
    const size_t row = (size_t)blockIdx.z * A + blockIdx.y;
    const int k = c - kPad;
    spad[row * Dp + c] = (k >= 0 && k < D) ? sino[row * D + k] : (raw)0.0f;
}

// ---------------------------------------------------------------- forward ------------------

// A lane's sums over its samples of one band. A sample with lower tap (row t, column i) and
// fraction fa along the row adds p0 += v(t, i), x0 += fa (v(t, i + 1) - v(t, i)), and p1, x1
// likewise on row t + 1. With u0 = p0 + x0 and u1 = p1 + x1 (the sums of the row-wise linear
// interpolants on the sample rows and on the rows below), the band's contribution is
// u0 + fb (u1 - u0), the sum of the samples' bilinear interpolants, because every sample of a
// ray has the same dominant weight fb.
struct BandSums {
    float p0 = 0.0f, x0 = 0.0f, p1 = 0.0f, x1 = 0.0f;
    // a0, a1: v(t, i), v(t, i + 1); b0, b1: v(t + 1, i), v(t + 1, i + 1)
    __device__ __forceinline__ void add(float a0, float a1, float b0, float b1, float fa) {
        p0 += a0;
        x0 = fmaf(fa, a1 - a0, x0);
        p1 += b0;
        x1 = fmaf(fa, b1 - b0, x1);
    }
    __device__ __forceinline__ float total(float fb) const {
        const float u0 = p0 + x0, u1 = p1 + x1;
        return fmaf(fb, u1 - u0, u0);
    }
};

// Row-synchronous march through LDS bands.
//
// A ray has one sample per dominant row: sample jf sits on row floor(md) + s jf. A workgroup
// holds adjacent detector cells of one angle and walks the dominant rows tau in increasing
// order: on row tau every lane whose dominant axis is the pass's takes its sample
// jf = s (tau - floor(md)), if jf lies in its range [q.z, q.z + N - 1], so all lanes read the
// same two rows. The rows come in bands of kBand, aligned to multiples of kBand: the workgroup
// stages rows [t0, t0 + kBand] of the copy whose rows run along the dominant axis, over the
// columns its samples in the band touch, into LDS (rows of kPitch floats, filled with
// 4-element row loads) and interpolates from there. The last two cells of every row are kept
// at zero and are what lanes without a sample on a row read. A band wider than kPitch - 4
// columns is sampled from global memory instead, with the same arithmetic. Bands in which
// every lane of the workgroup has a sample on every row skip the per-sample range test.
//
// Staged cells outside the copy (band corners beyond its edge) hold whatever the range-checked
// loads return; no sample reads them, because every tap of a sample lies inside the copy's zero
// border.
//
// Summation: a lane adds its samples of a band row by row into its BandSums, and the band's
// total into the ray's sum. Bands are absolute row numbers and both paths add the same values
// in the same order, so a ray's sum does not depend on which rays share its workgroup (and so
// not on the launch shape).
// rows per band: a band and the row below it make 32 staged rows, so that the common band
// widths (25 to 32 quads, four groups of threads) stage eight rows per thread, one group of
// loads in flight
constexpr int kBand = 31;
constexpr int kBandRows = kBand + 1;   // staged rows: the band and the row below it
// LDS row pitch in floats: a multiple of 4 (16-byte row stores) and below 255, so that one
// address reaches a tap pair and the pair below it with two ds_read2_b32 immediate offsets.
// 152 (19.5 KB per band) keeps eight workgroups per CU: 128 rays at up to about 0.9 texels per
// cell plus a band's slope and the tap margins fit; wider bands take the global path. At
// 1024^2 (0.67 texels per cell) a pitch of 152 gave 0.522 ms against 0.556 with 192 (six
// workgroups per CU), 176: 0.534, 144: 0.556, 136: 0.590.
constexpr int kPitch = 152;
constexpr int kFwdRays = 128;          // rays (threads) per workgroup of the forward
constexpr int kMaxWaves = kFwdRays / 64;

__device__ __forceinline__ int wave_min(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// Workgroup-wide minima of N values (callers negate to take maxima); every thread of the
// workgroup calls it. `red` holds N ints per wavefront.
template <int N>
__device__ __forceinline__ void block_min(int (&v)[N], int *red, int nwaves) {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = wave_min(v[i]);
    const int w = threadIdx.x >> 6;
    __syncthreads(); // previous readers of red are done
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int i = 0; i < N; ++i) red[w * N + i] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N; ++i) {
        int r = red[i];
        for (int k = 1; k < nwaves; ++k) r = min(r, red[k * N + i]);
        v[i] = __builtin_amdgcn_readfirstlane(r);
    }
}

__device__ __forceinline__ int floor_div(int a, int b) {
    return a >= 0 ? a / b : -((-a + b - 1) / b);
}

// One pass (YDOM: rays whose dominant axis is y, row-major copy; else x, transposed copy).
// All threads of the workgroup call it together; returns the lane's sum, 0 if not in the pass.
//   rstride      elements per row of the copy (rows run along the dominant axis)
//   cofs         element offset of the copy's centre (centred coordinate 0, 0)
// Dominant row range and column of a ray at row t: sample jf = s (t - floor(md)), column
// coordinate fma(jf, vc, mc). Used for the workgroup's extreme rays.
struct RayRow {
    float mc, vc;
    int r0, s;
};

template <bool YDOM>
__device__ __forceinline__ RayRow ray_row(float4 e) {
    const float md = YDOM ? e.y : e.x;
    return {YDOM ? e.x : e.y, YDOM ? e.z : e.w, (int)floorf(md), (YDOM ? e.w : e.z) > 0.0f ? 1 : -1};
}

__device__ __forceinline__ int ray_col(const RayRow &r, int t) {
    return floor_i32(fmaf((float)(r.s * (t - r.r0)), r.vc, r.mc));
}

template <class SI, bool YDOM>
__device__ __forceinline__ float march_band(rsrc_t R, float4 e, float4 q, bool mine, int rstride,
                                            int cofs, float *band, int *red, float4 *ext,
                                            int *flag, int nthreads) {
    const int n = __float_as_int(q.y);
    const bool act = mine && n > 0;
    const int nwaves = nthreads >> 6;
    const float md = YDOM ? e.y : e.x;   // dominant coordinate of the midpoint
    const float sd = YDOM ? e.w : e.z;   // dominant step, +-1
    const float mc = YDOM ? e.x : e.y;   // other coordinate of the midpoint
    const float vc = YDOM ? e.z : e.w;   // other step
    const int r0 = (int)floorf(md);
    const float fb = md - (float)r0;     // dominant weight of every sample (exact)
    const int s = sd > 0.0f ? 1 : -1;
    const float sf = (float)s;
    const int jlo = (int)q.z, jhi = jlo + n - 1;
    const int ta = s > 0 ? r0 + jlo : r0 - jhi; // first row of this lane
    const int tz = ta + n - 1;                  // last row
    const int k = (int)threadIdx.x;
    if (k == 0) flag[0] = flag[1] = 0; // ordered before every use by block_min's barriers
    // pass-wide: rows [tr0, tr1], the latest first row and earliest last row of any lane, the
    // first and last active lane, and whether every lane of the workgroup is in the pass
    int v[7] = {act ? ta : INT_MAX, act ? -tz : INT_MAX, act ? -ta : INT_MAX, act ? tz : INT_MAX,
                act ? k : INT_MAX,  act ? -k : INT_MAX,  act ? 1 : 0};
    block_min<7>(v, red, nwaves);
    if (v[0] == INT_MAX) return 0.0f; // no lane in this pass
    const int tr0 = v[0], tr1 = -v[1], ta_max = -v[2], tz_min = v[3];
    const bool all_in = v[6] == 1; // the minimum is 1 only if every lane is active
    // Rays of one angle come from one source and do not cross inside the image, so on any row
    // the columns of the workgroup's samples lie between those of its first and last active
    // rays (extended beyond their ends where needed). Every band checks this and falls back to
    // global memory if a lane's samples leave the staged columns.
    if (k == v[4]) ext[0] = e;
    if (k == -v[5]) ext[1] = e;
    __syncthreads();
    const RayRow e0 = ray_row<YDOM>(ext[0]), e1 = ray_row<YDOM>(ext[1]);
    const unsigned span = (unsigned)(n - 1);
    // a lane has a sample on row t iff (unsigned)t - ta_m <= span_m (never for inactive lanes:
    // rows are far from the offset 2^31)
    const unsigned ta_m = act ? (unsigned)ta : 0x80000000u, span_m = act ? span : 0u;
    const int soff1 = rstride * (int)sizeof(typename SI::raw);
    float acc = 0.0f;
    int bi = 0;
    // A band's staging plan: the columns it needs (bounded by the extreme rays, one column of
    // margin each side for rounding; the bounds are the same for every lane and live in scalar
    // registers) and this thread's share of the 16-byte row loads: rows [t0, t0 + kBandRows) x
    // columns [c_lo, c_lo + 4 Q) as FP32, with 4 Q <= kPitch - 4 (the zero cells stay
    // untouched). G' groups of Q threads, G' the largest power of two with G' Q <= nthreads:
    // thread g Q + q (g < G') takes quad q of the K = kBandRows / G' consecutive rows g K ..
    // g K + K - 1. A row then lies at a wave-uniform byte offset from the thread's first row in
    // the copy (one add; the offset stays in the load's range-checked vector offset, so rows
    // outside the copy read zeros) and at a constant one in LDS (the store's immediate offset),
    // and no row needs a bound. Quotients come from a float reciprocal: exact, as every
    // numerator is below 2^8 and every true quotient is at least 1/(2 Q) > 2^-8 away from the
    // next integer.
    struct Stage {
        int t0, b_lo, b_hi, c_lo, c_hi;
        bool lds, mine;
        int K, gb, lb;
    };
    constexpr int esz = (int)sizeof(typename SI::raw);
    const int rowb = __builtin_amdgcn_readfirstlane(rstride * esz); // bytes per row of the copy
    auto plan_band = [&](int t0) {
        Stage P{};
        P.t0 = t0;
        P.b_lo = max(tr0, t0);
        P.b_hi = min(tr1, t0 + kBand - 1);
        const int ea = ray_col(e0, P.b_lo), eb = ray_col(e0, P.b_hi);
        const int ec = ray_col(e1, P.b_lo), ed = ray_col(e1, P.b_hi);
        P.c_lo = __builtin_amdgcn_readfirstlane(min(min(ea, eb), min(ec, ed)) - 1);
        P.c_hi = __builtin_amdgcn_readfirstlane(max(max(ea, eb), max(ec, ed)) + 1);
        const int W = P.c_hi + 2 - P.c_lo; // columns [c_lo, c_hi + 1] hold the taps
        // staged when the band fits a row and at least two groups of Q threads exist (so that
        // K <= 16); a 64-thread workgroup stages bands of up to 128 columns
        P.lds = W <= kPitch - 4 && 2 * ((W + 3) >> 2) <= nthreads;
        if (P.lds) {
            const int Q = (W + 3) >> 2; // 1 .. 47, at most nthreads / 2 (a multiple of 32)
            const float rq = __builtin_amdgcn_rcpf((float)Q);
            const int G = __builtin_amdgcn_readfirstlane((int)(((float)nthreads + 0.5f) * rq));
            // G' = 2^lg <= min(G, kBandRows), K = kBandRows >> lg <= 16 (G >= 2 as 2 Q <= nthreads)
            const int lg = min(31 - __builtin_clz((unsigned)G), 5);
            P.K = kBandRows >> lg;
            const int g = (int)(((float)threadIdx.x + 0.5f) * rq);
            P.mine = g < (1 << lg);
            const int q4 = ((int)threadIdx.x - g * Q) << 2;
            // byte offsets of this thread's quad in its first row: copy and LDS
            P.gb = ((t0 + g * P.K) * rstride + P.c_lo + cofs + q4) * esz;
            P.lb = (g * P.K * kPitch + q4) * 4;
        }
        return P;
    };
    static_assert(kBandRows == 32, "row shares are K = 32 / G' rows");
    // rows k0 .. k0 + 7 of the thread's share (those below K), and their store into LDS
    auto load_rows = [&](const Stage &P, int k0, Quad (&v)[8]) {
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k0 + u < P.K) v[u] = buf_quad_at<SI>(R, P.gb + (k0 + u) * rowb);
    };
    auto store_rows = [&](const Stage &P, int k0, const Quad (&v)[8]) {
        char *l = reinterpret_cast<char *>(band) + P.lb;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k0 + u < P.K)
                *reinterpret_cast<float4 *>(l + (k0 + u) * kPitch * 4) =
                    make_float4(v[u].v[0], v[u].v[1], v[u].v[2], v[u].v[3]);
    };
    // The first eight rows of a thread's share of band i + 1 are loaded while band i is sampled
    // (registers `pre`), so their latency hides behind the sampling; K <= 16 rows in all.
    Stage P = plan_band(floor_div(tr0, kBand) * kBand);
    Quad pre[8];
    if (P.lds && P.mine) load_rows(P, 0, pre);
    for (; P.t0 <= tr1; ++bi) {
        // whether every lane has a sample on every row [b_lo, b_hi]
        const bool full = all_in && ta_max <= P.b_lo && tz_min >= P.b_hi;
        BandSums sums;
        float jf = (float)(s * (P.b_lo - r0));
        bool use_lds = false;
        if (P.lds) {
            if (P.mine) {
                store_rows(P, 0, pre);
                if (P.K > 8) {
                    Quad v[8];
                    load_rows(P, 8, v);
                    store_rows(P, 8, v);
                }
            }
            // every lane checks that its samples in the band stay in the staged columns (the
            // column is monotonic in the row, so its first and last row decide)
            const int f = max(ta, P.b_lo), l = min(tz, P.b_hi);
            bool out = false;
            if (act && f <= l) {
                const int c1 = floor_i32(fmaf((float)(s * (f - r0)), vc, mc));
                const int c2 = floor_i32(fmaf((float)(s * (l - r0)), vc, mc));
                out = min(c1, c2) < P.c_lo || max(c1, c2) > P.c_hi;
            }
            if (__builtin_amdgcn_ballot_w64(out) != 0 && (threadIdx.x & 63) == 0)
                flag[bi & 1] = 1;
            __syncthreads();
            use_lds = flag[bi & 1] == 0;
            // the other slot was last read in the previous band, and is next written in the
            // following one, after this band's closing barrier
            if (threadIdx.x == 0) flag[(bi + 1) & 1] = 0;
        }
        const Stage N = plan_band(P.t0 + kBand);
        if (N.t0 <= tr1 && N.lds && N.mine) load_rows(N, 0, pre);
        if (use_lds) {
            // float offset of the band's first row, column 0, held opaque in a scalar
            // register so that the compiler does not split a constant off every address
            int rofs = __builtin_amdgcn_readfirstlane((P.b_lo - P.t0) * kPitch - P.c_lo);
            asm("" : "+s"(rofs));
            const float *rowa = band + rofs;
            int t = P.b_lo;
            if (full) { // every lane has a sample on every row of the band
                for (; t + 3 <= P.b_hi; t += 4, jf += 4.0f * sf, rowa += 4 * kPitch) {
                    float c[4], x[4][4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        c[u] = fmaf(u == 0 ? jf : jf + (float)u * sf, vc, mc);
                        const float *p = rowa + u * kPitch + floor_i32(c[u]);
                        x[u][0] = p[0];
                        x[u][1] = p[1];
                        x[u][2] = p[kPitch];
                        x[u][3] = p[kPitch + 1];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        sums.add(x[u][0], x[u][1], x[u][2], x[u][3],
                                 __builtin_amdgcn_fractf(c[u]));
                }
                for (; t <= P.b_hi; ++t, jf += sf, rowa += kPitch) {
                    const float c = fmaf(jf, vc, mc);
                    const float *p = rowa + floor_i32(c);
                    sums.add(p[0], p[1], p[kPitch], p[kPitch + 1], __builtin_amdgcn_fractf(c));
                }
            } else {
                const int zc = P.c_lo + kPitch - 2; // the zero cells
                for (; t + 3 <= P.b_hi; t += 4, jf += 4.0f * sf, rowa += 4 * kPitch) {
                    float c[4], x[4][4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        c[u] = fmaf(u == 0 ? jf : jf + (float)u * sf, vc, mc);
                        const int i = floor_i32(c[u]);
                        const bool in = (unsigned)(t + u) - ta_m <= span_m;
                        const float *p = rowa + u * kPitch + (in ? i : zc);
                        x[u][0] = p[0];
                        x[u][1] = p[1];
                        x[u][2] = p[kPitch];
                        x[u][3] = p[kPitch + 1];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        sums.add(x[u][0], x[u][1], x[u][2], x[u][3],
                                 __builtin_amdgcn_fractf(c[u]));
                }
                for (; t <= P.b_hi; ++t, jf += sf, rowa += kPitch) {
                    const float c = fmaf(jf, vc, mc);
                    const int i = floor_i32(c);
                    const bool in = (unsigned)t - ta_m <= span_m;
                    const float *p = rowa + (in ? i : zc);
                    sums.add(p[0], p[1], p[kPitch], p[kPitch + 1], __builtin_amdgcn_fractf(c));
                }
            }
        }
        if (P.lds) __syncthreads(); // the band and its flag are read before the next is staged
        if (!use_lds) {
            // lanes without a sample on the row sample their nearest own sample and add zeros
            const float jmin = (float)jlo, jmax = (float)(act ? jhi : jlo);
            for (int t = P.b_lo; t <= P.b_hi; ++t, jf += sf) {
                const float j = __builtin_amdgcn_fmed3f(jf, jmin, jmax);
                const float c = fmaf(j, vc, mc);
                const int o = __mul24(r0 + s * (int)j, rstride) + floor_i32(c) + cofs;
                Pair a = buf_pair<SI>(R, o, 0);     // row, columns i and i + 1
                Pair b = buf_pair<SI>(R, o, soff1); // the row below
                if ((unsigned)t - ta_m > span_m) a = b = Pair{0.0f, 0.0f};
                sums.add(a.lo, a.hi, b.lo, b.hi, __builtin_amdgcn_fractf(c));
            }
        }
        acc += sums.total(fb);
        P = N;
    }
    return acc;
}

// One workgroup = nthreads <= kFwdRays adjacent detector cells of one angle, one thread per
// ray; LDS holds one band of kBandRows x kPitch floats (19.5 KB).
// sino[b][a][k] = scale * sum_{j<N} bilinear(img, e + j v)   (forward_2d.hpp:63-78)
template <class SI, class SS>
__global__ __launch_bounds__(kFwdRays) void k_fwd2d(const typename SI::raw *__restrict__ scratch,
                                              int Wp, int Hp, const float4 *__restrict__ ev,
                                              const float4 *__restrict__ ns,
                                              typename SS::raw *__restrict__ sino, int D, int A) {
    // one LDS object with the band first, so that cell addresses need no base offset
    __shared__ struct {
        float band[kBandRows * kPitch];
        float4 ext[2];
        int red[7 * kMaxWaves];
        int flag[2];
    } sh;
    float *band = sh.band;
    float4 *ext = sh.ext;
    int *red = sh.red, *flag = sh.flag;
    const int nthreads = blockDim.x;
    // The launch's linear workgroup index taken with angles fastest: consecutive workgroups
    // (dealt to the 8 XCDs in turn) hold the same cells at successive angles, so every XCD
    // gets every eighth angle of one cell block at a time, and workgroups resident together
    // march rays of similar length (0.593 -> 0.556 ms at 1024^2; cell blocks fastest, each
    // XCD a contiguous run: 0.635 ms).
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned hw = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const int a = (int)(hw % gy);
    const int k = (int)((hw / gy) % gx) * nthreads + threadIdx.x;
    const size_t b = hw / (gx * gy);
    const bool valid = k < D;
    const size_t r = (size_t)a * D + k;
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 e = valid ? ev[r] : zero;
    const float4 q = valid ? ns[r] : zero; // N = 0
    const int one = Wp * Hp;
    const rsrc_t R = make_rsrc(scratch + b * 2 * (size_t)one, bytes_of<SI>(2 * (uint64_t)one));
    const int c0x = Wp / 2, c0y = Hp / 2;
    const bool ydom = fabsf(e.w) == 1.0f; // y dominant: row-major copy; else the transposed one
    // the zero cells; ordered before their first read by the barriers at the start of a pass
    for (int i = threadIdx.x; i < 2 * kBandRows; i += nthreads)
        band[(i >> 1) * kPitch + kPitch - 2 + (i & 1)] = 0.0f;
    const float acc =
        march_band<SI, true>(R, e, q, valid && ydom, Wp, c0y * Wp + c0x, band, red, ext, flag,
                             nthreads) +
        march_band<SI, false>(R, e, q, valid && !ydom, Hp, one + c0x * Hp + c0y, band, red, ext,
                              flag, nthreads);
    if (valid) sino[b * (size_t)A * D + r] = from_f32<SS>(acc * q.x);
}

// ---------------------------------------------------------------- backprojection ---------

// One thread per pixel, all angles in one pass (backward_2d.hpp:31-99):
// vol[b][y][x] += (1/du) sum_a w_a^WPOW lerp(sino[b][a], w_a eta_a + D/2),  w_a = DSD/(xi_a + DSO).
// WPOW = 1 is the reference's weighting; WPOW = 2 the fan-beam FBP distance weight.
template <class SV, class SS, int WPOW>
__global__ __launch_bounds__(kMaxBlock) void k_bp2d(typename SV::raw *__restrict__ vol,
                                                    const typename SS::raw *__restrict__ spad,
                                                    const float4 *__restrict__ meta, int nx,
                                                    int ny, int D, int A, float sx, float sy,
                                                    float ox, float oy, float inv_du,
                                                    float half_d) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= nx || iy >= ny) return;
    const size_t b = blockIdx.z;
    const uint32_t sino_elems = (uint32_t)A * (D + 2 * kPadDet);
    const rsrc_t R = make_rsrc(spad + b * sino_elems, bytes_of<SS>(sino_elems));
    const float x = ((float)ix - (float)nx / 2.0f) * sx + ox + 0.5f;
    const float y = ((float)iy - (float)ny / 2.0f) * sy + oy + 0.5f;
    const float acc = bp2d_pixel<SS, WPOW>(R, meta, A, D, x, y, x * inv_du, y * inv_du, half_d);
    const size_t i = b * (size_t)nx * ny + (size_t)iy * nx + ix;
    vol[i] = from_f32<SV>(to_f32<SV>(vol[i]) + acc * inv_du);
}

// The same for a plan built from ASTRA-like views (geometry2d.hpp bp_view): per view,
// depth = n . x + n0, w = L / depth, cell coordinate w (e . x + e0) + c0, weight w^WPOW / |u|.
// vol[b][y][x] += sum_a w_a^WPOW / |u_a| lerp(sino[b][a], cell_a).
template <class SV, class SS, int WPOW>
__global__ __launch_bounds__(kMaxBlock) void k_bp2d_vec(typename SV::raw *__restrict__ vol,
                                                        const typename SS::raw *__restrict__ spad,
                                                        const float4 *__restrict__ bpv, int nx,
                                                        int ny, int D, int A, float sx, float sy,
                                                        float ox, float oy) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= nx || iy >= ny) return;
    const size_t b = blockIdx.z;
    const int Dp = D + 2 * kPadDet;
    const uint32_t sino_elems = (uint32_t)A * Dp;
    const rsrc_t R = make_rsrc(spad + b * sino_elems, bytes_of<SS>(sino_elems));
    const float x = ((float)ix - (float)nx / 2.0f) * sx + ox + 0.5f;
    const float y = ((float)iy - (float)ny / 2.0f) * sy + oy + 0.5f;
    const float hi = (float)D;
    float acc = 0.0f;
    int row = kPadDet;
#pragma unroll 4
    for (int a = 0; a < A; ++a, row += Dp) {
        const float4 m0 = bpv[3 * a], m1 = bpv[3 * a + 1]; // wave-uniform: scalar loads
        const float idu = bpv[3 * a + 2].x;
        const float w = m0.w * __builtin_amdgcn_rcpf(fmaf(m0.x, x, fmaf(m0.y, y, m0.z)));
        const float eta = fmaf(m1.x, x, fmaf(m1.y, y, m1.z));
        const float u = __builtin_amdgcn_fmed3f(fmaf(w, eta, m1.w), -2.0f, hi); // c0 - 0.5
        const float al = __builtin_amdgcn_fractf(u);
        const Pair p = buf_pair<SS>(R, row + floor_i32(u), 0);
        acc = fmaf(fmaf(al, p.hi - p.lo, p.lo), (WPOW == 2 ? w * w : w) * idu, acc);
    }
    const size_t i = b * (size_t)nx * ny + (size_t)iy * nx + ix;
    vol[i] = from_f32<SV>(to_f32<SV>(vol[i]) + acc);
}

// Backprojection by quarter turns, for square images (n even, unit pixels, no offset) whose
// angles come in quarter-turn quadruples (angle a + A/4 = angle a + pi/2, equal DSO and DSD):
// the pixel at world (y, -x) has at angle theta + pi/2 the depth xi + DSO and detector
// coordinate that the pixel (x, y) has at angle theta (with meta = (-sin, cos), xi = -sin x -
// cos y and eta = cos x - sin y turn into each other). So one thread evaluates the geometry of
// its pixel P = (x, y) once per angle a and adds its weighted
// sample of row a to P, of row a + A/4 to (y, -x), of row a + A/2 to (-x, -y) and of row
// a + 3A/4 to (-y, x): four pixels per geometry evaluation. In index terms, with h = n/2 and
// P = (ix, iy) in the quadrant [0, h)^2: (iy, n-1-ix), (n-1-ix, n-1-iy), (n-1-iy, ix).
// A workgroup owns a 16 x 16 tile of the quadrant. Per chunk of 16 angles, every wavefront
// projects the tile's corners (lanes 4j .. 4j+3 for angle j of the chunk), the four rows of each
// angle share that footprint (the four rotated tiles project alike), and the workgroup stages
// the four rows' footprint segments (64 cells) into LDS with 16-byte loads while the previous
// chunk is sampled from the other buffer. Samples read their tap pair from LDS at immediate
// offsets, one address per angle for all four rows. The plan checks the symmetry and that every
// footprint fits a segment; the per-lane clamp keeps every LDS read inside its segment.
constexpr int kQT = 16;            // tile edge, pixels
constexpr int kQThreads = kQT * kQT;
constexpr int kQChunk = 16;        // angles per staged chunk
constexpr int kQSeg = 64;          // staged cells per row
static_assert(kQThreads == 256 && kQChunk * 4 * (kQSeg / 4) == 4 * kQThreads,
              "each thread stages four quads per chunk");

template <class SV, class SS, int WPOW>
__global__ __launch_bounds__(kQThreads) void k_bp2d_qt(typename SV::raw *__restrict__ vol,
                                                       const typename SS::raw *__restrict__ spad,
                                                       const float4 *__restrict__ meta, int n,
                                                       int D, int A, float inv_du, float half_d) {
    __shared__ float seg[2][kQChunk * 4 * kQSeg]; // two chunks of [angle][row k][cell]
    const int h = n / 2, Q = A / 4, Dp = D + 2 * kPadDet;
    const size_t b = blockIdx.z;
    const uint32_t sino_elems = (uint32_t)A * Dp;
    const rsrc_t R = make_rsrc(spad + b * sino_elems, bytes_of<SS>(sino_elems));
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int bx = blockIdx.x * kQT, by = blockIdx.y * kQT;
    const int ix = bx + (t & (kQT - 1)), iy = by + t / kQT;
    const float x = (float)ix - (float)h + 0.5f, y = (float)iy - (float)h + 0.5f;
    const float px = x * inv_du, py = y * inv_du;
    const float hi = (float)D + 2.0f;     // padded coordinates: the row's zero border included
    const float c2 = half_d - 0.5f + 2.0f; // D/2, the lerp's texel-centre shift and the border
    // the tile corner (pixel centre) this lane projects for angle lane >> 2 of a chunk
    const int cxi = (lane & 1) ? min(bx + kQT - 1, h - 1) : bx;
    const int cyi = (lane & 2) ? min(by + kQT - 1, h - 1) : by;
    const float qx = (float)cxi - (float)h + 0.5f, qy = (float)cyi - (float)h + 0.5f;
    const float qpx = qx * inv_du, qpy = qy * inv_du;
    // staging share: quads q = t & 15 of row k = (t >> 4) & 3 of angles j = wv + 4 m, m < 4
    const int sq = t & 15, sk = (t >> 4) & 3;
    constexpr int esz = (int)sizeof(typename SS::raw);

    // Per chunk c, lanes 4 j .. 4 j + 3 of every wavefront hold angle j's segment: s0 (its first
    // padded cell), c2 - s0 (the sample coordinate's constant in segment coordinates) and the
    // largest lower-tap coordinate min(D + 2 - s0, kQSeg - 2) (the row's end, or the segment's).
    struct Seg {
        int s0;
        float c2s, hi;
    };
    auto footprint = [&](int c) -> Seg {
        const int a = min(c + (lane >> 2), A - 1);
        const float4 m = meta[a];
        const float depth = fmaf(m.x, qx, fmaf(-m.y, qy, m.w));
        const float eta = fmaf(m.y, qpx, m.x * qpy);
        const float w = m.z * __builtin_amdgcn_rcpf(depth);
        float v = __builtin_amdgcn_fmed3f(fmaf(w, eta, c2), 0.0f, hi);
        v = fminf(v, __shfl_xor(v, 1, 64));
        v = fminf(v, __shfl_xor(v, 2, 64));
        // at most Dp - kQSeg, so that a segment never runs past its row (the plan requires
        // Dp >= kQSeg): no staged quad reads another row or straddles the array's end
        const int s0 = min(max(floor_i32(v) - 1, 0), Dp - kQSeg);
        return {s0, c2 - (float)s0, fminf(hi - (float)s0, (float)(kQSeg - 2))};
    };
    auto load = [&](int c, int s0v, Quad (&qd)[4]) {
#pragma unroll
        for (int mm = 0; mm < 4; ++mm) {
            const int j = wv + 4 * mm;
            const int s0 = __builtin_amdgcn_readlane(s0v, 4 * j);
            // row of angle c + j (past the last angle: the last one's, loaded and unused),
            // turned by k quarters
            int r = min(c + j, A - 1) + sk * Q;
            r = r >= A ? r - A : r;
            qd[mm] = buf_quad_at<SS>(R, (r * Dp + s0 + 4 * sq) * esz);
        }
    };
    auto store = [&](int buf, const Quad (&qd)[4]) {
#pragma unroll
        for (int mm = 0; mm < 4; ++mm) {
            const int j = wv + 4 * mm;
            *reinterpret_cast<float4 *>(&seg[buf][(j * 4 + sk) * kQSeg + 4 * sq]) =
                make_float4(qd[mm].v[0], qd[mm].v[1], qd[mm].v[2], qd[mm].v[3]);
        }
    };
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // angle c + j of a chunk whose segments are in `sg`; TAIL: the angle may lie past the last
    // one (then it adds nothing)
    auto sample = [&](int c, int j, const Seg &sg, int bufofs, auto tail) {
        constexpr bool TAIL = decltype(tail)::value;
        if (TAIL && c + j >= A) return;
        const float4 m = meta[c + j]; // wave-uniform: scalar loads
        const float depth = fmaf(m.x, x, fmaf(-m.y, y, m.w)); // xi + DSO
        const float eta = fmaf(m.y, px, m.x * py);
        const float w = m.z * __builtin_amdgcn_rcpf(depth);
        const float c2s = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sg.c2s), 4 * j));
        const float hj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sg.hi), 4 * j));
        const float l = __builtin_amdgcn_fmed3f(fmaf(w, eta, c2s), 0.0f, hj);
        const float al = __builtin_amdgcn_fractf(l);
        const float wt = WPOW == 2 ? w * w : w;
        // byte offset of the angle's four rows, opaque in a scalar register so that the rows
        // stay immediate offsets of one address
        int jb = bufofs + j * 4 * kQSeg * 4;
        asm("" : "+s"(jb));
        const float *p = reinterpret_cast<const float *>(
            reinterpret_cast<const char *>(&seg[0][0]) + jb) + floor_i32(l);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float lo = p[k * kQSeg], up = p[k * kQSeg + 1];
            acc[k] = fmaf(fmaf(al, up - lo, lo), wt, acc[k]);
        }
    };
    using False = std::false_type;
    using True = std::true_type;
    Quad qd[4];
    Seg sg = footprint(0);
    load(0, sg.s0, qd);
    store(0, qd);
    __syncthreads();
    int cur = 0;
    for (int c = 0; c < A; c += kQChunk, cur ^= 1) {
        // next chunk's loads in flight while this one is sampled
        const int cn = c + kQChunk;
        Seg sn{};
        if (cn < A) {
            sn = footprint(cn);
            load(cn, sn.s0, qd);
        }
        const int bufofs = cur * kQChunk * 4 * kQSeg * 4;
        if (cn <= A) {
#pragma unroll
            for (int j = 0; j < kQChunk; ++j) sample(c, j, sg, bufofs, False{});
        } else {
#pragma unroll
            for (int j = 0; j < kQChunk; ++j) sample(c, j, sg, bufofs, True{});
        }
        if (cn < A) store(cur ^ 1, qd);
        sg = sn;
        __syncthreads();
    }
    if (ix >= h || iy >= h) return;
    const int m1 = n - 1;
    const int px4[4] = {ix, iy, m1 - ix, m1 - iy};
    const int py4[4] = {iy, m1 - ix, m1 - iy, ix};
    typename SV::raw *vb = vol + b * (size_t)n * n;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const size_t i = (size_t)py4[k] * n + px4[k];
        vb[i] = from_f32<SV>(to_f32<SV>(vb[i]) + acc[k] * inv_du);
    }
}

// ---------------------------------------------------------------- adjoint ------------------

// Largest |y q.x| over the rays of each sinogram of a batch (q.x the ray's scale), in two
// steps: per sinogram b, kAdjParts workgroups each write the maximum of their grid-stride
// share to parts[b][*]; the adjoint reduces the kAdjParts values. Each image of a batch gets
// its own fixed-point unit, so that its result does not depend on the other images.
constexpr int kAdjParts = 256;

template <class SS>
__global__ __launch_bounds__(kMaxBlock) void k_adj2d_absmax(const typename SS::raw *__restrict__ sino,
                                                           const float4 *__restrict__ ns,
                                                           size_t rays, float *__restrict__ parts) {
    __shared__ float red[kMaxBlock / 64];
    const size_t b = blockIdx.y;
    sino += b * rays;
    parts += b * kAdjParts;
    float m = 0.0f;
    for (size_t i = (size_t)blockIdx.x * kMaxBlock + threadIdx.x; i < rays;
         i += (size_t)kAdjParts * kMaxBlock)
        m = fmaxf(m, fabsf(to_f32<SS>(sino[i]) * ns[i].x));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kMaxBlock / 64; ++w) m = fmaxf(m, red[w]);
        parts[blockIdx.x] = m;
    }
}

__device__ __forceinline__ float adj2d_ymax(const float *__restrict__ parts) {
    float m = 0.0f;
    for (int i = threadIdx.x & 63; i < kAdjParts; i += 64) m = fmaxf(m, parts[i]);
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    return m;
}

// Matched adjoint (exact transpose of k_fwd2d), tile-driven. A workgroup owns a kT2 x kT2 texel
// tile and a group of angles. Per angle, the candidate rays are those through the tile's
// footprint box (every texel's [t - .5, t + 1.5) box, in the ray search frame, projected onto
// the detector through its corners); the candidates of kAdjAngles angles form one list, and
// each thread takes one ray, keeps its parameters in registers and walks the kT2 + 1 dominant
// rows whose samples reach the tile. A sample on row pd (lower tap) with column c adds y q.x
// times its forward weights to the texels (pd, floor c), (pd, floor c + 1), (pd + 1, ...)
// (fan2d.hip header), in fixed point with integer LDS atomics: gfx950 runs them about 45 times
// faster than FP32 LDS atomic adds, and integer sums are exact, so the result does not depend on
// the order of the additions. A term is rounded to units of max|y q.x| / 2^19 (so |term| <=
// 2^19) and added to a 32-bit LDS cell: half the LDS atomic traffic of 64-bit cells, which
// bounds this kernel. A candidate ray adds at most two terms to a cell (one sample per dominant
// row, and a cell takes the lower-row weights of one sample and the upper-row weights of the
// sample before), so a run of passes whose candidates number c adds at most 2 c 2^19 to a cell:
// each thread moves its share of the cells into 64-bit registers and clears them before that
// bound could pass 2^31 - 1 (runs of at most 2047 candidates), which keeps every 32-bit sum
// exact. y- and x-dominant rays add to separate tiles whose rows run along their dominant axis,
// padded by one row and column on each side: the four taps of a sample are then four immediate
// offsets from one LDS address and need no bounds checks. Each tile then adds its sums, in units
// of max|y q.x| / 2^30, to a 64-bit image with integer global atomics (several angle groups
// share a tile), and k_adj2d_finish adds that image to the volume.
constexpr int kT2 = 32;       // texel tile edge
constexpr int kAdjAngles = 8; // angles whose candidate rays are staged together
// Candidate order within an angle: phase p holds the rays p, p + kAdjPh, p + 2 kAdjPh, ..., so
// the lanes of a wavefront take rays kAdjPh cells apart and add to different LDS cells more
// often (padding the phases to equal size leaves holes, which are skipped). At 1024^2: kAdjPh
// 1 (adjacent rays): 1.23 ms, 2: 1.15, 3: 1.21, 4: 1.27; bit-identical.
constexpr int kAdjPh = 2;
static_assert(kAdjAngles <= 16, "lanes 4 s .. 4 s + 3 of a wavefront project angle s");
constexpr int kT2Pad = kT2 + 2;   // padded rows and columns
constexpr int kT2Pitch = kT2 + 3; // odd pitch in words
constexpr int kT2Cells = 2 * kT2Pad * kT2Pitch;
constexpr int kT2Own = (kT2Cells + kMaxBlock - 1) / kMaxBlock; // cells per thread
constexpr int kAdjBits = 19;        // a term's bound: 2^kAdjBits units
// candidates per 32-bit accumulation run: 2 kAdjBudget 2^kAdjBits <= 2^31 - 2^20
constexpr int kAdjBudget = (1 << (31 - kAdjBits - 1)) - 1;

template <class SS>
__global__ __launch_bounds__(kMaxBlock) void k_adj2d_tile(unsigned long long *__restrict__ accimg,
                                                         const typename SS::raw *__restrict__ sino,
                                                         const float *__restrict__ ageo,
                                                         const float4 *__restrict__ ev,
                                                         const float4 *__restrict__ ns,
                                                         const float *__restrict__ parts, int nx,
                                                         int ny, int D, int A, int G) {
    // tile[0]: y-dominant rays, rows along y; tile[1]: x-dominant rays, rows along x. Padded
    // by one row and column before and after, so every tap of a kept sample lands in the array.
    // 32-bit cells while the angles are added; 64-bit cells (the same memory) for the totals.
    __shared__ union {
        unsigned int c32[2][kT2Pad * kT2Pitch];
        unsigned long long c64[2][kT2Pad * kT2Pitch];
    } sh;
    auto &tile = sh.c32;
    const int tid = threadIdx.x;
    long long own[kT2Own]; // this thread's cells tid + k kMaxBlock, moved out of LDS
#pragma unroll
    for (int k = 0; k < kT2Own; ++k) own[k] = 0;
    // cells -> registers, then clear (every thread calls it)
    auto drain = [&]() {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kT2Own; ++k) {
            const int i = tid + k * kMaxBlock;
            if (i < kT2Cells) {
                own[k] += (int)(&tile[0][0])[i];
                (&tile[0][0])[i] = 0u;
            }
        }
        __syncthreads();
    };
    int run = 0; // candidates added since the last drain
    // XCD-aware order: workgroups are dealt to the 8 XCDs in turn, and logical index li gives
    // each XCD a contiguous run of tiles of one angle group (x fastest), whose candidate rays
    // share L2 (1.15 -> 1.08 ms at 1024^2; angle groups fastest: 1.17 ms)
    int g, tx0, ty0;
    size_t b;
    {
        const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
        const unsigned nwg = gx * gy * gz;
        const unsigned hw = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
        const unsigned n8 = nwg & ~7u;
        const unsigned li = hw < n8 ? (hw & 7u) * (n8 >> 3) + (hw >> 3) : hw;
        const unsigned per = gx * gy * (unsigned)G;
        b = li / per;
        const unsigned l = li % per;
        tx0 = (int)(l % gx) * kT2;
        ty0 = (int)((l / gx) % gy) * kT2;
        g = (int)(l / (gx * gy));
    }
    const int a0 = (int)((int64_t)g * A / G), a1 = (int)((int64_t)(g + 1) * A / G);
    for (int i = tid; i < kT2Cells; i += kMaxBlock) (&tile[0][0])[i] = 0u;
    __syncthreads(); // the tiles are clear before any wave adds to them
    const float ymax = adj2d_ymax(parts + b * kAdjParts);
    const float yscale = ymax > 0.0f ? (float)(1 << kAdjBits) / ymax : 0.0f;
    // the tile's first texel in the ray table's centred padded frame
    const int px0 = tx0 + kPad - (nx + 2 * kPad) / 2, py0 = ty0 + kPad - (ny + 2 * kPad) / 2;
    // footprint box corner this lane projects (lanes 0..3 of every wavefront)
    const float qx = (tid & 1) ? (float)tx0 + (float)kT2 + 0.5f : (float)tx0 - 0.5f;
    const float qy = (tid & 2) ? (float)ty0 + (float)kT2 + 0.5f : (float)ty0 - 0.5f;
    const typename SS::raw *sb = sino + b * (size_t)A * D;
    // kAdjAngles angles at a time: their candidate rays form one list, one ray per thread
    for (int ag = a0; ag < a1; ag += kAdjAngles) {
        const int na = min(kAdjAngles, a1 - ag);
        // candidate range per angle: lanes 4 s .. 4 s + 3 of every wavefront project the four
        // corners of the footprint box for angle ag + s
        const int slot = (tid & 63) >> 2;
        float kk = 0.0f;
        if (slot < na) {
            const float *gp = ageo + 6 * (ag + slot);
            const float wx = qx - gp[0], wy = qy - gp[1];
            kk = -(wx * gp[3] - wy * gp[2]) / (wx * gp[5] - wy * gp[4]);
        }
        float kmin = kk, kmax = kk;
        for (int off = 1; off < 4; off <<= 1) {
            kmin = fminf(kmin, __shfl_xor(kmin, off, 64));
            kmax = fmaxf(kmax, __shfl_xor(kmax, off, 64));
        }
        int k0s[kAdjAngles], offs[kAdjAngles], cnts[kAdjAngles];
        int total = 0;
#pragma unroll
        for (int q = 0; q < kAdjAngles; ++q) {
            const float lo = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(kmin), 4 * q));
            const float hi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(kmax), 4 * q));
            int cnt = 0, k0 = 0;
            // an angle whose box misses the detector (or gives NaN) has no candidates
            if (q < na && hi >= -1.0f && lo <= (float)D) {
                k0 = max((int)floorf(lo - 0.01f), 0);
                cnt = max(min((int)ceilf(hi + 0.01f), D - 1) - k0 + 1, 0);
            }
            k0s[q] = k0;
            offs[q] = total;
            cnts[q] = cnt;
            total += kAdjPh * ((cnt + kAdjPh - 1) / kAdjPh);
        }
        // one candidate ray per thread, its parameters in registers; `total` and the pass
        // sizes are the same for the whole workgroup
        for (int base = 0; base < total; base += kMaxBlock) {
            const int pass = min(kMaxBlock, total - base);
            if (run + pass > kAdjBudget) {
                drain();
                run = 0;
            }
            run += pass;
            const int ri = base + tid;
            if (ri >= total) continue;
            int q = 0;
#pragma unroll
            for (int u = 0; u < kAdjAngles; ++u)
                if (ri >= offs[u]) q = u;
            const int pm = (cnts[q] + kAdjPh - 1) / kAdjPh;
            const int pp = ri - offs[q];
            const int rl = (pp % pm) * kAdjPh + pp / pm;
            if (pm == 0 || rl >= cnts[q]) continue;
            const size_t r = (size_t)(ag + q) * D + k0s[q] + rl;
            const float4 qn = ns[r];
            const int n = __float_as_int(qn.y);
            if (n <= 0) continue;
            const float yv = to_f32<SS>(sb[r]) * qn.x * yscale; // |yv| <= 2^kAdjBits
            if (yv == 0.0f) continue;
            const float4 e = ev[r];
            const bool ydom = fabsf(e.w) == 1.0f;
            const float md = ydom ? e.y : e.x, sd = ydom ? e.w : e.z;
            const float mc = ydom ? e.x : e.y, vc = ydom ? e.z : e.w;
            const int r0 = (int)floorf(md);
            const float fb = md - (float)r0;
            const float wl = yv * (1.0f - fb), wh = yv * fb; // rows pd, pd + 1
            const int s = sd > 0.0f ? 1 : -1;
            const int jlo = (int)qn.z, jhi = jlo + n - 1;
            const int d0 = ydom ? py0 : px0, c0 = ydom ? px0 : py0;
            // Step t is the sample whose dominant lower tap is pd = d0 - 1 + t (tile row t - 1):
            // jf = s (pd - r0). Keep the steps with jf in [jlo, jhi] that reach the tile's rows.
            const int ta = s > 0 ? jlo + r0 - d0 + 1 : r0 - jhi - d0 + 1;
            const int t_lo = max(ta, 0), t_hi = min(ta + n - 1, kT2);
            // Step t adds wl [1 - fa, fa] to tile row t - 1 and wh [1 - fa, fa] to row t at the
            // columns lc, lc + 1: padded rows t, t + 1 and columns lc + 1, lc + 2, four
            // immediate offsets from one address. Steps with lc outside [-1, kT2) miss the tile.
            unsigned int *row = &tile[ydom ? 0 : 1][t_lo * kT2Pitch + 1];
            float jf = (float)(s * (d0 - 1 + t_lo - r0));
            const float fs = (float)s;
            for (int t = t_lo; t <= t_hi; ++t) {
                const float c = fmaf(jf, vc, mc);
                jf += fs;
                const int lc = floor_i32(c) - c0; // the lower tap's column in the tile
                if ((unsigned)(lc + 1) <= (unsigned)kT2) {
                    const float fa = __builtin_amdgcn_fractf(c);
                    const float v1 = wl * fa, v0 = wl - v1, u1 = wh * fa, u0 = wh - u1;
                    unsigned int *q = row + lc;
                    atomicAdd(q, (unsigned int)round_i32(v0));
                    atomicAdd(q + 1, (unsigned int)round_i32(v1));
                    atomicAdd(q + kT2Pitch, (unsigned int)round_i32(u0));
                    atomicAdd(q + kT2Pitch + 1, (unsigned int)round_i32(u1));
                }
                row += kT2Pitch;
            }
        }
    }
    drain();
    // totals in units of max|y q.x| / 2^30
#pragma unroll
    for (int k = 0; k < kT2Own; ++k) {
        const int i = tid + k * kMaxBlock;
        if (i < kT2Cells) (&sh.c64[0][0])[i] = (unsigned long long)(own[k] * (1ll << (30 - kAdjBits)));
    }
    __syncthreads();
    unsigned long long *out = accimg + b * (size_t)nx * ny;
    for (int i = tid; i < kT2 * kT2; i += kMaxBlock) {
        const int lx = i % kT2, ly = i / kT2;
        const int X = tx0 + lx, Y = ty0 + ly;
        const unsigned long long v =
            sh.c64[0][(ly + 1) * kT2Pitch + lx + 1] + sh.c64[1][(lx + 1) * kT2Pitch + ly + 1];
        if (X < nx && Y < ny && v != 0ull) atomicAdd(&out[(size_t)Y * nx + X], v);
    }
}

// vol += accimg * max|y q.x| / 2^30.
template <class SV>
__global__ __launch_bounds__(kMaxBlock) void k_adj2d_finish(typename SV::raw *__restrict__ vol,
                                                           unsigned long long *__restrict__ accimg,
                                                           const float *__restrict__ parts,
                                                           size_t img) {
    const size_t b = blockIdx.y;
    const float ymax = adj2d_ymax(parts + b * kAdjParts);
    const double unit = (double)ymax * 0x1p-30;
    const size_t j = (size_t)blockIdx.x * kMaxBlock + threadIdx.x;
    if (j >= img) return;
    const size_t i = b * img + j;
    const long long q = (long long)accimg[i];
    vol[i] = from_f32<SV>(to_f32<SV>(vol[i]) + (float)((double)q * unit));
}

__global__ __launch_bounds__(kMaxBlock) void k_fill_u64(unsigned long long *__restrict__ p,
                                                        size_t n) {
    const size_t i = (size_t)blockIdx.x * kMaxBlock + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ------------------------------------------------ transpose of the backprojection -------

// Bᵀ: the transpose of the voxel-driven backprojection (k_bp2d / k_bp2d_vec), the gradient of
// TorchBackprojection. The backprojection gives pixel p, per projection a, wt_a(p) times the
// linear interpolation of row a at u_a(p) (cells floor(u) and floor(u) + 1, weights 1 - al and
// al; cells outside the row read zero), times 1/du (circular plans; the view's 1/|u| is part of
// wt for plans from views). So sino[a][c] = sum_p (those weights) v[p], with k_bp2d's FP32
// positions and weights.
// A workgroup owns a segment of kGS detector cells of one projection and gathers it. The pixels
// whose terms reach the segment's cells are those with u in [c0 - 1, c0 + kGS): a wedge from
// the source between the lines u = c0 - 1 and u = c0 + kGS (u is a ratio of two functions
// affine in the pixel). The workgroup marches through the image along the axis K (rows or
// columns) that the segment's central line follows most closely, in chunks of kGSlab lines;
// per line the wedge is the interval between the two boundary lines' crossings (one pixel of
// margin each side, clipped to the image), walked by the workgroup's threads along the line,
// from the zero-bordered copy whose rows run along it (row-major for rows, transposed for
// columns), so that a wavefront reads runs of consecutive pixels. Every pixel computes its own
// u and weight with the backprojector's arithmetic and adds its terms that fall inside the
// segment to kGS + 2 LDS cells with integer atomics. A segment whose boundary lines meet the
// lines at grazing angles walks whole lines.
// Fixed point: terms are rounded to units of max|v| * bound / 2^kGBits, where `bound` is the
// plan's bound on wt (times 1/du) over the image and all projections, so |term| <= 2^kGBits
// and a 32-bit LDS cell holds the sum of up to 2^(31 - kGBits) - 1 terms. A pixel adds at most
// one term to a cell, so the cells are moved into 64-bit registers (one per thread of the first
// kGS) and cleared before the pixel visits since the last such drain could pass that count.
// Integer sums are exact: the result does not depend on the order of the additions. Each
// segment is written once, in the sinogram's storage type.
constexpr int kGS = 64;                                      // detector cells per segment
constexpr int kGC = kGS + 2;                                 // LDS cells: c0 - 1 .. c0 + kGS
constexpr int kGSlab = 64;                                   // image lines per chunk
constexpr int kGBits = 18;                                   // a term's bound: 2^kGBits units
constexpr int kGBudget = (1 << (31 - kGBits)) - 1;           // pixel visits between drains
// pixels per thread at a time (at 1024^2: 1: 0.95 ms, 2: 0.91, 4: 0.97, 8: 1.16)
constexpr int kGBatch = 2;
constexpr int kGPass = kGBudget / (kGBatch * kMaxBlock) * (kGBatch * kMaxBlock); // per pass
static_assert(kGS <= kMaxBlock && kGSlab <= 64, "one owned cell per thread, one lane per line");

// Largest |v| of each image of a batch, in kAdjParts partial maxima per image.
template <class SV>
__global__ __launch_bounds__(kMaxBlock) void k_absmax_img2d(const typename SV::raw *__restrict__ v,
                                                            size_t n, float *__restrict__ parts) {
    __shared__ float red[kMaxBlock / 64];
    const size_t b = blockIdx.y;
    v += b * n;
    parts += b * kAdjParts;
    float m = 0.0f;
    for (size_t i = (size_t)blockIdx.x * kMaxBlock + threadIdx.x; i < n;
         i += (size_t)kAdjParts * kMaxBlock)
        m = fmaxf(m, fabsf(to_f32<SV>(v[i])));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kMaxBlock / 64; ++w) m = fmaxf(m, red[w]);
        parts[blockIdx.x] = m;
    }
}

// One projection's detector coordinate u (lower-tap cell floor(u), clamped to [-2, D] as the
// backprojector clamps it) and weight wt of the pixel at world (x, y): k_bp2d's arithmetic for
// circular plans (tab = meta, wt without the 1/du), k_bp2d_vec's for plans from views
// (tab = three float4 per projection, wt including 1/|u|).
template <bool VIEWS, int WPOW>
__device__ __forceinline__ void bpT_geom(const float4 *__restrict__ tab, int a, float x, float y,
                                         float px, float py, float c0, float hi, float &u,
                                         float &wt) {
    if (VIEWS) {
        const float4 m0 = tab[3 * a], m1 = tab[3 * a + 1];
        const float idu = tab[3 * a + 2].x;
        const float w = m0.w * __builtin_amdgcn_rcpf(fmaf(m0.x, x, fmaf(m0.y, y, m0.z)));
        const float eta = fmaf(m1.x, x, fmaf(m1.y, y, m1.z));
        u = __builtin_amdgcn_fmed3f(fmaf(w, eta, m1.w), -2.0f, hi);
        wt = (WPOW == 2 ? w * w : w) * idu;
    } else {
        const float4 m = tab[a];
        const float depth = fmaf(m.x, x, fmaf(-m.y, y, m.w)); // xi + DSO
        const float eta = fmaf(m.y, px, m.x * py);
        const float w = m.z * __builtin_amdgcn_rcpf(depth);
        u = __builtin_amdgcn_fmed3f(fmaf(w, eta, c0), -2.0f, hi);
        wt = WPOW == 2 ? w * w : w;
    }
}

// The line u(x, y) = uc of projection a as l0 x + l1 y + l2 = 0 in world coordinates (FP64),
// from the same table as bpT_geom: u = w eta + c with w and eta ratios / sums of affine terms.
template <bool VIEWS>
__device__ __forceinline__ void bpT_line(const float4 *__restrict__ tab, int a, double uc,
                                         double inv_du, double c0, double l[3]) {
    if (VIEWS) {
        const float4 m0 = tab[3 * a], m1 = tab[3 * a + 1];
        const double k = uc - (double)m1.w; // k (m0 . (x, y, 1)) = m0.w (m1 . (x, y, 1))
        l[0] = k * m0.x - (double)m0.w * m1.x;
        l[1] = k * m0.y - (double)m0.w * m1.y;
        l[2] = k * m0.z - (double)m0.w * m1.z;
    } else {
        const float4 m = tab[a];
        const double k = uc - c0; // k (m.x x - m.y y + m.w) = m.z (m.y x + m.x y) / du
        l[0] = k * m.x - (double)m.z * inv_du * m.y;
        l[1] = -k * m.y - (double)m.z * inv_du * m.x;
        l[2] = k * m.w;
    }
}

template <class SV, class SS, int WPOW, bool VIEWS>
__global__ __launch_bounds__(kMaxBlock) void k_bpT2d(typename SS::raw *__restrict__ sino,
                                                     const typename SV::raw *__restrict__ staged,
                                                     const float4 *__restrict__ tab,
                                                     const float *__restrict__ parts, float bound,
                                                     int nx, int ny, int D, int A, float sx,
                                                     float sy, float ox, float oy, float inv_du,
                                                     float half_d) {
    __shared__ unsigned int cells[kGC]; // cell c0 - 1 + i at i
    __shared__ int2 span[2][kGSlab];    // per line of a chunk: pixel range along it (inclusive)
    __shared__ int width[2];            // per chunk: the widest range
    const int tid = threadIdx.x, lane = tid & 63;
    const int a = (int)(blockIdx.x % (unsigned)A);
    const int c0 = (int)(blockIdx.x / (unsigned)A) * kGS;
    const size_t b = blockIdx.y;
    const float hi = (float)D, uoff = half_d - 0.5f;
    // fixed-point scale 2^18 / (max|v| bound) as an FP32 mantissa and an exponent
    static_assert(kGBits == 18, "the scale uses 2^18");
    const float vmax = adj2d_ymax(parts + b * kAdjParts);
    const double unit_inv = vmax > 0.0f ? 0x1p18 / ((double)vmax * (double)bound) : 0.0;
    const int sexp = unit_inv > 0.0 ? ilogb(unit_inv) : 0;
    const float smant = (float)ldexp(unit_inv, -sexp);
    // pixel index -> world: x = ix sx + bx (the world() map of the backprojector)
    const double bx = (double)ox + 0.5 - (double)((float)nx / 2.0f) * sx;
    const double by = (double)oy + 0.5 - (double)((float)ny / 2.0f) * sy;
    // the segment's central and boundary lines in index coordinates: n0 ix + n1 iy + n2 = 0
    auto index_line = [&](double uc, double (&n)[3]) {
        double l[3];
        bpT_line<VIEWS>(tab, a, uc, (double)inv_du, (double)uoff, l);
        n[0] = l[0] * sx;
        n[1] = l[1] * sy;
        n[2] = l[0] * bx + l[1] * by + l[2];
    };
    // march along rows (K = y, ranges along x) when the central line runs closer to y
    int rows;
    {
        double n[3];
        index_line((double)c0 + 0.5 * kGS, n);
        rows = fabs(n[0]) >= fabs(n[1]) ? 1 : 0;
        rows = __builtin_amdgcn_readfirstlane(rows);
    }
    const int nK = rows ? ny : nx, nF = rows ? nx : ny;
    // Boundary lines as F = f0 + f1 K. The wedge's section of a line of K is the interval
    // between the two crossings when both boundary rays cross the lines of K steeply and in the
    // same direction. A line's coefficients are (uc Q - P) for u = P / Q with Q > 0 in front of
    // the source, so rotating its normal by a right angle gives a direction whose orientation
    // relative to the rays' forward direction is the same for every uc (the 2 x 2 determinant
    // of P and Q does not depend on uc): equal signs of the normals' F components mean equal
    // directions along K.
    float f0[2], f1[2];
    bool whole = false;
    double sgn[2];
    for (int e = 0; e < 2; ++e) {
        double n[3];
        index_line(e ? (double)(c0 + kGS) : (double)(c0 - 1), n);
        const double nF_ = rows ? n[0] : n[1], nK_ = rows ? n[1] : n[0];
        sgn[e] = nF_;
        if (!(fabs(nF_) >= 0.25 * sqrt(nF_ * nF_ + nK_ * nK_)) || nF_ == 0.0) {
            whole = true;
            f0[e] = f1[e] = 0.0f;
        } else {
            f0[e] = (float)(-n[2] / nF_);
            f1[e] = (float)(-nK_ / nF_);
        }
    }
    if ((sgn[0] > 0.0) != (sgn[1] > 0.0)) whole = true;
    const int Wp = nx + 2 * kPad, Hp = ny + 2 * kPad;
    // pixel (f, k) along (F, K) lies at src[f + k strK] of the copy whose rows run along F
    const typename SV::raw *src = staged + b * 2 * (size_t)Wp * Hp +
                                  (rows ? (size_t)0 : (size_t)Wp * Hp) +
                                  (size_t)kPad * (rows ? Wp : Hp) + kPad;
    const int strK = rows ? Wp : Hp;
    for (int i = tid; i < kGC; i += kMaxBlock) cells[i] = 0u;
    long long own = 0;
    auto drain = [&]() {
        __syncthreads();
        if (tid < kGS) {
            own += (long long)(int)cells[tid + 1];
            cells[tid + 1] = 0u;
        }
        __syncthreads();
    };
    int run = 0; // pixel visits since the last drain
    int buf = 0;
    for (int k0 = 0; k0 < nK; k0 += kGSlab, buf ^= 1) {
        const int S = min(kGSlab, nK - k0);
        if (tid < 64) {
            int lo = 0, up = -1;
            if (lane < S) {
                if (whole) {
                    up = nF - 1;
                } else {
                    const float k = (float)(k0 + lane);
                    const float g0 = fmaf(f1[0], k, f0[0]), g1 = fmaf(f1[1], k, f0[1]);
                    constexpr float big = 0x1p24f;
                    lo = max(0, floor_i32(__builtin_amdgcn_fmed3f(fminf(g0, g1), -big, big)) - 1);
                    up = min(nF - 1,
                             floor_i32(__builtin_amdgcn_fmed3f(fmaxf(g0, g1), -big, big)) + 1);
                }
                span[buf][lane] = make_int2(lo, up);
            }
            int w = max(up - lo + 1, 0);
            for (int off = 32; off > 0; off >>= 1) w = max(w, __shfl_xor(w, off, 64));
            if (lane == 0) width[buf] = w;
        }
        __syncthreads();
        const int NW = width[buf];
        const int total = NW * S;
        if (total == 0) continue;
        // element e = tid + n kMaxBlock of the chunk as (dF, dk), dF fastest; a step of
        // kMaxBlock elements is (sF, sK). Quotients by NW through an FP32 reciprocal: (m + 0.5)
        // / NW lies at least 0.5 / NW from an integer, far more than the rounding for m < 2^20.
        const float rNW = 1.0f / (float)NW;
        int dk = (int)(((float)tid + 0.5f) * rNW), dF = tid - dk * NW;
        const int sK = (int)(((float)kMaxBlock + 0.5f) * rNW), sF = kMaxBlock - sK * NW;
        int e = tid;
        for (int base = 0; base < total; base += kGPass) {
            const int n = min(kGPass, total - base);
            if (run + n > kGBudget) {
                drain();
                run = 0;
            }
            run += n;
            // kGBatch elements per thread at a time, their loads issued together; passes start
            // at multiples of kGBatch kMaxBlock elements, so a batch never straddles two
            for (; e < base + n; e += kGBatch * kMaxBlock) {
                int fq[kGBatch], kq[kGBatch];
                bool live[kGBatch];
                float val[kGBatch];
#pragma unroll
                for (int q = 0; q < kGBatch; ++q) {
                    const int2 sp = span[buf][min(dk, S - 1)];
                    fq[q] = sp.x + dF;
                    kq[q] = k0 + dk;
                    live[q] = e + q * kMaxBlock < base + n && fq[q] <= sp.y;
                    val[q] = to_f32<SV>(src[live[q] ? fq[q] + kq[q] * strK : 0]);
                    dF += sF;
                    if (dF >= NW) {
                        dF -= NW;
                        ++dk;
                    }
                    dk += sK;
                }
#pragma unroll
                for (int q = 0; q < kGBatch; ++q) {
                    const float vs = live[q] ? ldexpf(val[q] * smant, sexp) : 0.0f;
                    if (vs == 0.0f) continue;
                    const int ix = rows ? fq[q] : kq[q], iy = rows ? kq[q] : fq[q];
                    const float x = ((float)ix - (float)nx / 2.0f) * sx + ox + 0.5f;
                    const float y = ((float)iy - (float)ny / 2.0f) * sy + oy + 0.5f;
                    float u, wt;
                    bpT_geom<VIEWS, WPOW>(tab, a, x, y, x * inv_du, y * inv_du, uoff, hi, u, wt);
                    const int i = floor_i32(u) - (c0 - 1);
                    if ((unsigned)i > (unsigned)kGS) continue;
                    const float al = __builtin_amdgcn_fractf(u);
                    const float t = vs * wt * (VIEWS ? 1.0f : inv_du);
                    const float t1 = t * al, t0 = t - t1;
                    atomicAdd(&cells[i], (unsigned)round_i32(t0));
                    atomicAdd(&cells[i + 1], (unsigned)round_i32(t1));
                }
            }
        }
    }
    __syncthreads();
    if (tid < kGS) {
        own += (long long)(int)cells[tid + 1];
        const int c = c0 + tid;
        if (c < D) {
            const double unit = (double)vmax * (double)bound * 0x1p-18;
            sino[(b * A + a) * (size_t)D + c] = from_f32<SS>((float)((double)own * unit));
        }
    }
}

// ---------------------------------------------------------------- dispatch -----------------

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

cmi_status check_call(const cmi_plan2d *p, const void *vol, int vol_dtype, const void *sino,
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
    const uint64_t img = (uint64_t)p->g.nx * p->g.ny, sn = (uint64_t)p->A * (p->g.D + 2 * kPad);
    const uint64_t pad = 2 * (uint64_t)(p->g.nx + 2 * kPad) * (p->g.ny + 2 * kPad);
    if (std::max({img, sn, pad}) * 4 >= (1ull << 31))
        return fail(CMI_ERR_INVALID_ARGUMENT, "%s: one image or sinogram exceeds 2 GiB", fn);
    if (batch > 65535) return fail(CMI_ERR_INVALID_ARGUMENT, "%s: batch exceeds 65535", fn);
    CMI_TRY(check_device_ptr(vol, p->device, "volume"));
    CMI_TRY(check_device_ptr(sino, p->device, "sinogram"));
    return CMI_OK;
}

// Checks and fields common to both plan constructors (Python order (y, x) -> (x, y), as
// plan.cu:48-50).
cmi_status check_plan_shape(int vol_prec, int det_prec, const uint64_t vol_shape[2],
                            const double vol_spacing[2], uint64_t det_count, uint64_t nangles,
                            const char *fn) {
    if (!dtype_ok(vol_prec) || !dtype_ok(det_prec))
        return fail(CMI_ERR_DTYPE, "%s: precision must be f32, f16 or bf16", fn);
    if (vol_shape[0] == 0 || vol_shape[1] == 0 || det_count == 0 || nangles == 0)
        return fail(CMI_ERR_INVALID_ARGUMENT,
                    "%s: empty shape (vol %llu x %llu, det %llu, angles %llu)", fn,
                    (unsigned long long)vol_shape[0], (unsigned long long)vol_shape[1],
                    (unsigned long long)det_count, (unsigned long long)nangles);
    if (vol_shape[0] > (1u << 20) || vol_shape[1] > (1u << 20) || det_count > (1u << 24) ||
        nangles > (1u << 24) || det_count * nangles >= (1ull << 31))
        return fail(CMI_ERR_INVALID_ARGUMENT, "%s: shape too large", fn);
    if (!(vol_spacing[0] > 0) || !(vol_spacing[1] > 0))
        return fail(CMI_ERR_INVALID_ARGUMENT, "%s: spacings must be positive", fn);
    return CMI_OK;
}

void set_plan_shape(cmi_plan2d &p, int device, int vol_prec, const uint64_t vol_shape[2],
                    const double vol_spacing[2], const double vol_offset[2], int det_prec,
                    uint64_t det_count, uint64_t nangles) {
    p.device = device;
    p.vol_prec = vol_prec;
    p.det_prec = det_prec;
    p.g.nx = (int64_t)vol_shape[1];
    p.g.ny = (int64_t)vol_shape[0];
    p.g.sx = vol_spacing[1];
    p.g.sy = vol_spacing[0];
    p.g.ox = vol_offset[1];
    p.g.oy = vol_offset[0];
    p.g.D = (int64_t)det_count;
    p.A = (int64_t)nangles;
}

// Whether the backprojection by quarter turns (k_bp2d_qt) applies: a square image with an even
// side, unit pixels and no offset (the pixel centres, at (i - n/2) + 1/2 as the reference places
// them, backward_2d.hpp:58-59, are then symmetric under quarter turns), angles in quadruples
// a + k A/4 that are quarter turns apart with equal DSO and DSD, a padded detector row at least
// as long as a staged segment, and every 16 x 16 tile's detector footprint within a staged
// segment: the projection of a segment of length L at
// depth >= DSO - r (r the largest pixel-centre radius) spans at most
// L DSD / (DSO - r) sqrt(1 + (r / (DSO - r))^2) / du cells.
bool quarter_turns(const fan2d::Scalars &g, const double *angles, const double *DSO,
                   const double *DSD, int64_t A) {
    if (g.nx != g.ny || g.nx % 2 != 0 || g.nx < 2 || g.sx != 1.0 || g.sy != 1.0 ||
        g.ox != 0.0 || g.oy != 0.0 || A % 4 != 0 || A < 4 || g.D + 2 * kPad < kQSeg)
        return false;
    const int64_t Q = A / 4;
    const double r = ((double)(g.nx / 2) - 0.5) * std::sqrt(2.0);
    double wmax = 0.0;
    for (int64_t a = 0; a < A; ++a) {
        if (!(DSO[a] > r + 1.0)) return false;
        const double ratio = r / (DSO[a] - r);
        wmax = std::max(wmax, DSD[a] / (DSO[a] - r) * std::sqrt(1.0 + ratio * ratio));
        if (a >= Q) continue;
        for (int64_t k = 1; k < 4; ++k) {
            const int64_t b = a + k * Q;
            if (DSO[b] != DSO[a] || DSD[b] != DSD[a]) return false;
            double d = std::remainder(angles[b] - angles[a] - (double)k * M_PI / 2.0, 2.0 * M_PI);
            if (std::fabs(d) > 1e-9) return false;
        }
    }
    const double span = wmax * (kQT - 1) * std::sqrt(2.0) / g.du;
    return span + 4.0 <= (double)kQSeg;
}

// Shared tail of the plan constructors: per-angle tables to the device, then the FP64 ray
// table (k_ray_table). `meta` (circular plans) or `bpv` (plans from views) is empty.
cmi_status finish_plan2d(std::unique_ptr<cmi_plan2d> p, const std::vector<AngleGeom> &geo,
                         const std::vector<float4> &meta, const std::vector<float4> &bpv,
                         cmi_plan2d **plan) {
    const int64_t A = p->A, D = p->g.D;
    // the backprojection's weight bounds: depth is affine over the image, so its minimum over
    // the pixel centres is at a corner; 1 % of margin for the kernels' FP32 arithmetic
    {
        const double nx = (double)p->g.nx, ny = (double)p->g.ny;
        const double xs[2] = {(0.0 - nx / 2.0) * p->g.sx + p->g.ox + 0.5,
                              (nx - 1.0 - nx / 2.0) * p->g.sx + p->g.ox + 0.5};
        const double ys[2] = {(0.0 - ny / 2.0) * p->g.sy + p->g.oy + 0.5,
                              (ny - 1.0 - ny / 2.0) * p->g.sy + p->g.oy + 0.5};
        double b1 = 0.0, b2 = 0.0;
        bool ok = true;
        for (int64_t a = 0; a < A && ok; ++a) {
            for (int c = 0; c < 4; ++c) {
                const double x = xs[c & 1], y = ys[c >> 1];
                double depth, L, scale;
                if (!meta.empty()) {
                    const float4 m = meta[a];
                    depth = (double)m.x * x - (double)m.y * y + (double)m.w;
                    L = m.z;
                    scale = 1.0 / p->g.du;
                } else {
                    const float4 m0 = bpv[3 * a];
                    depth = (double)m0.x * x + (double)m0.y * y + (double)m0.z;
                    L = m0.w;
                    scale = bpv[3 * a + 2].x;
                }
                if (!(depth > 0.0)) {
                    ok = false;
                    break;
                }
                const double w = std::fabs(L) / depth;
                b1 = std::max(b1, w * scale);
                b2 = std::max(b2, w * w * scale);
            }
        }
        if (ok && b1 > 0.0 && std::isfinite(b2)) {
            p->bt_bound[0] = 1.01 * b1;
            p->bt_bound[1] = 1.01 * b2;
        }
    }
    std::vector<float> ageo(6 * A);
    for (int64_t a = 0; a < A; ++a)
        for (int d = 0; d < 2; ++d) {
            ageo[6 * a + d] = (float)geo[a].src[d];
            ageo[6 * a + 2 + d] = (float)geo[a].org[d];
            ageo[6 * a + 4 + d] = (float)geo[a].del[d];
        }
    DeviceGuard guard;
    CMI_TRY(guard.enter(p->device));
    struct Cleanup {
        cmi_plan2d *p;
        bool armed = true;
        ~Cleanup() {
            if (!armed) return;
            (void)hipFree(p->d_geo);
            (void)hipFree(p->d_meta);
            (void)hipFree(p->d_bpv);
            (void)hipFree(p->d_ageo);
            (void)hipFree(p->d_ev);
            (void)hipFree(p->d_ns);
        }
    } cleanup{p.get()};
    const size_t rays = (size_t)A * D;
    CMI_HIP_TRY(hipMalloc(&p->d_geo, sizeof(AngleGeom) * A));
    if (!meta.empty()) CMI_HIP_TRY(hipMalloc(&p->d_meta, sizeof(float4) * A));
    if (!bpv.empty()) CMI_HIP_TRY(hipMalloc(&p->d_bpv, sizeof(float4) * bpv.size()));
    CMI_HIP_TRY(hipMalloc(&p->d_ageo, sizeof(float) * 6 * A));
    CMI_HIP_TRY(hipMalloc(&p->d_ev, sizeof(float4) * rays));
    CMI_HIP_TRY(hipMalloc(&p->d_ns, sizeof(float4) * rays));
    p->table_bytes = sizeof(AngleGeom) * A + sizeof(float4) * (meta.size() + bpv.size()) +
                     sizeof(float) * 6 * A + 2 * sizeof(float4) * rays;
    unsigned long long *d_total = nullptr;
    CMI_HIP_TRY(hipMalloc(&d_total, sizeof(unsigned long long)));
    std::unique_ptr<unsigned long long, void (*)(unsigned long long *)> total_guard(
        d_total, [](unsigned long long *q) { (void)hipFree(q); });
    hipStream_t s;
    CMI_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::unique_ptr<ihipStream_t, void (*)(hipStream_t)> stream_guard(
        s, [](hipStream_t q) { (void)hipStreamDestroy(q); });
    CMI_HIP_TRY(hipMemcpyAsync(p->d_geo, geo.data(), sizeof(AngleGeom) * A,
                               hipMemcpyHostToDevice, s));
    if (!meta.empty())
        CMI_HIP_TRY(hipMemcpyAsync(p->d_meta, meta.data(), sizeof(float4) * A,
                                   hipMemcpyHostToDevice, s));
    if (!bpv.empty())
        CMI_HIP_TRY(hipMemcpyAsync(p->d_bpv, bpv.data(), sizeof(float4) * bpv.size(),
                                   hipMemcpyHostToDevice, s));
    CMI_HIP_TRY(hipMemcpyAsync(p->d_ageo, ageo.data(), sizeof(float) * 6 * A,
                               hipMemcpyHostToDevice, s));
    CMI_HIP_TRY(hipMemsetAsync(d_total, 0, sizeof(unsigned long long), s));
    dim3 blk(256), grd((unsigned)((D + 255) / 256), (unsigned)A);
    hipLaunchKernelGGL(k_ray_table, grd, blk, 0, s, p->d_geo, (int)A, (int)D, p->g.nx, p->g.ny,
                       p->g.sx, p->g.sy, p->d_ev, p->d_ns, d_total);
    CMI_HIP_TRY(hipGetLastError());
    unsigned long long total = 0;
    CMI_HIP_TRY(hipMemcpyAsync(&total, d_total, sizeof total, hipMemcpyDeviceToHost, s));
    CMI_HIP_TRY(hipStreamSynchronize(s));
    p->total_samples = total;
    cleanup.armed = false;
    *plan = p.release();
    return CMI_OK;
}

// Plans from ASTRA-like views (cmi_plan2d_create_vec, and cmi_plan2d_create_ex with exact
// geometry, which builds its views first).
cmi_status plan2d_from_views(int device, int vol_prec, const uint64_t vol_shape[2],
                             const double vol_spacing[2], const double vol_offset[2],
                             int det_prec, uint64_t det_count, const double *vecs,
                             uint64_t nangles, const char *fn, cmi_plan2d **plan) {
    CMI_TRY(check_plan_shape(vol_prec, det_prec, vol_shape, vol_spacing, det_count, nangles, fn));
    for (uint64_t a = 0; a < nangles; ++a) {
        const double *v = vecs + 6 * a;
        for (int i = 0; i < 6; ++i)
            if (!std::isfinite(v[i]))
                return fail(CMI_ERR_INVALID_ARGUMENT, "%s: non-finite view %llu", fn,
                            (unsigned long long)a);
        const fan2d::BpView bv = fan2d::bp_view(v, (int64_t)det_count);
        if (!(v[4] * v[4] + v[5] * v[5] > 0.0) || !(bv.L > 0.0))
            return fail(CMI_ERR_INVALID_ARGUMENT,
                        "%s: view %llu has a zero cell step or its source on the detector line",
                        fn, (unsigned long long)a);
    }
    auto p = std::make_unique<cmi_plan2d>();
    set_plan_shape(*p, device, vol_prec, vol_shape, vol_spacing, vol_offset, det_prec, det_count,
                   nangles);
    p->g.du = std::sqrt(vecs[4] * vecs[4] + vecs[5] * vecs[5]); // view 0's, for get_info
    const int64_t A = p->A;
    std::vector<AngleGeom> geo(A);
    std::vector<float4> bpv(3 * A);
    for (int64_t a = 0; a < A; ++a) {
        geo[a] = fan2d::plan_angle_vec(p->g, vecs + 6 * a);
        const fan2d::BpView bv = fan2d::bp_view(vecs + 6 * a, p->g.D);
        bpv[3 * a] = make_float4((float)bv.n[0], (float)bv.n[1], (float)bv.n0, (float)bv.L);
        // the cell offset carries the lerp's texel-centre shift (-0.5)
        bpv[3 * a + 1] =
            make_float4((float)bv.e[0], (float)bv.e[1], (float)bv.e0, (float)(bv.c0 - 0.5));
        bpv[3 * a + 2] = make_float4((float)bv.inv_du, 0.0f, 0.0f, 0.0f);
    }
    return finish_plan2d(std::move(p), geo, {}, bpv, plan);
}

} // namespace

// ------------------------------------------------------ forward, shared by the entry points --

namespace cmi {

cmi_status fan2d_forward_check(const cmi_plan2d *plan, const void *volume, int vol_dtype,
                               const void *sinogram, int sino_dtype, uint64_t batch) {
    return check_call(plan, volume, vol_dtype, sinogram, sino_dtype, batch, "forward_2d");
}

int fan2d_plan_device(const cmi_plan2d *plan) { return plan->device; }

// Two zero-bordered copies of each image of the batch (row-major and transposed).
size_t fan2d_forward_scratch_bytes(const cmi_plan2d *plan, int vol_dtype, uint64_t batch) {
    const size_t Wp = (size_t)plan->g.nx + 2 * kPad, Hp = (size_t)plan->g.ny + 2 * kPad;
    return 2 * Wp * Hp * dtype_size(vol_dtype) * batch;
}

// Stages the batch into `scratch` (at least fan2d_forward_scratch_bytes bytes of device memory
// on the plan's device, which is current) and launches the march, on `s`. Every staged element,
// border included, is written, so the scratch needs no clearing between calls.
cmi_status fan2d_forward_enqueue(const cmi_plan2d *plan, const void *volume, int vol_dtype,
                                 void *sinogram, int sino_dtype, uint64_t batch, void *scratch,
                                 hipStream_t s) {
    const int nx = (int)plan->g.nx, ny = (int)plan->g.ny, D = (int)plan->g.D, A = (int)plan->A;
    const int Wp = nx + 2 * kPad, Hp = ny + 2 * kPad;
    return dispatch2(vol_dtype, sino_dtype, [&](auto si, auto ss) -> cmi_status {
        using SI = decltype(si);
        using SS = decltype(ss);
        using RI = typename SI::raw;
        dim3 sb(32, 8), sg((unsigned)((Wp + 31) / 32), (unsigned)((Hp + 31) / 32),
                           (unsigned)batch);
        hipLaunchKernelGGL((k_stage2d<SI>), sg, sb, 0, s, (const RI *)volume, (RI *)scratch, nx,
                           ny, Wp, Hp);
        // block_x * block_y (at most kFwdRays) adjacent cells of one angle per workgroup
        const unsigned nt = std::min<unsigned>(plan->block_x * plan->block_y, kFwdRays);
        dim3 fb(nt), fg((unsigned)((D + nt - 1) / nt), (unsigned)A, (unsigned)batch);
        hipLaunchKernelGGL((k_fwd2d<SI, SS>), fg, fb, 0, s, (const RI *)scratch, Wp, Hp,
                           plan->d_ev, plan->d_ns, (typename SS::raw *)sinogram, D, A);
        CMI_HIP_TRY(hipGetLastError());
        return CMI_OK;
    });
}

} // namespace cmi

// ===================================================================== C-ABI ===============

extern "C" {

CMI_API cmi_status cmi_plan2d_create(int device, int vol_prec, const uint64_t vol_shape[2],
                                     const double vol_spacing[2], const double vol_offset[2],
                                     int det_prec, uint64_t det_count, double det_spacing,
                                     double det_offset, const double *DSO, const double *DSD,
                                     const double *angles, uint64_t nangles, double det_pitch,
                                     double COR, cmi_plan2d **plan) {
    clear_error();
    if (!plan) return fail(CMI_ERR_INVALID_ARGUMENT, "plan2d_create: output pointer is null");
    *plan = nullptr;
    if (!vol_shape || !vol_spacing || !vol_offset || !DSO || !DSD || !angles)
        return fail(CMI_ERR_INVALID_ARGUMENT, "plan2d_create: null geometry array");
    CMI_TRY(check_plan_shape(vol_prec, det_prec, vol_shape, vol_spacing, det_count, nangles,
                             "plan2d_create"));
    if (!(det_spacing > 0))
        return fail(CMI_ERR_INVALID_ARGUMENT, "plan2d_create: spacings must be positive");
    for (uint64_t a = 0; a < nangles; ++a)
        if (!std::isfinite(angles[a]) || !std::isfinite(DSO[a]) || !std::isfinite(DSD[a]))
            return fail(CMI_ERR_INVALID_ARGUMENT, "plan2d_create: non-finite geometry at angle %llu",
                        (unsigned long long)a);

    auto p = std::make_unique<cmi_plan2d>();
    set_plan_shape(*p, device, vol_prec, vol_shape, vol_spacing, vol_offset, det_prec, det_count,
                   nangles);
    p->g.du = det_spacing;
    p->g.det_offset = det_offset;
    p->pitch = det_pitch;
    p->cor = COR;

    const int64_t A = p->A;
    std::vector<AngleGeom> geo(A);
    std::vector<float4> meta(A);
    for (int64_t a = 0; a < A; ++a) {
        geo[a] = fan2d::plan_angle(p->g, angles[a], DSO[a], DSD[a]);
        meta[a] = bp2d_meta(angles[a], DSO[a], DSD[a]);
    }
    p->quarter = quarter_turns(p->g, angles, DSO, DSD, A);
    return finish_plan2d(std::move(p), geo, meta, {}, plan);
}

CMI_API cmi_status cmi_plan2d_create_vec(int device, int vol_prec, const uint64_t vol_shape[2],
                                         const double vol_spacing[2], const double vol_offset[2],
                                         int det_prec, uint64_t det_count, const double *vecs,
                                         uint64_t nangles, cmi_plan2d **plan) {
    clear_error();
    if (!plan) return fail(CMI_ERR_INVALID_ARGUMENT, "plan2d_create_vec: output pointer is null");
    *plan = nullptr;
    if (!vol_shape || !vol_spacing || !vol_offset || !vecs)
        return fail(CMI_ERR_INVALID_ARGUMENT, "plan2d_create_vec: null geometry array");
    return plan2d_from_views(device, vol_prec, vol_shape, vol_spacing, vol_offset, det_prec,
                             det_count, vecs, nangles, "plan2d_create_vec", plan);
}

CMI_API cmi_status cmi_plan2d_create_ex(int device, int vol_prec, const uint64_t vol_shape[2],
                                        const double vol_spacing[2], const double vol_offset[2],
                                        int det_prec, uint64_t det_count, double det_spacing,
                                        double det_offset, const double *DSO, const double *DSD,
                                        const double *angles, uint64_t nangles,
                                        double det_pitch, double COR, uint32_t flags,
                                        cmi_plan2d **plan) {
    if ((flags & ~(uint32_t)CMI_PLAN2D_EXACT_GEOMETRY) != 0) {
        clear_error();
        if (plan) *plan = nullptr;
        return fail(CMI_ERR_INVALID_ARGUMENT, "plan2d_create_ex: unknown flags 0x%x", flags);
    }
    if (!(flags & CMI_PLAN2D_EXACT_GEOMETRY))
        return cmi_plan2d_create(device, vol_prec, vol_shape, vol_spacing, vol_offset, det_prec,
                                 det_count, det_spacing, det_offset, DSO, DSD, angles, nangles,
                                 det_pitch, COR, plan);
    clear_error();
    if (!plan) return fail(CMI_ERR_INVALID_ARGUMENT, "plan2d_create_ex: output pointer is null");
    *plan = nullptr;
    if (!vol_shape || !vol_spacing || !vol_offset || !DSO || !DSD || !angles)
        return fail(CMI_ERR_INVALID_ARGUMENT, "plan2d_create_ex: null geometry array");
    CMI_TRY(check_plan_shape(vol_prec, det_prec, vol_shape, vol_spacing, det_count, nangles,
                             "plan2d_create_ex"));
    if (!(det_spacing > 0))
        return fail(CMI_ERR_INVALID_ARGUMENT, "plan2d_create_ex: spacings must be positive");
    if (!std::isfinite(det_offset) || !std::isfinite(det_pitch) || !std::isfinite(COR))
        return fail(CMI_ERR_INVALID_ARGUMENT, "plan2d_create_ex: non-finite detector geometry");
    fan2d::Scalars g{};
    g.du = det_spacing;
    g.det_offset = det_offset;
    std::vector<double> vecs(6 * nangles);
    for (uint64_t a = 0; a < nangles; ++a) {
        if (!std::isfinite(angles[a]) || !std::isfinite(DSO[a]) || !std::isfinite(DSD[a]))
            return fail(CMI_ERR_INVALID_ARGUMENT,
                        "plan2d_create_ex: non-finite geometry at angle %llu",
                        (unsigned long long)a);
        fan2d::exact_view(g, angles[a], DSO[a], DSD[a], det_pitch, COR, vecs.data() + 6 * a);
    }
    return plan2d_from_views(device, vol_prec, vol_shape, vol_spacing, vol_offset, det_prec,
                             det_count, vecs.data(), nangles, "plan2d_create_ex", plan);
}

CMI_API cmi_status cmi_fan2d_bp_table(uint64_t nangles, const double *angles, const double *DSO,
                                      const double *DSD, float *table) {
    clear_error();
    if (!angles || !DSO || !DSD || !table)
        return fail(CMI_ERR_INVALID_ARGUMENT, "fan2d_bp_table: null argument");
    for (uint64_t a = 0; a < nangles; ++a) {
        const float4 m = bp2d_meta(angles[a], DSO[a], DSD[a]);
        table[4 * a + 0] = m.x;
        table[4 * a + 1] = m.y;
        table[4 * a + 2] = m.z;
        table[4 * a + 3] = m.w;
    }
    return CMI_OK;
}

CMI_API cmi_status cmi_plan2d_destroy(cmi_plan2d *plan) {
    clear_error();
    if (!plan) return CMI_OK;
    DeviceGuard guard;
    const cmi_status st = guard.enter(plan->device);
    if (st == CMI_OK) {
        // Work enqueued with this plan may still be running on any stream.
        (void)hipDeviceSynchronize();
        (void)hipFree(plan->d_geo);
        (void)hipFree(plan->d_meta);
        (void)hipFree(plan->d_bpv);
        (void)hipFree(plan->d_ageo);
        (void)hipFree(plan->d_ev);
        (void)hipFree(plan->d_ns);
    }
    delete plan;
    return st;
}

CMI_API cmi_status cmi_plan2d_set_block(cmi_plan2d *plan, uint32_t block_x, uint32_t block_y) {
    clear_error();
    if (!plan) return fail(CMI_ERR_INVALID_ARGUMENT, "plan2d_set_block: plan is null");
    // The forward kernel is compiled for at most kMaxBlock threads per workgroup
    // (__launch_bounds__); a larger workgroup must never reach the launch.
    if (block_x == 0 || block_x % 64 != 0 || block_y == 0 || block_x * block_y > kMaxBlock)
        return fail(CMI_ERR_INVALID_ARGUMENT,
                    "plan2d_set_block: need block_x a multiple of 64 and block_x*block_y <= "
                    "%d (got %u x %u)",
                    kMaxBlock, block_x, block_y);
    plan->block_x = block_x;
    plan->block_y = block_y;
    return CMI_OK;
}

CMI_API cmi_status cmi_plan2d_get_block(const cmi_plan2d *plan, uint32_t *block_x,
                                        uint32_t *block_y) {
    clear_error();
    if (!plan || !block_x || !block_y)
        return fail(CMI_ERR_INVALID_ARGUMENT, "plan2d_get_block: null argument");
    *block_x = plan->block_x;
    *block_y = plan->block_y;
    return CMI_OK;
}

CMI_API cmi_status cmi_plan2d_get_info(const cmi_plan2d *plan, cmi_plan2d_info *info) {
    clear_error();
    if (!plan || !info) return fail(CMI_ERR_INVALID_ARGUMENT, "plan2d_get_info: null argument");
    info->device = plan->device;
    info->vol_prec = plan->vol_prec;
    info->det_prec = plan->det_prec;
    info->nx = (uint64_t)plan->g.nx;
    info->ny = (uint64_t)plan->g.ny;
    info->det_count = (uint64_t)plan->g.D;
    info->nangles = (uint64_t)plan->A;
    info->total_samples = plan->total_samples;
    info->table_bytes = plan->table_bytes;
    return CMI_OK;
}

CMI_API cmi_status cmi_forward_2d(const cmi_plan2d *plan, const void *volume, int vol_dtype,
                                  void *sinogram, int sino_dtype, uint64_t batch,
                                  void *stream) {
    clear_error();
    CMI_TRY(fan2d_forward_check(plan, volume, vol_dtype, sinogram, sino_dtype, batch));
    DeviceGuard guard;
    CMI_TRY(guard.enter(plan->device));
    const hipStream_t s = (hipStream_t)stream;
    void *scratch = nullptr;
    CMI_HIP_TRY(hipMallocAsync(&scratch, fan2d_forward_scratch_bytes(plan, vol_dtype, batch), s));
    const cmi_status st =
        fan2d_forward_enqueue(plan, volume, vol_dtype, sinogram, sino_dtype, batch, scratch, s);
    const hipError_t fe = hipFreeAsync(scratch, s);
    if (st != CMI_OK) return st;
    CMI_HIP_TRY(fe);
    return CMI_OK;
}

CMI_API cmi_status cmi_backward_2d_weighted(const cmi_plan2d *plan, void *volume, int vol_dtype,
                                            const void *sinogram, int sino_dtype,
                                            uint64_t batch, int weighting, void *stream) {
    clear_error();
    CMI_TRY(check_call(plan, volume, vol_dtype, sinogram, sino_dtype, batch, "backward_2d"));
    if (weighting != CMI_BP2D_REFERENCE && weighting != CMI_BP2D_FBP)
        return fail(CMI_ERR_INVALID_ARGUMENT, "backward_2d: unknown weighting %d", weighting);
    DeviceGuard guard;
    CMI_TRY(guard.enter(plan->device));
    const hipStream_t s = (hipStream_t)stream;
    const int nx = (int)plan->g.nx, ny = (int)plan->g.ny, D = (int)plan->g.D, A = (int)plan->A;
    const int Dp = D + 2 * kPad;
    void *scratch = nullptr;
    CMI_HIP_TRY(hipMallocAsync(&scratch, (size_t)Dp * A * dtype_size(sino_dtype) * batch, s));
    const cmi_status st = dispatch2(vol_dtype, sino_dtype, [&](auto sv, auto ss) -> cmi_status {
        using SV = decltype(sv);
        using SS = decltype(ss);
        using RS = typename SS::raw;
        dim3 pb(256), pg((unsigned)((Dp + 255) / 256), (unsigned)A, (unsigned)batch);
        hipLaunchKernelGGL((k_stage_sino2d<SS>), pg, pb, 0, s, (const RS *)sinogram,
                           (RS *)scratch, D, A);
        dim3 b(64, 4), g((unsigned)((nx + 63) / 64), (unsigned)((ny + 3) / 4), (unsigned)batch);
        if (plan->d_bpv) {
            auto kern = weighting == CMI_BP2D_FBP ? k_bp2d_vec<SV, SS, 2> : k_bp2d_vec<SV, SS, 1>;
            hipLaunchKernelGGL(kern, g, b, 0, s, (typename SV::raw *)volume, (const RS *)scratch,
                               plan->d_bpv, nx, ny, D, A, (float)plan->g.sx, (float)plan->g.sy,
                               (float)plan->g.ox, (float)plan->g.oy);
        } else if (plan->quarter && weighting == CMI_BP2D_REFERENCE) {
            // (the FBP weighting keeps the per-pixel body, which the serving kernels share, so
            // that a served FBP volume equals the plain one bit for bit)
            const unsigned th = (unsigned)((nx / 2 + kQT - 1) / kQT);
            auto kern = k_bp2d_qt<SV, SS, 1>;
            hipLaunchKernelGGL(kern, dim3(th, th, (unsigned)batch), dim3(kQThreads), 0, s,
                               (typename SV::raw *)volume, (const RS *)scratch, plan->d_meta, nx,
                               D, A, (float)(1.0 / plan->g.du), (float)D / 2.0f);
        } else {
            auto kern = weighting == CMI_BP2D_FBP ? k_bp2d<SV, SS, 2> : k_bp2d<SV, SS, 1>;
            hipLaunchKernelGGL(kern, g, b, 0, s, (typename SV::raw *)volume, (const RS *)scratch,
                               plan->d_meta, nx, ny, D, A, (float)plan->g.sx, (float)plan->g.sy,
                               (float)plan->g.ox, (float)plan->g.oy, (float)(1.0 / plan->g.du),
                               (float)D / 2.0f);
        }
        CMI_HIP_TRY(hipGetLastError());
        return CMI_OK;
    });
    const hipError_t fe = hipFreeAsync(scratch, s);
    if (st != CMI_OK) return st;
    CMI_HIP_TRY(fe);
    return CMI_OK;
}

CMI_API cmi_status cmi_backward_2d(const cmi_plan2d *plan, void *volume, int vol_dtype,
                                   const void *sinogram, int sino_dtype, uint64_t batch,
                                   void *stream) {
    return cmi_backward_2d_weighted(plan, volume, vol_dtype, sinogram, sino_dtype, batch,
                                    CMI_BP2D_REFERENCE, stream);
}

CMI_API cmi_status cmi_backward_2d_transpose(const cmi_plan2d *plan, void *sinogram,
                                             int sino_dtype, const void *volume, int vol_dtype,
                                             uint64_t batch, int weighting, void *stream) {
    clear_error();
    CMI_TRY(check_call(plan, volume, vol_dtype, sinogram, sino_dtype, batch,
                       "backward_2d_transpose"));
    if (weighting != CMI_BP2D_REFERENCE && weighting != CMI_BP2D_FBP)
        return fail(CMI_ERR_INVALID_ARGUMENT, "backward_2d_transpose: unknown weighting %d",
                    weighting);
    const double bound = plan->bt_bound[weighting == CMI_BP2D_FBP ? 1 : 0];
    if (!std::isfinite(bound))
        return fail(CMI_ERR_NOT_SUPPORTED,
                    "backward_2d_transpose: a source lies inside the image's bounding box, so "
                    "the backprojection's weight is unbounded");
    DeviceGuard guard;
    CMI_TRY(guard.enter(plan->device));
    const hipStream_t s = (hipStream_t)stream;
    const int nx = (int)plan->g.nx, ny = (int)plan->g.ny, D = (int)plan->g.D, A = (int)plan->A;
    const size_t img = (size_t)nx * ny;
    const int Wp = nx + 2 * kPad, Hp = ny + 2 * kPad;
    const uint64_t nseg = (uint64_t)(D + kGS - 1) / kGS;
    if ((uint64_t)A * nseg * kMaxBlock >= (1ull << 32) || batch > 65535)
        return fail(CMI_ERR_INVALID_ARGUMENT,
                    "backward_2d_transpose: too many projections, cells or images for one "
                    "launch");
    // scratch: per image its zero-bordered row-major and transposed copies, then kAdjParts
    // partial maxima per image
    const size_t copy_bytes =
        (2 * (size_t)Wp * Hp * dtype_size(vol_dtype) * batch + 255) / 256 * 256;
    void *scratch = nullptr;
    CMI_HIP_TRY(hipMallocAsync(&scratch, copy_bytes + kAdjParts * batch * sizeof(float), s));
    float *parts = (float *)((char *)scratch + copy_bytes);
    const cmi_status st = dispatch2(vol_dtype, sino_dtype, [&](auto sv, auto ss) -> cmi_status {
        using SV = decltype(sv);
        using SS = decltype(ss);
        using RV = typename SV::raw;
        dim3 sb(32, 8), sg((unsigned)((Wp + 31) / 32), (unsigned)((Hp + 31) / 32),
                           (unsigned)batch);
        hipLaunchKernelGGL((k_stage2d<SV>), sg, sb, 0, s, (const RV *)volume, (RV *)scratch, nx,
                           ny, Wp, Hp);
        hipLaunchKernelGGL((k_absmax_img2d<SV>), dim3(kAdjParts, (unsigned)batch),
                           dim3(kMaxBlock), 0, s, (const RV *)volume, img, parts);
        const bool views = plan->d_bpv != nullptr;
        const bool fbp = weighting == CMI_BP2D_FBP;
        auto kern = views ? (fbp ? k_bpT2d<SV, SS, 2, true> : k_bpT2d<SV, SS, 1, true>)
                          : (fbp ? k_bpT2d<SV, SS, 2, false> : k_bpT2d<SV, SS, 1, false>);
        // projections fastest: neighbouring workgroups gather overlapping wedges
        hipLaunchKernelGGL(kern, dim3((unsigned)(A * nseg), (unsigned)batch), dim3(kMaxBlock),
                           0, s, (typename SS::raw *)sinogram, (const RV *)scratch,
                           views ? plan->d_bpv : plan->d_meta, parts, (float)bound, nx, ny, D,
                           A, (float)plan->g.sx, (float)plan->g.sy, (float)plan->g.ox,
                           (float)plan->g.oy, (float)(1.0 / plan->g.du), (float)D / 2.0f);
        CMI_HIP_TRY(hipGetLastError());
        return CMI_OK;
    });
    const hipError_t fe = hipFreeAsync(scratch, s);
    if (st != CMI_OK) return st;
    CMI_HIP_TRY(fe);
    return CMI_OK;
}

CMI_API cmi_status cmi_adjoint_2d(const cmi_plan2d *plan, void *volume, int vol_dtype,
                                  const void *sinogram, int sino_dtype, uint64_t batch,
                                  void *stream) {
    clear_error();
    CMI_TRY(check_call(plan, volume, vol_dtype, sinogram, sino_dtype, batch, "adjoint_2d"));
    DeviceGuard guard;
    CMI_TRY(guard.enter(plan->device));
    const hipStream_t s = (hipStream_t)stream;
    const int nx = (int)plan->g.nx, ny = (int)plan->g.ny, D = (int)plan->g.D, A = (int)plan->A;
    const size_t img = (size_t)nx * ny * batch, rays = (size_t)A * D;
    // scratch: the 64-bit sum image, then kAdjParts partial maxima
    void *scratch = nullptr;
    CMI_HIP_TRY(hipMallocAsync(&scratch, img * 8 + kAdjParts * batch * sizeof(float), s));
    unsigned long long *accimg = (unsigned long long *)scratch;
    float *parts = (float *)(accimg + img);
    const cmi_status st = dispatch2(vol_dtype, sino_dtype, [&](auto sv, auto ss) -> cmi_status {
        using SV = decltype(sv);
        using SS = decltype(ss);
        const unsigned nb = (unsigned)((img + kMaxBlock - 1) / kMaxBlock);
        hipLaunchKernelGGL(k_fill_u64, dim3(nb), dim3(kMaxBlock), 0, s, accimg, img);
        hipLaunchKernelGGL((k_adj2d_absmax<SS>), dim3(kAdjParts, (unsigned)batch),
                           dim3(kMaxBlock), 0, s, (const typename SS::raw *)sinogram, plan->d_ns,
                           rays, parts);
        // angle groups: enough workgroups to fill the GPU many times over (about 96 per CU; at
        // 1024^2 a target of 8192 workgroups gave 1.09 ms, 16384: 1.06, 24576: 1.04, 32768: 1.05)
        const unsigned tx = (unsigned)((nx + kT2 - 1) / kT2), ty = (unsigned)((ny + kT2 - 1) / kT2);
        const uint64_t tiles = (uint64_t)tx * ty * batch;
        const int G = (int)std::max<uint64_t>(
            1, std::min<uint64_t>((uint64_t)A, (24576 + tiles - 1) / tiles));
        hipLaunchKernelGGL((k_adj2d_tile<SS>), dim3(tx, ty, (unsigned)(G * batch)),
                           dim3(kMaxBlock), 0, s, accimg, (const typename SS::raw *)sinogram,
                           plan->d_ageo, plan->d_ev, plan->d_ns, parts, nx, ny, D, A, G);
        const size_t one = (size_t)nx * ny;
        hipLaunchKernelGGL((k_adj2d_finish<SV>),
                           dim3((unsigned)((one + kMaxBlock - 1) / kMaxBlock), (unsigned)batch),
                           dim3(kMaxBlock), 0, s, (typename SV::raw *)volume, accimg, parts, one);
        CMI_HIP_TRY(hipGetLastError());
        return CMI_OK;
    });
    const hipError_t fe = hipFreeAsync(scratch, s);
    if (st != CMI_OK) return st;
    CMI_HIP_TRY(fe);
    return CMI_OK;
}

} // extern "C"
