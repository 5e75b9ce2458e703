// Serving backprojector for gfx950: one persistent launch over the tiles of many jobs of
// mixed geometry and size, and the staging kernel that writes filtered projections into the
// zero-bordered layout it reads.
//
// Work distribution
//   The host lists the jobs of a launch with a tile range each; the fan-beam and the
//   cone-beam jobs go to two persistent kernels on the same stream (fan-beam tiles need no
//   LDS and few registers, so their kernel keeps more workgroups resident). Within a kernel,
//   job j owns launch tiles [first_j, first_j + (tile_end_j - tile_begin_j)). The grid is
//   sized to what the device holds resident of that kernel; each workgroup takes the next
//   launch tile from the kernel's atomic counter, finds its job by binary search over
//   `first`, backprojects that tile and repeats until the counter passes the total. Every
//   workgroup exits after at most total + grid iterations of the counter, so the launch
//   drains.
//
// Per-tile arithmetic is the plain backprojectors' own (bp2d_pixel / bp3d_column with the
// same scalar inputs), so a voxel's value does not depend on the launch it ran in.
//
// Data layout in HBM (per server)
//   job tables  [kRing][max_jobs] DevJob      one slot per launch in flight
//   counters    [kRing][2] uint32             tile counters of each slot (fan, cone)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "bp_tiles.hpp"
#include "stage_index.hpp"
#include "cmi_internal.hpp"

using namespace cmi;

namespace {

constexpr int kRing = 4;       // launches whose job tables may be in flight at once
constexpr int kThreads = 256;  // threads per workgroup
constexpr int kTile2X = 64, kTile2Y = 4;   // fan-beam tile: pixels
constexpr int kTile3X = 16, kTile3Y = 16;  // cone-beam tile: columns of kZB voxels
// cone-beam tile of y runs: kTileX x kYB x kTileZ voxels (bp_tiles.hpp)

// Device copy of a job with the derived values the kernel needs.
struct DevJob {
    int32_t kind, mode, columns;
    const void *sino;
    const float *table;
    void *volume;
    int32_t nx, ny, nz, nu, nv, nangles;
    int32_t tiles_x, tiles_y;  // tiles along x / along y of the volume
    uint32_t first;            // first launch tile of this job
    uint32_t tile_begin;       // first tile of the job's own numbering
    uint32_t sino_bytes;       // size of the staged projections, for the range check
    float sx, sy, ox, oy, inv_du, half_d; // fan beam
    float cx, cy, cz;                     // cone beam: centre of the volume in index units
};

uint64_t tiles_of(const cmi_bp_job &j) {
    if (j.kind == CMI_JOB_FAN2D)
        return (uint64_t)((j.nx + kTile2X - 1) / kTile2X) * ((j.ny + kTile2Y - 1) / kTile2Y);
    if (j.columns == 1)
        return (uint64_t)((j.nx + kTileX - 1) / kTileX) * ((j.ny + kYB - 1) / kYB) *
               ((j.nz + kTileZ - 1) / kTileZ);
    return (uint64_t)((j.nx + kTile3X - 1) / kTile3X) * ((j.ny + kTile3Y - 1) / kTile3Y) *
           ((j.nz + kZB - 1) / kZB);
}

// ---------------------------------------------------------------- staging -----------------

// dst[a][vp][up] = scale * row_scale[a] * src[a][vp - vpad][up - 2], 0 in the border. Flat
// over the n destination elements, four consecutive ones per thread; the source row is
// followed element by element. VEC: the four are written with one vector store, which needs
// dst aligned to four elements (the host checks; otherwise element stores).
template <class S, bool VEC>
__global__ __launch_bounds__(kThreads) void k_stage_projections(
    const float *__restrict__ src, uint64_t stride, int V, int U, int vpad,
    const float *__restrict__ row_scale, float scale, uint64_t n,
    typename S::raw *__restrict__ dst) {
    using raw = typename S::raw;
    struct alignas(4 * sizeof(raw)) Four {
        raw v[4];
    };
    const uint64_t i0 = ((uint64_t)blockIdx.x * kThreads + threadIdx.x) * 4;
    if (i0 >= n) return;
    const int Up = U + 2 * kPadDet, Vp = V + 2 * vpad;
    StageIndex at = stage_index(i0, Up, Vp, n <= 0xffffffffull);
    Four f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int u = at.up - kPadDet, v = at.vp - vpad;
        float val = 0.0f;
        if (i0 + k < n && u >= 0 && u < U && v >= 0 && v < V) {
            val = src[(at.a * V + v) * stride + u] * scale;
            if (row_scale) val *= row_scale[at.a];
        }
        f.v[k] = from_f32<S>(val);
        stage_next(at, Up, Vp);
    }
    if (VEC && i0 + 4 <= n) {
        *reinterpret_cast<Four *>(dst + i0) = f;
    } else {
        for (uint64_t k = 0; k < 4 && i0 + k < n; ++k) dst[i0 + k] = f.v[k];
    }
}

// Ramp filtering of projection rows through complex FFTs of row pairs (the filter's response
// is real and even, so the real and imaginary parts of a filtered pair are the filtered rows).
// Pack: complex row r of length L holds rows 2r and 2r + 1 of `src` times their cosine-weight
// rows (row i uses weight row i mod cw_rows), zero past D; an odd last row pairs with zeros.
__global__ __launch_bounds__(kThreads) void k_filter_pack(const float *__restrict__ src,
                                                          uint64_t rows, int D,
                                                          const float *__restrict__ cw,
                                                          uint64_t cw_rows, int L,
                                                          float2 *__restrict__ dst) {
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= L) return;
    const uint64_t pairs = (rows + 1) / 2;
    for (uint64_t r = blockIdx.y; r < pairs; r += gridDim.y) {
        const uint64_t i0 = 2 * r, i1 = i0 + 1;
        float a = 0.0f, b = 0.0f;
        if (k < D) {
            a = src[i0 * D + k] * cw[(i0 % cw_rows) * D + k];
            if (i1 < rows) b = src[i1 * D + k] * cw[(i1 % cw_rows) * D + k];
        }
        dst[r * L + k] = make_float2(a, b);
    }
}

// Unpack: dst row 2r (2r + 1) = the real (imaginary) part of the first D entries of row r.
__global__ __launch_bounds__(kThreads) void k_filter_unpack(const float2 *__restrict__ src,
                                                            uint64_t rows, int D, int L,
                                                            float *__restrict__ dst) {
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= D) return;
    const uint64_t pairs = (rows + 1) / 2;
    for (uint64_t r = blockIdx.y; r < pairs; r += gridDim.y) {
        const float2 v = src[r * L + k];
        dst[2 * r * D + k] = v.x;
        if (2 * r + 1 < rows) dst[(2 * r + 1) * D + k] = v.y;
    }
}

__global__ __launch_bounds__(64) void k_zero_counter2(unsigned *__restrict__ c) {
    if (threadIdx.x < 2) c[threadIdx.x] = 0u;
}

// ---------------------------------------------------------------- tiles -------------------

// Fan-beam tile: pixel (ix, iy) as k_bp2d computes it.
template <class SV, class SS>
__device__ __forceinline__ void serve_tile_2d(const DevJob &J, uint32_t t) {
    const int tx = (int)(t % (uint32_t)J.tiles_x), ty = (int)(t / (uint32_t)J.tiles_x);
    const int ix = tx * kTile2X + (int)(threadIdx.x % kTile2X);
    const int iy = ty * kTile2Y + (int)(threadIdx.x / kTile2X);
    if (ix >= J.nx || iy >= J.ny) return;
    const rsrc_t R = make_rsrc(J.sino, J.sino_bytes);
    const ctab meta = const_table(J.table);
    const float x = ((float)ix - (float)J.nx / 2.0f) * J.sx + J.ox + 0.5f;
    const float y = ((float)iy - (float)J.ny / 2.0f) * J.sy + J.oy + 0.5f;
    const float px = x * J.inv_du, py = y * J.inv_du;
    const float acc = J.mode == CMI_BP2D_FBP
                          ? bp2d_pixel<SS, 2>(R, meta, J.nangles, J.nu, x, y, px, py, J.half_d)
                          : bp2d_pixel<SS, 1>(R, meta, J.nangles, J.nu, x, y, px, py, J.half_d);
    auto *vol = static_cast<typename SV::raw *>(J.volume);
    const size_t i = (size_t)iy * J.nx + ix;
    vol[i] = from_f32<SV>(to_f32<SV>(vol[i]) + acc * J.inv_du);
}

// Cone-beam tile: the column of kZB voxels k_bp3d gives thread (ix, iy) of z block zb.
template <class SV, class SS>
__device__ __forceinline__ void serve_tile_3d(const DevJob &J, uint32_t t) {
    const uint32_t per_slab = (uint32_t)J.tiles_x * (uint32_t)J.tiles_y;
    const int zb = (int)(t / per_slab);
    const uint32_t r = t % per_slab;
    const int ix = (int)(r % (uint32_t)J.tiles_x) * kTile3X + (int)(threadIdx.x % kTile3X);
    const int iy = (int)(r / (uint32_t)J.tiles_x) * kTile3Y + (int)(threadIdx.x / kTile3X);
    if (ix >= J.nx || iy >= J.ny) return;
    const int iz0 = zb * kZB;
    const rsrc_t R = make_rsrc(J.sino, J.sino_bytes);
    float acc[kZB];
#pragma unroll
    for (int z = 0; z < kZB; ++z) acc[z] = 0.0f;
    const float hx = (float)ix - J.cx, hy = (float)iy - J.cy, hz = (float)iz0 - J.cz;
    if (J.mode == CMI_BP3D_FDK)
        bp3d_column<SS, true, true>(R, const_table(J.table), J.nangles, J.nu, J.nv, hx, hy, hz, acc);
    else if (J.mode == CMI_BP3D_MATCHED)
        bp3d_column<SS, true, false>(R, const_table(J.table), J.nangles, J.nu, J.nv, hx, hy, hz, acc);
    else
        bp3d_column<SS, false, false>(R, const_table(J.table), J.nangles, J.nu, J.nv, hx, hy, hz, acc);
    auto *vol = static_cast<typename SV::raw *>(J.volume);
#pragma unroll
    for (int z = 0; z < kZB; ++z) {
        const int iz = iz0 + z;
        if (iz < J.nz) {
            const size_t i = ((size_t)iz * J.ny + iy) * (size_t)J.nx + ix;
            vol[i] = from_f32<SV>(to_f32<SV>(vol[i]) + acc[z]);
        }
    }
}

// Cone-beam tile of y runs: the tile and runs k_bp3d_y gives block (tx, ty, tz). Every thread
// of the workgroup calls this (the LDS staging inside synchronises the workgroup).
template <class SV, class SS>
__device__ __forceinline__ void serve_tile_3d_y(const DevJob &J, uint32_t t) {
    const uint32_t per_slab = (uint32_t)J.tiles_x * (uint32_t)J.tiles_y;
    const int zt = (int)(t / per_slab);
    const uint32_t r = t % per_slab;
    const int ix0 = (int)(r % (uint32_t)J.tiles_x) * kTileX;
    const int iy0 = (int)(r / (uint32_t)J.tiles_x) * kYB;
    const int iz0 = zt * kTileZ;
    const int ix = ix0 + (int)(threadIdx.x % kTileX), iz = iz0 + (int)(threadIdx.x / kTileX);
    const rsrc_t R = make_rsrc(J.sino, J.sino_bytes);
    float acc[kYB];
#pragma unroll
    for (int y = 0; y < kYB; ++y) acc[y] = 0.0f;
    const float hx = (float)ix - J.cx, hy = (float)iy0 - J.cy, hz = (float)iz - J.cz;
    const float tx = (float)ix0 - J.cx, tz = (float)iz0 - J.cz;
    if (J.mode == CMI_BP3D_FDK)
        bp3d_yrun_lds<SS, true, true>(R, const_table(J.table), J.nangles, J.nu, J.nv, hx, hy, hz, tx, hy, tz,
                                      acc);
    else if (J.mode == CMI_BP3D_MATCHED)
        bp3d_yrun_lds<SS, true, false>(R, const_table(J.table), J.nangles, J.nu, J.nv, hx, hy, hz, tx, hy, tz,
                                       acc);
    else
        bp3d_yrun_lds<SS, false, false>(R, const_table(J.table), J.nangles, J.nu, J.nv, hx, hy, hz, tx, hy, tz,
                                        acc);
    if (ix >= J.nx || iz >= J.nz) return;
    auto *vol = static_cast<typename SV::raw *>(J.volume);
#pragma unroll
    for (int y = 0; y < kYB; ++y) {
        const int iy = iy0 + y;
        if (iy < J.ny) {
            const size_t i = ((size_t)iz * J.ny + iy) * (size_t)J.nx + ix;
            vol[i] = from_f32<SV>(to_f32<SV>(vol[i]) + acc[y]);
        }
    }
}

// Next launch tile from the counter, the same for every thread of the workgroup.
__device__ __forceinline__ uint32_t next_tile(unsigned *counter, unsigned *s_tile) {
    __syncthreads(); // every thread has read the previous tile
    if (threadIdx.x == 0) *s_tile = atomicAdd(counter, 1u);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*s_tile);
}

// The job owning launch tile t: the last with first <= t.
__device__ __forceinline__ const DevJob &job_of(const DevJob *__restrict__ jobs, int njobs,
                                               uint32_t t) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first <= t) lo = mid;
        else hi = mid - 1;
    }
    return jobs[lo];
}

// Fan-beam tiles. A kernel of its own: it needs no LDS and few registers, so more workgroups
// stay resident than for the cone-beam tiles (the 2D body is bound by issue and latency).
template <class SV, class SS>
__global__ __launch_bounds__(kThreads) void k_serve_bp2d(const DevJob *__restrict__ jobs,
                                                         int njobs, uint32_t total,
                                                         unsigned *__restrict__ counter) {
    __shared__ unsigned s_tile;
    for (;;) {
        const uint32_t t = next_tile(counter, &s_tile);
        if (t >= total) return;
        const DevJob &J = job_of(jobs, njobs, t);
        serve_tile_2d<SV, SS>(J, t - J.first + J.tile_begin);
    }
}

// Cone-beam tiles. YRUNS: every job of the launch takes the y-run body (columns == 1), and the
// general column body is left out. Five wavefronts per SIMD: the persistent loop around the
// body's three modes took 97 VGPRs (four wavefronts per SIMD); 96 fit without spills, 80 spilled.
template <class SV, class SS, bool YRUNS>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_serve_bp3d(const DevJob *__restrict__ jobs,
                                                         int njobs, uint32_t total,
                                                         unsigned *__restrict__ counter) {
    __shared__ unsigned s_tile;
    for (;;) {
        const uint32_t t = next_tile(counter, &s_tile);
        if (t >= total) return;
        const DevJob &J = job_of(jobs, njobs, t);
        const uint32_t lt = t - J.first + J.tile_begin;
        if (YRUNS || J.columns == 1) serve_tile_3d_y<SV, SS>(J, lt);
        else serve_tile_3d<SV, SS>(J, lt);
    }
}

template <class F>
cmi_status dispatch2(int vdt, int sdt, F &&f) {
    auto with_s = [&](auto sv) -> cmi_status {
        switch (sdt) {
        case CMI_F32: return f(sv, F32{});
        case CMI_F16: return f(sv, F16{});
        case CMI_BF16: return f(sv, BF16{});
        }
        return fail(CMI_ERR_DTYPE, "unsupported sinogram dtype %d", sdt);
    };
    switch (vdt) {
    case CMI_F32: return with_s(F32{});
    case CMI_F16: return with_s(F16{});
    case CMI_BF16: return with_s(BF16{});
    }
    return fail(CMI_ERR_DTYPE, "unsupported volume dtype %d", vdt);
}

cmi_status check_job(const cmi_bp_job &j, uint32_t i, int sino_dtype) {
    if (j.kind != CMI_JOB_FAN2D && j.kind != CMI_JOB_CONE3D)
        return fail(CMI_ERR_INVALID_ARGUMENT, "serve: job %u has unknown kind %d", i, j.kind);
    if (!j.sino || !j.table || !j.volume)
        return fail(CMI_ERR_INVALID_ARGUMENT, "serve: job %u has a null pointer", i);
    const bool fan = j.kind == CMI_JOB_FAN2D;
    if (fan ? (j.mode != CMI_BP2D_REFERENCE && j.mode != CMI_BP2D_FBP)
            : (j.mode != CMI_BP3D_REFERENCE && j.mode != CMI_BP3D_MATCHED &&
               j.mode != CMI_BP3D_FDK))
        return fail(CMI_ERR_INVALID_ARGUMENT, "serve: job %u has unknown mode %d", i, j.mode);
    if (j.nx == 0 || j.ny == 0 || j.nz == 0 || j.nu == 0 || j.nv == 0 || j.nangles == 0)
        return fail(CMI_ERR_INVALID_ARGUMENT, "serve: job %u has an empty dimension", i);
    if (j.nx > 16384 || j.ny > 16384 || j.nz > 16384 || j.nu > 65536 || j.nv > 65536 ||
        j.nangles > 65535)
        return fail(CMI_ERR_INVALID_ARGUMENT, "serve: job %u exceeds the size limits", i);
    if (!fan && j.columns != 0 && j.columns != 1)
        return fail(CMI_ERR_INVALID_ARGUMENT, "serve: job %u has unknown columns %d", i,
                    j.columns);
    if (fan && (j.nz != 1 || j.nv != 1))
        return fail(CMI_ERR_INVALID_ARGUMENT, "serve: fan-beam job %u needs nz = nv = 1", i);
    if (fan && !(std::isfinite(j.sx) && std::isfinite(j.sy) && std::isfinite(j.ox) &&
                 std::isfinite(j.oy) && std::isfinite(j.inv_du) && j.inv_du > 0.0f))
        return fail(CMI_ERR_INVALID_ARGUMENT, "serve: fan-beam job %u has bad spacing", i);
    const uint64_t elems = (uint64_t)j.nangles * (j.nu + 2 * kPadDet) *
                           (fan ? 1 : (uint64_t)j.nv + 2 * kPadDet);
    if (elems * dtype_size(sino_dtype) >= (1ull << 31))
        return fail(CMI_ERR_INVALID_ARGUMENT, "serve: job %u projections exceed 2 GiB", i);
    const uint64_t vox = (uint64_t)j.nx * j.ny * j.nz;
    if (vox >= (1ull << 31))
        return fail(CMI_ERR_INVALID_ARGUMENT, "serve: job %u volume exceeds 2^31 voxels", i);
    const uint64_t nt = tiles_of(j);
    if (j.tile_begin >= j.tile_end || j.tile_end > nt)
        return fail(CMI_ERR_INVALID_ARGUMENT,
                    "serve: job %u tile range [%u, %u) invalid for %llu tiles", i, j.tile_begin,
                    j.tile_end, (unsigned long long)nt);
    return CMI_OK;
}

} // namespace

struct cmi_server {
    int device = 0;
    uint32_t max_jobs = 0;
    uint32_t grid = 0;   // resident workgroups of the cone-beam serving kernel (any job)
    uint32_t grid3y = 0; // resident workgroups of the cone-beam serving kernel (y runs only)
    uint32_t grid2d = 0; // resident workgroups of the fan-beam serving kernel
    DevJob *d_jobs = nullptr;
    unsigned *d_counter = nullptr;
    DevJob *h_jobs = nullptr; // pinned
    hipEvent_t done[kRing] = {};
    bool used[kRing] = {};
    int next = 0;
    int cus = 0;
};

extern "C" {

CMI_API cmi_status cmi_stage_projections(int device, const float *src, uint64_t src_stride,
                                         uint64_t A, uint64_t V, uint64_t U, uint32_t vpad,
                                         const float *row_scale, float scale, void *dst,
                                         int dst_dtype, void *stream) {
    clear_error();
    if (!src || !dst) return fail(CMI_ERR_INVALID_ARGUMENT, "stage_projections: null pointer");
    if (!dtype_ok(dst_dtype))
        return fail(CMI_ERR_DTYPE, "stage_projections: unsupported dtype %d", dst_dtype);
    if (A == 0 || V == 0 || U == 0 || src_stride < U || A > 65535 || V > 65535 ||
        U > (1u << 24) || (vpad != 0 && vpad != (uint32_t)kPadDet) || (V > 1 && vpad == 0))
        return fail(CMI_ERR_INVALID_ARGUMENT, "stage_projections: bad shape");
    if (!std::isfinite(scale))
        return fail(CMI_ERR_INVALID_ARGUMENT, "stage_projections: non-finite scale");
    DeviceGuard guard;
    CMI_TRY(guard.enter(device));
    CMI_TRY(check_device_ptr(src, device, "source"));
    CMI_TRY(check_device_ptr(dst, device, "destination"));
    if (row_scale) CMI_TRY(check_device_ptr(row_scale, device, "row_scale"));
    const hipStream_t s = (hipStream_t)stream;
    const int Up = (int)U + 2 * kPadDet, Vp = (int)V + 2 * (int)vpad;
    const uint64_t n = (uint64_t)Up * Vp * A;
    if ((n + 4 * kThreads - 1) / (4 * kThreads) > 0x7fffffffull)
        return fail(CMI_ERR_INVALID_ARGUMENT, "stage_projections: too many elements");
    const dim3 b(kThreads), g((unsigned)((n + 4 * kThreads - 1) / (4 * kThreads)));
    // one vector store per four elements needs dst aligned to four elements
    const bool vec = (uintptr_t)dst % (4 * dtype_size(dst_dtype)) == 0;
    auto launch = [&](auto tag) {
        using S = decltype(tag);
        auto kern = vec ? k_stage_projections<S, true> : k_stage_projections<S, false>;
        hipLaunchKernelGGL(kern, g, b, 0, s, src, src_stride, (int)V, (int)U, (int)vpad,
                           row_scale, scale, n, (typename S::raw *)dst);
    };
    if (dst_dtype == CMI_F32) launch(F32{});
    else if (dst_dtype == CMI_F16) launch(F16{});
    else launch(BF16{});
    CMI_HIP_TRY(hipGetLastError());
    return CMI_OK;
}

CMI_API cmi_status cmi_filter_pack(int device, const float *src, uint64_t rows, uint64_t D,
                                   const float *cw, uint64_t cw_rows, uint64_t L, float *dst,
                                   void *stream) {
    clear_error();
    if (!src || !cw || !dst) return fail(CMI_ERR_INVALID_ARGUMENT, "filter_pack: null pointer");
    if (rows == 0 || D == 0 || cw_rows == 0 || L < D || L > (1u << 24))
        return fail(CMI_ERR_INVALID_ARGUMENT, "filter_pack: bad shape");
    DeviceGuard guard;
    CMI_TRY(guard.enter(device));
    CMI_TRY(check_device_ptr(src, device, "source"));
    CMI_TRY(check_device_ptr(cw, device, "weights"));
    CMI_TRY(check_device_ptr(dst, device, "destination"));
    const uint64_t pairs = (rows + 1) / 2;
    dim3 g((unsigned)((L + kThreads - 1) / kThreads), (unsigned)std::min<uint64_t>(pairs, 65535));
    hipLaunchKernelGGL(k_filter_pack, g, dim3(kThreads), 0, (hipStream_t)stream, src, rows,
                       (int)D, cw, cw_rows, (int)L, reinterpret_cast<float2 *>(dst));
    CMI_HIP_TRY(hipGetLastError());
    return CMI_OK;
}

CMI_API cmi_status cmi_filter_unpack(int device, const float *src, uint64_t rows, uint64_t D,
                                     uint64_t L, float *dst, void *stream) {
    clear_error();
    if (!src || !dst) return fail(CMI_ERR_INVALID_ARGUMENT, "filter_unpack: null pointer");
    if (rows == 0 || D == 0 || L < D || L > (1u << 24))
        return fail(CMI_ERR_INVALID_ARGUMENT, "filter_unpack: bad shape");
    DeviceGuard guard;
    CMI_TRY(guard.enter(device));
    CMI_TRY(check_device_ptr(src, device, "source"));
    CMI_TRY(check_device_ptr(dst, device, "destination"));
    const uint64_t pairs = (rows + 1) / 2;
    dim3 g((unsigned)((D + kThreads - 1) / kThreads), (unsigned)std::min<uint64_t>(pairs, 65535));
    hipLaunchKernelGGL(k_filter_unpack, g, dim3(kThreads), 0, (hipStream_t)stream,
                       reinterpret_cast<const float2 *>(src), rows, (int)D, (int)L, dst);
    CMI_HIP_TRY(hipGetLastError());
    return CMI_OK;
}

CMI_API cmi_status cmi_bp_job_tiles(const cmi_bp_job *job, uint64_t *tiles) {
    clear_error();
    if (!job || !tiles) return fail(CMI_ERR_INVALID_ARGUMENT, "bp_job_tiles: null pointer");
    if (job->kind != CMI_JOB_FAN2D && job->kind != CMI_JOB_CONE3D)
        return fail(CMI_ERR_INVALID_ARGUMENT, "bp_job_tiles: unknown kind %d", job->kind);
    *tiles = tiles_of(*job);
    return CMI_OK;
}

CMI_API cmi_status cmi_server_create(int device, uint32_t max_jobs, cmi_server **server) {
    clear_error();
    if (!server) return fail(CMI_ERR_INVALID_ARGUMENT, "server_create: null output");
    *server = nullptr;
    if (max_jobs == 0 || max_jobs > 4096)
        return fail(CMI_ERR_INVALID_ARGUMENT, "server_create: max_jobs must be in [1, 4096]");
    DeviceGuard guard;
    CMI_TRY(guard.enter(device));
    auto p = std::make_unique<cmi_server>();
    p->device = device;
    p->max_jobs = max_jobs;
    CMI_HIP_TRY(hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, device));
    int per_cu = 0, per_cu3y = 0, per_cu2 = 0;
    CMI_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, k_serve_bp3d<F32, F32, false>, kThreads, 0));
    CMI_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu3y, k_serve_bp3d<F32, F32, true>, kThreads, 0));
    CMI_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu2, k_serve_bp2d<F32, F32>,
                                                             kThreads, 0));
    p->grid = (uint32_t)std::max(1, per_cu) * (uint32_t)p->cus;
    p->grid3y = (uint32_t)std::max(1, per_cu3y) * (uint32_t)p->cus;
    p->grid2d = (uint32_t)std::max(1, per_cu2) * (uint32_t)p->cus;
    struct Cleanup {
        cmi_server *p;
        bool armed = true;
        ~Cleanup() {
            if (!armed) return;
            (void)hipFree(p->d_jobs);
            (void)hipFree(p->d_counter);
            (void)hipHostFree(p->h_jobs);
            for (auto &e : p->done)
                if (e) (void)hipEventDestroy(e);
        }
    } cleanup{p.get()};
    const size_t table = sizeof(DevJob) * max_jobs * kRing;
    CMI_HIP_TRY(hipMalloc(&p->d_jobs, table));
    CMI_HIP_TRY(hipMalloc(&p->d_counter, sizeof(unsigned) * 2 * kRing));
    CMI_HIP_TRY(hipHostMalloc(&p->h_jobs, table, hipHostMallocDefault));
    for (auto &e : p->done) CMI_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    cleanup.armed = false;
    *server = p.release();
    return CMI_OK;
}

CMI_API cmi_status cmi_server_get_capacity(const cmi_server *p, uint32_t *workgroups) {
    clear_error();
    if (!p || !workgroups) return fail(CMI_ERR_INVALID_ARGUMENT, "server_get_capacity: null");
    *workgroups = p->grid;
    return CMI_OK;
}

CMI_API cmi_status cmi_server_destroy(cmi_server *p) {
    clear_error();
    if (!p) return CMI_OK;
    DeviceGuard guard;
    CMI_TRY(guard.enter(p->device));
    for (int i = 0; i < kRing; ++i)
        if (p->used[i]) (void)hipEventSynchronize(p->done[i]);
    (void)hipFree(p->d_jobs);
    (void)hipFree(p->d_counter);
    (void)hipHostFree(p->h_jobs);
    for (auto &e : p->done) (void)hipEventDestroy(e);
    delete p;
    return CMI_OK;
}

CMI_API cmi_status cmi_serve_backproject(cmi_server *p, const cmi_bp_job *jobs, uint32_t njobs,
                                         int vol_dtype, int sino_dtype, void *stream) {
    clear_error();
    if (!p) return fail(CMI_ERR_INVALID_ARGUMENT, "serve: server is null");
    if (!jobs || njobs == 0) return fail(CMI_ERR_INVALID_ARGUMENT, "serve: no jobs");
    if (njobs > p->max_jobs)
        return fail(CMI_ERR_INVALID_ARGUMENT, "serve: %u jobs exceed the server's %u", njobs,
                    p->max_jobs);
    if (!dtype_ok(vol_dtype) || !dtype_ok(sino_dtype))
        return fail(CMI_ERR_DTYPE, "serve: unsupported dtype (%d, %d)", vol_dtype, sino_dtype);
    uint64_t total = 0;
    for (uint32_t i = 0; i < njobs; ++i) {
        CMI_TRY(check_job(jobs[i], i, sino_dtype));
        total += jobs[i].tile_end - jobs[i].tile_begin;
    }
    if (total >= (1ull << 31))
        return fail(CMI_ERR_INVALID_ARGUMENT, "serve: too many tiles in one launch");
    DeviceGuard guard;
    CMI_TRY(guard.enter(p->device));
    for (uint32_t i = 0; i < njobs; ++i) {
        CMI_TRY(check_device_ptr(jobs[i].sino, p->device, "job projections"));
        CMI_TRY(check_device_ptr(jobs[i].table, p->device, "job table"));
        CMI_TRY(check_device_ptr(jobs[i].volume, p->device, "job volume"));
    }
    const hipStream_t s = (hipStream_t)stream;
    const int slot = p->next;
    // the slot's previous launch has finished reading its table and counter
    if (p->used[slot]) CMI_HIP_TRY(hipEventSynchronize(p->done[slot]));
    // the slot's table: fan-beam jobs first, then cone-beam ones, each kind with its own
    // launch-tile numbering (the two kinds run in separate launches)
    DevJob *h = p->h_jobs + (size_t)slot * p->max_jobs;
    uint32_t n2 = 0, total2 = 0, total3 = 0;
    bool yruns = true; // every cone-beam job takes the y-run body
    for (uint32_t i = 0; i < njobs; ++i) {
        n2 += jobs[i].kind == CMI_JOB_FAN2D ? 1u : 0u;
        if (jobs[i].kind != CMI_JOB_FAN2D && jobs[i].columns != 1) yruns = false;
    }
    uint32_t k2 = 0, k3 = n2;
    for (uint32_t i = 0; i < njobs; ++i) {
        const cmi_bp_job &j = jobs[i];
        DevJob d{};
        d.kind = j.kind;
        d.mode = j.mode;
        d.columns = j.kind == CMI_JOB_FAN2D ? 0 : j.columns;
        d.sino = j.sino;
        d.table = static_cast<const float *>(j.table);
        d.volume = j.volume;
        d.nx = (int)j.nx, d.ny = (int)j.ny, d.nz = (int)j.nz;
        d.nu = (int)j.nu, d.nv = (int)j.nv, d.nangles = (int)j.nangles;
        const bool fan = j.kind == CMI_JOB_FAN2D;
        if (fan) {
            d.tiles_x = (d.nx + kTile2X - 1) / kTile2X;
            d.tiles_y = (d.ny + kTile2Y - 1) / kTile2Y;
        } else if (d.columns == 1) {
            d.tiles_x = (d.nx + kTileX - 1) / kTileX;
            d.tiles_y = (d.ny + kYB - 1) / kYB;
        } else {
            d.tiles_x = (d.nx + kTile3X - 1) / kTile3X;
            d.tiles_y = (d.ny + kTile3Y - 1) / kTile3Y;
        }
        uint32_t &first = fan ? total2 : total3;
        d.first = first;
        d.tile_begin = j.tile_begin;
        const uint64_t elems = (uint64_t)j.nangles * (j.nu + 2 * kPadDet) *
                               (fan ? 1 : (uint64_t)j.nv + 2 * kPadDet);
        d.sino_bytes = (uint32_t)(elems * dtype_size(sino_dtype));
        // the plain launchers' scalar arguments, computed the same way
        d.sx = j.sx, d.sy = j.sy, d.ox = j.ox, d.oy = j.oy, d.inv_du = j.inv_du;
        d.half_d = (float)(int)j.nu / 2.0f;
        d.cx = (float)(((int)j.nx - 1) / 2.0);
        d.cy = (float)(((int)j.ny - 1) / 2.0);
        d.cz = (float)(((int)j.nz - 1) / 2.0);
        h[fan ? k2++ : k3++] = d;
        first += j.tile_end - j.tile_begin;
    }
    DevJob *dj = p->d_jobs + (size_t)slot * p->max_jobs;
    unsigned *counter = p->d_counter + 2 * slot;
    CMI_HIP_TRY(hipMemcpyAsync(dj, h, sizeof(DevJob) * njobs, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_zero_counter2, dim3(1), dim3(64), 0, s, counter);
    CMI_HIP_TRY(hipGetLastError());
    const cmi_status st = dispatch2(vol_dtype, sino_dtype, [&](auto sv, auto ss) -> cmi_status {
        using SV = decltype(sv);
        using SS = decltype(ss);
        if (total3) {
            const unsigned grid =
                (unsigned)std::min<uint64_t>(total3, yruns ? p->grid3y : p->grid);
            auto kern = yruns ? k_serve_bp3d<SV, SS, true> : k_serve_bp3d<SV, SS, false>;
            hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), 0, s,
                               (const DevJob *)(dj + n2), (int)(njobs - n2), total3,
                               counter + 1);
            CMI_HIP_TRY(hipGetLastError());
        }
        if (total2) {
            const unsigned grid = (unsigned)std::min<uint64_t>(total2, p->grid2d);
            hipLaunchKernelGGL((k_serve_bp2d<SV, SS>), dim3(grid), dim3(kThreads), 0, s,
                               (const DevJob *)dj, (int)n2, total2, counter);
            CMI_HIP_TRY(hipGetLastError());
        }
        return CMI_OK;
    });
    CMI_TRY(st);
    CMI_HIP_TRY(hipEventRecord(p->done[slot], s));
    p->used[slot] = true;
    p->next = (slot + 1) % kRing;
    return CMI_OK;
}

} // extern "C"
