/*
 * curadon_mi.h -- C-ABI of the MI355X (gfx950) projection operators.
 *
 * This is the drop-in boundary for curadon's native extension `curadon_ext`
 * (python/standalone/src/bind.cpp:58-66). Every entry point takes plain pointers, sizes and
 * an explicit HIP stream (passed as void*), returns a status code, and never calls exit().
 * Device arrays are caller-owned; plans own their device tables and are immutable after
 * creation except for the launch-shape tunables, so one plan may be used from several host
 * threads and streams at once.
 *
 * Layouts are the reference's (python/standalone/src/forward.cu:47-58, 85-87):
 *   2D image     (ny, nx)            x fastest          -- batch of B images: (B, ny, nx)
 *   2D sinogram  (A, D)              detector fastest   -- batch: (B, A, D)
 *   3D volume    (nz, ny, nx)        x fastest          -- batch: (B, nz, ny, nx)
 *   3D sinogram  (A, V, U)           u fastest          -- batch: (B, A, V, U)
 * Geometry arguments use the Python axis order of the reference ((y, x), (z, y, x), (v, u)).
 *
 * Semantics kept from the reference:
 *   forward  overwrites the sinogram (forward_2d.hpp:58-61,78);
 *   backward accumulates into the volume, += (backward_2d.hpp:99, backward_3d.hpp:100-107).
 * Differences, all stated where the entry point is declared: no host synchronisation (work
 * is enqueued on the given stream), 3D missed rays write 0, and no texture arrays: instead of
 * the reference's per-call copy into a texture array, a call stages zero-bordered copies of
 * its input in a stream-ordered scratch allocation (cmi_forward_2d_noalloc uses a static one).
 */
#ifndef CURADON_MI_H
#define CURADON_MI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define CMI_API __attribute__((visibility("default")))
#else
#define CMI_API
#endif

/* Status codes. The message of the last failure on the calling host thread is returned by
 * cmi_last_error(). Replaces gpuErrchk -> exit() (include/curadon/detail/error.h:6-18) and
 * the nanobind attribute_error / type_error raised by the binding
 * (python/standalone/src/forward.cu:95-105, 107-135; plan.cu:20-28). */
typedef enum cmi_status {
    CMI_OK = 0,
    CMI_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad shape, size or geometry value */
    CMI_ERR_DTYPE = 2,            /* unsupported dtype, or dtype differs from the plan's */
    CMI_ERR_DEVICE = 3,           /* array or plan on another device than expected */
    CMI_ERR_HIP = 4,              /* HIP runtime failure; message holds the HIP error text */
    CMI_ERR_OUT_OF_MEMORY = 5,
    CMI_ERR_NOT_SUPPORTED = 6
} cmi_status;

/* Storage precisions. F32 and F16 keep the numeric values of curad::precision
 * (include/curadon/detail/texture.hpp:18-21); BF16 is an addition. Arithmetic is FP32. */
typedef enum cmi_dtype { CMI_F16 = 16, CMI_F32 = 32, CMI_BF16 = 1016 } cmi_dtype;

/* Backprojector variants for 3D (see cmi_backward_3d). */
typedef enum cmi_bp3d_mode {
    CMI_BP3D_REFERENCE = 0, /* the reference's projection, unweighted */
    CMI_BP3D_MATCHED = 1,   /* the forward model's projection, unweighted */
    CMI_BP3D_FDK = 2        /* the forward model's projection, weight (DSO / depth)^2 */
} cmi_bp3d_mode;

/* Weightings of the 2D backprojector (see cmi_backward_2d_weighted). */
typedef enum cmi_bp2d_weighting {
    CMI_BP2D_REFERENCE = 0, /* w = DSD / (DSO + xi), the reference's magnification weight */
    CMI_BP2D_FBP = 1        /* w^2, the fan-beam FBP distance weight */
} cmi_bp2d_weighting;

/* Message of the last failed call on this host thread ("" if none). */
CMI_API const char *cmi_last_error(void);

/* Library version, major * 10000 + minor * 100 + patch. */
CMI_API int cmi_version(void);

/* ====================================================================== 2D fan beam == */

typedef struct cmi_plan2d cmi_plan2d;

/*
 * Replaces the plan_2d constructor bound at python/standalone/src/plan.cu:30-59
 * (curad::plan_2d, include/curadon/detail/plan/plan_2d.hpp:103-126). Same argument order:
 * device, volume precision, vol_shape (y, x), vol_spacing (y, x), vol_offset (y, x),
 * detector precision, det_count, det_spacing, det_offset, per-angle DSO / DSD / angles,
 * det_pitch, COR. The reference takes FP32 geometry arrays; these take FP64 (FP32 values
 * widen exactly) because the per-angle and per-ray geometry is computed in FP64.
 * det_pitch and COR are stored and unused, as in the reference (plan_2d.hpp:121-122); see
 * cmi_plan2d_create_ex for a plan that honours them.
 * Builds the per-angle geometry on the host and the per-ray table on `device`.
 */
CMI_API cmi_status cmi_plan2d_create(int device, int vol_prec, const uint64_t vol_shape[2],
                                     const double vol_spacing[2], const double vol_offset[2],
                                     int det_prec, uint64_t det_count, double det_spacing,
                                     double det_offset, const double *DSO, const double *DSD,
                                     const double *angles, uint64_t nangles, double det_pitch,
                                     double COR, cmi_plan2d **plan);

/*
 * A 2D plan from ASTRA-like views, one per projection: vecs[6 a .. 6 a + 5] = (source x, y,
 * detector centre x, y, u x, y) in world coordinates, u the step from cell 0's centre to cell
 * 1's. Every view may move the source, shift or tilt the detector and change the cell
 * spacing. No counterpart in the reference: its FanGeometry.from_vec keeps only DSO, DSD and
 * the angle of each view and rejects unequal spacings (python/standalone/src/curadon/
 * geometry.py:84-133). Forward, backprojection (per-view magnification, weight divided by the
 * view's |u|) and adjoint all follow the views; the serving tables (cmi_fan2d_bp_table) do not.
 */
CMI_API cmi_status cmi_plan2d_create_vec(int device, int vol_prec, const uint64_t vol_shape[2],
                                         const double vol_spacing[2], const double vol_offset[2],
                                         int det_prec, uint64_t det_count, const double *vecs,
                                         uint64_t nangles, cmi_plan2d **plan);

/*
 * cmi_plan2d_create with flags. flags = 0 is cmi_plan2d_create (the reference's semantics: the
 * forward moves the source by det_offset, the backprojection ignores det_offset, and det_pitch
 * and COR are unused; plan_2d.hpp:76-80, 121-122, backward_2d.hpp:58-92).
 * CMI_PLAN2D_EXACT_GEOMETRY: every projection becomes an ASTRA-like view (the plan is then a
 * plan from views, see cmi_plan2d_create_vec) in which, before the clockwise rotation by the
 * angle (plan_2d.hpp:17-80), the source sits at (-COR, DSO), the detector centre at
 * (det_offset - COR, -(DSD - DSO)) and the cell step is det_spacing (cos det_pitch,
 * sin det_pitch): det_offset shifts the detector along its own axis, det_pitch turns it about
 * its centre, and the rotation centre lies COR along the detector axis from the line through
 * the source and the detector centre; forward, backprojection and adjoint all use this
 * geometry. No counterpart in the reference.
 */
#define CMI_PLAN2D_EXACT_GEOMETRY 1u
CMI_API cmi_status cmi_plan2d_create_ex(int device, int vol_prec, const uint64_t vol_shape[2],
                                        const double vol_spacing[2], const double vol_offset[2],
                                        int det_prec, uint64_t det_count, double det_spacing,
                                        double det_offset, const double *DSO, const double *DSD,
                                        const double *angles, uint64_t nangles,
                                        double det_pitch, double COR, uint32_t flags,
                                        cmi_plan2d **plan);

CMI_API cmi_status cmi_plan2d_destroy(cmi_plan2d *plan);

/* Launch-shape tunables, the analogue of the r/w forward_block_x / forward_block_y
 * properties (plan.cu:57-58, plan_2d.hpp:184-186). block_x = detector cells per
 * workgroup (a multiple of 64), block_y = angles per workgroup, with
 * block_x * block_y <= 256; other shapes are rejected with CMI_ERR_INVALID_ARGUMENT. */
CMI_API cmi_status cmi_plan2d_set_block(cmi_plan2d *plan, uint32_t block_x, uint32_t block_y);
CMI_API cmi_status cmi_plan2d_get_block(const cmi_plan2d *plan, uint32_t *block_x,
                                        uint32_t *block_y);

/* Shape queries. total_samples = sum over rays of the forward sample count N. */
typedef struct cmi_plan2d_info {
    int device;
    int vol_prec, det_prec;
    uint64_t nx, ny, det_count, nangles;
    uint64_t total_samples;
    uint64_t table_bytes;
} cmi_plan2d_info;
CMI_API cmi_status cmi_plan2d_get_info(const cmi_plan2d *plan, cmi_plan2d_info *info);

/*
 * Ray-driven forward projection, sinogram := A volume, for `batch` images.
 * Replaces curadon_ext.forward_2d (python/standalone/src/forward.cu:92-136 ->
 * curad::fp::forward_2d, include/curadon/detail/forward_2d.hpp:82-139).
 * vol_dtype must equal the plan's volume precision (forward.cu:107-123); sino_dtype must
 * equal the plan's detector precision (not checked by the reference). Enqueued on `stream`;
 * does not block. Scratch: the call stages each image into zero-padded row-major and
 * transposed copies from a stream-ordered allocation that is released on the same stream.
 */
CMI_API cmi_status cmi_forward_2d(const cmi_plan2d *plan, const void *volume, int vol_dtype,
                                  void *sinogram, int sino_dtype, uint64_t batch, void *stream);

/*
 * cmi_forward_2d without any allocation: no counterpart in the reference. For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Same arguments, validation and result as cmi_forward_2d (the same
 * kernels, bit-identical output); the staging scratch (two zero-bordered copies per image,
 * 2 (nx + 4) (ny + 4) bytes-per-element per image) is a static device array of the library of
 * CMI_NOALLOC_SCRATCH_BYTES, which each GPU holds for the life of the process. A call that needs
 * more returns CMI_ERR_INVALID_ARGUMENT (for FP32, single images up to about 2890 x 2890).
 * All calls on one GPU share that GPU's array: calls that overlap on one GPU (several streams
 * or host threads) race, so callers serialise them.
 */
#define CMI_NOALLOC_SCRATCH_BYTES (64ull << 20)
CMI_API cmi_status cmi_forward_2d_noalloc(const cmi_plan2d *plan, const void *volume,
                                          int vol_dtype, void *sinogram, int sino_dtype,
                                          uint64_t batch, void *stream);

/*
 * Voxel-driven backprojection with the reference's weighting, volume += B sinogram.
 * Replaces curadon_ext.backward_2d (python/standalone/src/backward.cu:89-133 ->
 * curad::bp::backproject_2d, include/curadon/detail/backward_2d.hpp:103-166). All angles
 * are processed in one launch (no 1024-angle chunking). Enqueued on `stream`.
 */
CMI_API cmi_status cmi_backward_2d(const cmi_plan2d *plan, void *volume, int vol_dtype,
                                   const void *sinogram, int sino_dtype, uint64_t batch,
                                   void *stream);

/*
 * cmi_backward_2d with a choice of weighting: CMI_BP2D_REFERENCE is cmi_backward_2d,
 * CMI_BP2D_FBP squares the magnification weight, which a fan-beam filtered backprojection needs
 * (the reference notes its weighting is not the FBP one, backward_2d.hpp:94-95).
 */
CMI_API cmi_status cmi_backward_2d_weighted(const cmi_plan2d *plan, void *volume, int vol_dtype,
                                            const void *sinogram, int sino_dtype,
                                            uint64_t batch, int weighting, void *stream);

/*
 * Matched adjoint of cmi_forward_2d, volume += A^T sinogram: the exact transpose of the
 * ray-march above (same sample positions and bilinear weights). Sums are fixed point in units
 * of max|y scale| / 2^19 per image, kept exactly (32-bit LDS cells moved into 64-bit sums
 * before they could overflow), so the result is deterministic and a batched call equals
 * single calls bit for bit. Scratch (stream-ordered): 8 bytes per volume
 * element plus 1 KB per image. No counterpart in the reference, whose autograd pairs the
 * forward with the voxel-driven backprojector (python/standalone/src/curadon/diffable.py:22-24).
 */
CMI_API cmi_status cmi_adjoint_2d(const cmi_plan2d *plan, void *volume, int vol_dtype,
                                  const void *sinogram, int sino_dtype, uint64_t batch,
                                  void *stream);

/*
 * Transpose of cmi_backward_2d_weighted, sinogram := B^T volume (overwrites): the gradient of
 * the voxel-driven backprojection, with the backprojector's own FP32 detector coordinates and
 * weights (reference or FBP weighting; circular plans and plans from views). Replaces the
 * reference's approximate gradient of TorchBackprojection, the forward projection
 * (python/standalone/src/curadon/diffable.py:35-38). Sums are fixed point in units of
 * max|volume| times the plan's weight bound over 2^18, kept exactly (32-bit LDS cells moved
 * into 64-bit registers before they could overflow), so the result is deterministic.
 * CMI_ERR_NOT_SUPPORTED when a source lies inside the image's bounding box. Scratch
 * (stream-ordered): two zero-bordered copies of each image plus 1 KB per image.
 */
CMI_API cmi_status cmi_backward_2d_transpose(const cmi_plan2d *plan, void *sinogram,
                                             int sino_dtype, const void *volume, int vol_dtype,
                                             uint64_t batch, int weighting, void *stream);

/* ===================================================================== 3D cone beam == */

/*
 * Cone-beam description, in the argument order and axis order of curadon_ext.forward_3d /
 * backward_3d (python/standalone/src/forward.cu:23-38, backward.cu:20-35): vol_shape /
 * vol_spacing / vol_offset in (z, y, x), det_shape / det_spacing / det_offset in (v, u),
 * det_rotation as (pitch, roll, yaw), per-angle phi / theta / psi, scalar DSO, DSD, COR.
 */
typedef struct cmi_cone3d_desc {
    uint64_t vol_shape[3];
    double vol_spacing[3];
    double vol_offset[3];
    uint64_t det_shape[2];
    double det_spacing[2];
    double det_offset[2];
    double det_rotation[3];
    double DSO, DSD, COR;
    uint64_t nangles;
    const double *phi;
    const double *theta;
    const double *psi;
} cmi_cone3d_desc;

typedef struct cmi_plan3d cmi_plan3d;

/* The reference has no 3D plan: forward_3d / backproject_3d rebuild their geometry, allocate
 * a cudaArray and upload __constant__ tables on every call (forward_3d.hpp:214-274,
 * backward_3d.hpp:181-253). This plan holds the per-angle tables on `device` instead. */
CMI_API cmi_status cmi_plan3d_create(int device, int vol_prec, int det_prec,
                                     const cmi_cone3d_desc *desc, cmi_plan3d **plan);

/*
 * A 3D plan from ASTRA-like cone views, one per projection: vecs[12 a .. 12 a + 11] = (source
 * x, y, z, panel centre x, y, z, u-step x, y, z, v-step x, y, z) in world coordinates (x, y, z;
 * the volume's centre sits at vol_offset, as in cmi_plan3d_create), the u-step going from pixel
 * (iu, iv) to (iu + 1, iv) and the v-step to (iu, iv + 1); pixel (0, 0)'s centre is
 * centre - (U - 1)/2 u - (V - 1)/2 v. Shapes and spacings in the Python axis order of
 * cmi_cone3d_desc: vol_shape / vol_spacing / vol_offset (z, y, x), det_shape (v, u). Every
 * view may move the source and place, tilt and scale the panel freely (any trajectory, a
 * projection-matrix geometry given by its source and panel). No counterpart in the reference,
 * whose cone geometry takes Euler angles only and lists projection matrices as future work
 * (README.md:58-59). cmi_forward_3d, cmi_adjoint_3d and cmi_backward_3d in CMI_BP3D_MATCHED
 * mode follow the views; CMI_BP3D_REFERENCE mode projects like CMI_BP3D_MATCHED (the
 * reference's own backprojection frame exists only for Euler angles) and CMI_BP3D_FDK returns
 * CMI_ERR_NOT_SUPPORTED (its weight needs a circular orbit's DSO and DSD).
 */
CMI_API cmi_status cmi_plan3d_create_vec(int device, int vol_prec, int det_prec,
                                         const uint64_t vol_shape[3], const double vol_spacing[3],
                                         const double vol_offset[3], const uint64_t det_shape[2],
                                         const double *vecs, uint64_t nangles,
                                         cmi_plan3d **plan);

CMI_API cmi_status cmi_plan3d_destroy(cmi_plan3d *plan);

/* Shape queries. total_samples = sum over rays that hit the clip box of the sample count
 * nst + 1 of the forward march. */
typedef struct cmi_plan3d_info {
    int device;
    int vol_prec, det_prec;
    uint64_t nx, ny, nz, nu, nv, nangles;
    uint64_t total_samples;
    uint64_t table_bytes;
} cmi_plan3d_info;
CMI_API cmi_status cmi_plan3d_get_info(const cmi_plan3d *plan, cmi_plan3d_info *info);

/*
 * Cone-beam forward projection, sinogram := A volume (unscaled sum of trilinear samples, as
 * forward_3d.hpp:87-93). Replaces curadon_ext.forward_3d (forward.cu:23-75 ->
 * curad::fp::forward_3d, forward_3d.hpp:214-274). Rays that miss the volume write 0 (the
 * reference leaves them unwritten, forward_3d.hpp:74-76).
 */
CMI_API cmi_status cmi_forward_3d(const cmi_plan3d *plan, const void *volume, int vol_dtype,
                                  void *sinogram, int sino_dtype, uint64_t batch, void *stream);

/*
 * Voxel-driven backprojection, volume += B sinogram. Replaces curadon_ext.backward_3d
 * (backward.cu:20-72 -> curad::bp::backproject_3d, backward_3d.hpp:181-253).
 * mode CMI_BP3D_REFERENCE reproduces the reference's projection (detector plane at
 * z = DSO - DSD, integer detector centre, detector spacing ignored; backward_3d.hpp:79-94).
 * mode CMI_BP3D_MATCHED projects each voxel through the forward model's geometry instead.
 * mode CMI_BP3D_FDK projects like CMI_BP3D_MATCHED and weights each sample by
 * (DSO / (t DSD))^2, t the voxel's parameter on its ray (t = 1 at the panel): the FDK distance
 * weight for an untilted panel.
 */
CMI_API cmi_status cmi_backward_3d(const cmi_plan3d *plan, void *volume, int vol_dtype,
                                   const void *sinogram, int sino_dtype, uint64_t batch,
                                   int mode, void *stream);

/* Matched adjoint of cmi_forward_3d, volume += A^T sinogram: every forward sample scatters
 * its trilinear weights times the ray's value, at the forward march's own positions. The sums
 * are fixed point in units of max|y| / 2^19 per volume, kept exactly (32-bit LDS cells moved
 * into 64-bit sums before they could overflow), so the result is deterministic and a batched
 * call equals single calls bit for bit. Scratch (stream-ordered):
 * a 32-byte march table entry per ray, 8 bytes per voxel (rounded up to 16^3 tiles) and 1 KB per
 * volume. No counterpart in the reference. */
CMI_API cmi_status cmi_adjoint_3d(const cmi_plan3d *plan, void *volume, int vol_dtype,
                                  const void *sinogram, int sino_dtype, uint64_t batch,
                                  void *stream);

/*
 * Transpose of cmi_backward_3d, sinogram := B^T volume (overwrites), in the same mode: the
 * gradient of the voxel-driven backprojection, with the backprojector's FP32 projection
 * matrices and bilinear weights (times (DSO / (t DSD))^2 for CMI_BP3D_FDK). Replaces the
 * reference's approximate gradient of TorchBackprojection, the forward projection
 * (python/standalone/src/curadon/diffable.py:35-38). Sums are fixed point in units of
 * max|volume| times the weight bound over 2^18, kept exactly (32-bit LDS cells moved into 64-bit
 * registers before they could overflow), so the result is deterministic. CMI_ERR_NOT_SUPPORTED
 * when W <= 0 at a corner voxel of the volume for some projection (a voxel at or behind the
 * source), and for CMI_BP3D_FDK on plans from views. Scratch (stream-ordered): a y-fastest
 * copy of each volume with a 2-voxel border, plus 1 KB per volume.
 */
CMI_API cmi_status cmi_backward_3d_transpose(const cmi_plan3d *plan, void *sinogram,
                                             int sino_dtype, const void *volume, int vol_dtype,
                                             uint64_t batch, int mode, void *stream);

/* ============================================================ serving: multi-job launch == */

/*
 * Host helpers that build the backprojection tables a serving job carries, from the same
 * geometry arguments as the plans (no counterpart in the reference, whose backprojectors
 * rebuild them per call: backward_2d.hpp:58-62, backward_3d.hpp:110-178).
 *   cmi_fan2d_bp_table: table[4a..4a+3] = (-sin a, cos a, DSD[a], DSO[a]) in FP32.
 *   cmi_cone3d_bp_matrices: 12 floats per angle, the 3x4 matrix of `mode` (cmi_bp3d_mode)
 *   that cmi_backward_3d uses for the same description; *y_invariant (if not null) is set to
 *   1 when no angle's U or W row depends on y (circular orbits about the y axis, untilted
 *   panel), the case in which the backprojectors run along y (see cmi_bp_job.columns).
 */
CMI_API cmi_status cmi_fan2d_bp_table(uint64_t nangles, const double *angles, const double *DSO,
                                      const double *DSD, float *table);
CMI_API cmi_status cmi_cone3d_bp_matrices(const cmi_cone3d_desc *desc, int mode, float *out,
                                          int *y_invariant);

/*
 * Copies filtered projections into the zero-bordered layout the serving backprojector reads,
 * converting to `dst_dtype`:
 *   dst[a][vp][up] = scale * row_scale[a] * src[(a * V + vp - vpad) * src_stride + up - 2]
 * for vp - vpad in [0, V) and up - 2 in [0, U), and 0 elsewhere; dst is [A][V + 2 vpad][U + 4].
 * Fan-beam jobs pass V = 1, vpad = 0 ([A][D + 4]); cone-beam jobs V = panel rows, vpad = 2.
 * `src` is FP32 device memory (for instance the output of an inverse real FFT of length
 * src_stride >= U); row_scale may be null (all ones). `dst` may have any element alignment
 * (a view with a storage offset, say): when it is aligned to four elements, four are written
 * with one vector store, otherwise one at a time. Enqueued on `stream`.
 */
CMI_API cmi_status cmi_stage_projections(int device, const float *src, uint64_t src_stride,
                                         uint64_t A, uint64_t V, uint64_t U, uint32_t vpad,
                                         const float *row_scale, float scale, void *dst,
                                         int dst_dtype, void *stream);

/*
 * Ramp filtering of projection rows through complex FFTs of row pairs. No counterpart in the
 * reference, which filters with a circular FFT per row in Python (examples/fbp.py:51-56); these
 * are the native halves of curadon_amd.reconstruction.filter_rows_weighted. The filter's
 * response is real and even, so a complex row whose real and imaginary parts are two projection
 * rows filters both at once.
 * cmi_filter_pack: dst (complex, [ceil(rows / 2)][L] interleaved FP32) row r := (src[2r] * cw[2r
 * mod cw_rows], src[2r + 1] * cw[2r + 1 mod cw_rows]) in its first D entries, 0 after; an odd
 * last row pairs with zeros. src is FP32 [rows][D], cw FP32 [cw_rows][D]; L >= D.
 * cmi_filter_unpack: dst (FP32 [rows][D]) rows 2r, 2r + 1 := the real and imaginary parts of
 * the first D entries of src row r (complex [ceil(rows / 2)][L]).
 * Device memory on `device`, enqueued on `stream`.
 */
CMI_API cmi_status cmi_filter_pack(int device, const float *src, uint64_t rows, uint64_t D,
                                   const float *cw, uint64_t cw_rows, uint64_t L, float *dst,
                                   void *stream);
CMI_API cmi_status cmi_filter_unpack(int device, const float *src, uint64_t rows, uint64_t D,
                                     uint64_t L, float *dst, void *stream);

/* Job kinds of the serving backprojector. */
typedef enum cmi_job_kind { CMI_JOB_FAN2D = 0, CMI_JOB_CONE3D = 1 } cmi_job_kind;

/*
 * One backprojection job of a multi-job launch: volume += B sinogram over the tile range
 * [tile_begin, tile_end) of the job's volume. Fan-beam tiles are 64 x 4 pixels (row-major
 * over ceil(nx / 64) x ceil(ny / 4)). Cone-beam tiles are 16 x 16 x 8 voxels (x fastest, then
 * y, then z blocks) with columns = 0, or 16 x 16 x 16 (x, y, z; x fastest, then y, then z
 * blocks) with columns = 1, which the caller may set only for y_invariant matrices. The
 * arithmetic of every voxel equals cmi_backward_2d_weighted / cmi_backward_3d for the same
 * tables (cmi_backward_3d runs along y exactly when the matrices are y_invariant), so a job's
 * result does not depend on the other jobs of the launch or on how its tiles are split
 * between launches.
 *   sino    staged projections: [nangles][nu + 4] (fan) or [nangles][nv + 4][nu + 4] (cone)
 *   table   device copy of cmi_fan2d_bp_table / cmi_cone3d_bp_matrices output
 *   volume  [ny][nx] / [nz][ny][nx], accumulated into
 *   mode    fan: cmi_bp2d_weighting; cone: cmi_bp3d_mode (the mode the matrices were made for)
 *   columns cone only: 0 = z columns (any geometry), 1 = y runs (y_invariant matrices only)
 *   fan only: sx, sy pixel spacing, ox, oy volume offset, inv_du = 1 / detector spacing
 * No counterpart in the reference, which runs one operator call per image.
 */
typedef struct cmi_bp_job {
    int32_t kind;
    int32_t mode;
    int32_t columns;
    const void *sino;
    const void *table;
    void *volume;
    uint32_t nx, ny, nz;
    uint32_t nu, nv, nangles;
    float sx, sy, ox, oy, inv_du;
    uint32_t tile_begin, tile_end;
} cmi_bp_job;

/* Number of tiles of a job's volume (independent of tile_begin / tile_end). */
CMI_API cmi_status cmi_bp_job_tiles(const cmi_bp_job *job, uint64_t *tiles);

typedef struct cmi_server cmi_server;

/* A multi-job launcher on `device`: owns a ring of device job tables (max_jobs entries each)
 * and tile counters. */
CMI_API cmi_status cmi_server_create(int device, uint32_t max_jobs, cmi_server **server);
CMI_API cmi_status cmi_server_destroy(cmi_server *server);

/* Workgroups of the cone-beam serving kernel for jobs of either tile layout that the device
 * holds resident at once (the persistent grid of a launch with at least that many cone-beam
 * tiles). Launches whose cone-beam jobs all take y runs, and the fan-beam tiles, use kernels of
 * their own with their own resident counts. */
CMI_API cmi_status cmi_server_get_capacity(const cmi_server *server, uint32_t *workgroups);

/*
 * One persistent launch over `njobs` jobs whose volumes are stored as vol_dtype and sinograms
 * as sino_dtype: a grid sized to the device's resident capacity pulls tiles from an atomic
 * counter (zeroed by a fill kernel enqueued before it) until all tiles of all jobs are done.
 * Tiles are handed out in job order, so callers list the costliest tiles first. Enqueued on
 * `stream`; the job array is copied, the caller may reuse it on return. Jobs of one launch
 * must write disjoint volume regions.
 */
CMI_API cmi_status cmi_serve_backproject(cmi_server *server, const cmi_bp_job *jobs,
                                         uint32_t njobs, int vol_dtype, int sino_dtype,
                                         void *stream);

#ifdef __cplusplus
}
#endif
#endif
