/*
 * curadon_oracle.h -- CPU restatement of curadon's projector arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP projectors in
 * curadon_amd/csrc: only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark load it. The product path never calls it.
 *
 * Every routine evaluates the reference's per-ray / per-voxel arithmetic in IEEE FP64
 * (the *_f32 variants repeat the 2D routines in FP32 and exist only to time a CPU port of
 * the reference math). Interpolation uses exact weights with the reference's texture
 * semantics: unnormalised coordinates, texel i centred at i + 0.5, linear filtering and
 * border addressing that reads 0 outside the array
 * (include/curadon/detail/texture.hpp:81-87, texture_utils.hpp:31-34).
 *
 * Memory layouts are the reference's: images/volumes x fastest, sinograms (A, D) and
 * (A, V, U) row-major with the detector coordinate fastest.
 *
 * Which check pins each routine is listed in the header comment of each function group.
 */
#ifndef CURADON_ORACLE_H
#define CURADON_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------------------------------------
 * 2D fan beam, flat detector.
 * Geometry as held by plan_2d (include/curadon/detail/plan/plan_2d.hpp:143-167): image
 * shape/spacing/offset in (x, y) order, detector count/spacing/offset, per-angle arrays.
 * ------------------------------------------------------------------------------------- */
typedef struct {
    int64_t nx, ny;       /* image shape, x fastest */
    double sx, sy;        /* pixel spacing */
    double ox, oy;        /* image offset */
    int64_t det_count;    /* D */
    double du;            /* detector cell spacing */
    double det_offset;    /* moves the source, as plan_2d.hpp:76-80 does */
    int64_t nangles;      /* A */
    const double *angles; /* [A] radians */
    const double *dso;    /* [A] source to isocentre */
    const double *dsd;    /* [A] source to detector */
    const double *vecs;   /* NULL, or [A][6] ASTRA-like views (source x, y, detector centre x, y,
                             u x, y; world coordinates, u = cell 0 centre to cell 1 centre),
                             which then replace angles / DSO / DSD / du / det_offset */
} orc_fan2d;

/* Per-angle source position, vector from the source to cell 0 and the cell-to-cell step,
 * all in texel coordinates (plan_2d.hpp:17-80, 189-241). Outputs are [A][2]. */
void orc_plan2d(const orc_fan2d *g, double *src, double *origin, double *delta);

/* Ray (a, k) after the slab clip (intersection.h:30-46) and the march set-up of
 * kernel_forward_2d (forward_2d.hpp:52-77): entry point e, unit-L-inf step v, sample count
 * N = rint(L-inf chord) and the world step length. Returns N; 0 for a missed ray. */
int64_t orc_ray2d(const orc_fan2d *g, const double *src, const double *origin,
                  const double *delta, int64_t a, int64_t k, double e[2], double v[2],
                  double *scale);

/* Forward projection of rows [a0, a1) of the sinogram (forward_2d.hpp:21-79). Missed rays
 * write 0. If nsamples is non-null it receives the total sample count of those rows. */
void orc_fwd2d(const orc_fan2d *g, const double *img, double *sino, int64_t a0, int64_t a1,
               int64_t *nsamples);

/* Exact transpose of orc_fwd2d restricted to rows [a0, a1): img += A^T sino. */
void orc_adj2d(const orc_fan2d *g, const double *sino, double *img, int64_t a0, int64_t a1);

/* Voxel-driven backprojection (backward_2d.hpp:21-100) of image rows [y0, y1):
 * img[y][x] += (1/du) * sum_a w_a * lerp(sino[a], u_a(x, y)). */
void orc_bp2d(const orc_fan2d *g, const double *sino, double *img, int64_t y0, int64_t y1);

/* Same pixel positions and detector coordinates as orc_bp2d with the magnification weight
 * raised to the power wpow: wpow = 1 is orc_bp2d, wpow = 2 the fan-beam FBP distance weight
 * (DSD / (DSO + xi))^2. */
void orc_bp2d_w(const orc_fan2d *g, const double *sino, double *img, int64_t y0, int64_t y1,
                int wpow);
/* Transpose of orc_bp2d_w (sino[a][c] += the backprojection's weights times img), projections
 * a0 .. a1 - 1; the gradient of the voxel-driven backprojection. */
void orc_bp2d_T_w(const orc_fan2d *g, const double *img, double *sino, int64_t a0, int64_t a1,
                  int wpow);

/* FP32 ports of the two hot loops, same arithmetic in single precision, OpenMP over
 * angles / image rows. Used only to time the reference math on host cores. */
void orc_fwd2d_f32(const orc_fan2d *g, const float *img, float *sino, int64_t a0, int64_t a1);
void orc_bp2d_f32(const orc_fan2d *g, const float *sino, float *img, int64_t y0, int64_t y1);

/* OpenMP threads the parallel routines use (1 when built without OpenMP), and their setter
 * (n > 0; no effect without OpenMP). */
int orc_num_threads(void);
void orc_set_num_threads(int n);

/* Exact-weight interpolation helpers (texture semantics), exported for known-answer tests. */
double orc_lerp1(const double *row, int64_t n, double u);
double orc_bilerp(const double *img, int64_t nx, int64_t ny, double x, double y);
double orc_trilerp(const double *vol, int64_t nx, int64_t ny, int64_t nz, double x, double y,
                   double z);

/* ----------------------------------------------------------------------------------------
 * 3D cone beam, flat panel, circular or general (phi, theta, psi) trajectory.
 * Field order follows device_measurement / device_volume after the binding's axis reversal
 * (python/standalone/src/forward.cu:47-58): volume (x, y, z), detector (u, v).
 * ------------------------------------------------------------------------------------- */
typedef struct {
    int64_t nx, ny, nz;
    double sx, sy, sz;
    double ox, oy, oz;
    int64_t nu, nv;
    double du, dv;
    double off_u, off_v;
    double pitch, roll, yaw; /* detector rotation */
    double dso, dsd, cor;
    int64_t nangles;
    const double *phi, *theta, *psi; /* [A] each */
    /* ASTRA-like views [A][12] = (source, panel centre, u-step, v-step) in world (x, y, z), or
     * NULL. When set, they replace the Euler-angle set-up (phi / theta / psi, DSO, DSD, COR,
     * detector offset and rotation are then unused), and the reference-mode backprojection
     * projects like the matched one. */
    const double *vecs;
} orc_cone3d;

/* Rotations of rotation.hpp:7-48, exported for the orthogonality / special-case tests. */
void orc_rotate_yzy(const double v[3], double phi, double theta, double psi, double out[3]);
void orc_rotate_rpy(const double v[3], double roll, double pitch, double yaw, double out[3]);

/* Forward set-up (forward_3d.hpp:97-211): per angle the detector pixel (0,0) position, the
 * u- and v-steps and the source, in voxel coordinates (voxel i centred at i). [A][3] each.
 * With views: pixel (0,0) at centre - (U-1)/2 u - (V-1)/2 v, every point taken into the voxel
 * frame as the reference's steps 4-5 do (forward_3d.hpp:158-169). */
void orc_plan3d_fwd(const orc_cone3d *g, double *origin, double *du, double *dv, double *src);

/* Ray (a, iu, iv) of kernel_forward_3d (forward_3d.hpp:57-90): returns the number of
 * march steps nst (the ray takes nst + 1 samples t_i = tmin + i * dt, i = 0..nst), or -1 for
 * a ray that misses the clip box [-1, n+1]^3. */
int64_t orc_ray3d(const orc_cone3d *g, const double *origin, const double *du,
                  const double *dv, const double *src, int64_t a, int64_t iu, int64_t iv,
                  double s[3], double dir[3], double *tmin, double *dt);

/* Forward projection of projections [a0, a1): unscaled sum of trilinear samples
 * (forward_3d.hpp:87-93). Missed rays write 0. nsamples as in orc_fwd2d. */
void orc_fwd3d(const orc_cone3d *g, const double *vol, double *sino, int64_t a0, int64_t a1,
               int64_t *nsamples);

/* Exact transpose of orc_fwd3d over projections [a0, a1): vol += A^T sino. */
void orc_adj3d(const orc_cone3d *g, const double *sino, double *vol, int64_t a0, int64_t a1);

/* Reference-mode voxel-driven backprojection (backward_3d.hpp:28-108, 110-178, 218-227):
 * per angle the volume origin and voxel steps are rotated into the detector frame and each
 * voxel is projected onto the plane z = DSO - DSD; unweighted bilinear sum.
 * Outputs [A][3] for origin / steps and [3] for the source. */
void orc_plan3d_bp(const orc_cone3d *g, double *vorigin, double *vdx, double *vdy, double *vdz,
                   double src[3]);
void orc_bp3d(const orc_cone3d *g, const double *sino, double *vol, int64_t z0, int64_t z1);

/* Geometry-consistent voxel-driven backprojection: each voxel centre is projected through
 * the forward geometry of orc_plan3d_fwd onto the panel, and the panel is sampled
 * bilinearly there; unweighted sum. This is the reference's backprojector with the
 * detector plane, rotation direction and detector spacing taken from the forward model. */
void orc_bp3d_matched(const orc_cone3d *g, const double *sino, double *vol, int64_t z0,
                      int64_t z1);

/* orc_bp3d_matched with the FDK distance weight (DSO / (t DSD))^2, where t is the voxel's
 * parameter on its ray (t = 1 at the panel), i.e. (DSO / depth)^2 for an untilted panel. */
void orc_bp3d_fdk(const orc_cone3d *g, const double *sino, double *vol, int64_t z0, int64_t z1);

/* Transpose of the 3D backprojection of `mode` (0 reference, 1 matched, 2 FDK):
 * sino[a] += B_a^T vol for projections a0 .. a1 - 1. */
void orc_bp3d_T(const orc_cone3d *g, int mode, const double *vol, double *sino, int64_t a0,
                int64_t a1);

/* FP32 ports of orc_fwd3d and orc_bp3d (reference mode), used only to time a CPU port of the
 * reference arithmetic. */
void orc_fwd3d_f32(const orc_cone3d *g, const float *vol, float *sino, int64_t a0, int64_t a1);
void orc_bp3d_f32(const orc_cone3d *g, const float *sino, float *vol, int64_t z0, int64_t z1);

#ifdef __cplusplus
}
#endif
#endif
