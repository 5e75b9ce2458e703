/*
 * curadon_oracle.c -- FP64 CPU restatement of curadon's projector arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY (see curadon_oracle.h). Each function cites the reference
 * file:line it restates. Built by oracle/Makefile into oracle/_build/libcuradon_oracle.so.
 *
 * Reading conventions taken from the reference and kept here:
 *  - the slab clip starts from t in [FLT_MIN, FLT_MAX] and combines the per-axis bounds with
 *    IEEE minNum/maxNum (intersection.h:19-25, 37-43), so a 0/0 = NaN bound is ignored;
 *  - the 2D miss test is tmin > tmax - 1e-6 with the constant in double
 *    (forward_2d.hpp:58);
 *  - the 2D sample count is round-half-even of the L-inf chord (forward_2d.hpp:66);
 *  - the 2D recurrence ro += v (forward_2d.hpp:72-75) is evaluated in closed form e + j v;
 *  - the 3D float loop for (t = tmin; t <= tmax; t += dt) (forward_3d.hpp:87) is read in
 *    exact arithmetic: nst + 1 samples t_i = tmin + i dt, i = 0..nst. Its two end samples lie
 *    on the clip-box faces, one voxel outside the data, where the trilinear value is 0, so
 *    the reading does not change the result.
 */
#include "curadon_oracle.h"

#include <float.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ---------------------------------------------------------------- interpolation (A.0) -- */

/* Texel fetch with border addressing: 0 outside [0, n). */
static inline double fetch1(const double *r, int64_t n, int64_t i) {
    return (i >= 0 && i < n) ? r[i] : 0.0;
}

/* tex1DLayered with linear filter at unnormalised coordinate u (texel i centred at i+0.5). */
double orc_lerp1(const double *row, int64_t n, double u) {
    const double ub = u - 0.5;
    const double fl = floor(ub);
    const double al = ub - fl;
    const int64_t i = (int64_t)fl;
    return (1.0 - al) * fetch1(row, n, i) + al * fetch1(row, n, i + 1);
}

static inline double fetch2(const double *img, int64_t nx, int64_t ny, int64_t i, int64_t j) {
    return (i >= 0 && i < nx && j >= 0 && j < ny) ? img[j * nx + i] : 0.0;
}

/* tex2D with linear filter; texel (i, j) is img[j][i]. */
double orc_bilerp(const double *img, int64_t nx, int64_t ny, double x, double y) {
    const double xb = x - 0.5, yb = y - 0.5;
    const double fx = floor(xb), fy = floor(yb);
    const double a = xb - fx, b = yb - fy;
    const int64_t i = (int64_t)fx, j = (int64_t)fy;
    const double r0 = (1.0 - a) * fetch2(img, nx, ny, i, j) + a * fetch2(img, nx, ny, i + 1, j);
    const double r1 =
        (1.0 - a) * fetch2(img, nx, ny, i, j + 1) + a * fetch2(img, nx, ny, i + 1, j + 1);
    return (1.0 - b) * r0 + b * r1;
}

static inline double fetch3(const double *v, int64_t nx, int64_t ny, int64_t nz, int64_t i,
                            int64_t j, int64_t k) {
    return (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz)
               ? v[(k * ny + j) * nx + i]
               : 0.0;
}

/* tex3D with linear filter at texel coordinates (x, y, z). */
double orc_trilerp(const double *vol, int64_t nx, int64_t ny, int64_t nz, double x, double y,
                   double z) {
    const double xb = x - 0.5, yb = y - 0.5, zb = z - 0.5;
    const double fx = floor(xb), fy = floor(yb), fz = floor(zb);
    const double a = xb - fx, b = yb - fy, c = zb - fz;
    const int64_t i = (int64_t)fx, j = (int64_t)fy, k = (int64_t)fz;
    double acc = 0.0;
    for (int dk = 0; dk < 2; ++dk) {
        const double wz = dk ? c : 1.0 - c;
        for (int dj = 0; dj < 2; ++dj) {
            const double wy = dj ? b : 1.0 - b;
            const double r = (1.0 - a) * fetch3(vol, nx, ny, nz, i, j + dj, k + dk) +
                             a * fetch3(vol, nx, ny, nz, i + 1, j + dj, k + dk);
            acc += wz * wy * r;
        }
    }
    return acc;
}

/* Scatter form of orc_bilerp: adds s * weight to every texel the sample reads. */
static void bilerp_scatter(double *img, int64_t nx, int64_t ny, double x, double y, double s) {
    const double xb = x - 0.5, yb = y - 0.5;
    const double fx = floor(xb), fy = floor(yb);
    const double a = xb - fx, b = yb - fy;
    const int64_t i = (int64_t)fx, j = (int64_t)fy;
    const double w[4] = {(1.0 - a) * (1.0 - b), a * (1.0 - b), (1.0 - a) * b, a * b};
    const int64_t ii[4] = {i, i + 1, i, i + 1}, jj[4] = {j, j, j + 1, j + 1};
    for (int t = 0; t < 4; ++t)
        if (ii[t] >= 0 && ii[t] < nx && jj[t] >= 0 && jj[t] < ny) img[jj[t] * nx + ii[t]] += s * w[t];
}

static void trilerp_scatter(double *vol, int64_t nx, int64_t ny, int64_t nz, double x, double y,
                            double z, double s) {
    const double xb = x - 0.5, yb = y - 0.5, zb = z - 0.5;
    const double fx = floor(xb), fy = floor(yb), fz = floor(zb);
    const double a = xb - fx, b = yb - fy, c = zb - fz;
    const int64_t i = (int64_t)fx, j = (int64_t)fy, k = (int64_t)fz;
    for (int dk = 0; dk < 2; ++dk)
        for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di) {
                const int64_t I = i + di, J = j + dj, K = k + dk;
                if (I < 0 || I >= nx || J < 0 || J >= ny || K < 0 || K >= nz) continue;
                const double w = (di ? a : 1.0 - a) * (dj ? b : 1.0 - b) * (dk ? c : 1.0 - c);
                vol[(K * ny + J) * nx + I] += s * w;
            }
}

/* ------------------------------------------------------------------- 2D plan (A.1) ----- */

/* Clockwise rotation by phi followed by a translation by -t (rotation.hpp:51-61). */
static inline void rot2(const double v[2], double c, double s, const double t[2], double out[2]) {
    out[0] = v[0] * c + v[1] * s - t[0];
    out[1] = -v[0] * s + v[1] * c - t[1];
}

/* Vector form (no counterpart in the reference, whose from_vec keeps only DSO / DSD / angle,
 * python/standalone/src/curadon/geometry.py:84-133): the view gives the source s, the detector
 * centre c and the cell step u in world coordinates, so cell 0's centre is c - (D - 1)/2 u, and
 * the texel-frame quantities follow exactly as in plan_2d.hpp:28-48, 226-235. */
static void plan2d_vec(const orc_fan2d *g, double *src, double *origin, double *delta) {
    const double ext[2] = {(double)g->nx * g->sx, (double)g->ny * g->sy};
    const double sp[2] = {g->sx, g->sy}, off[2] = {g->ox, g->oy};
    const double h = ((double)g->det_count - 1.0) / 2.0;
    for (int64_t a = 0; a < g->nangles; ++a) {
        const double *v = g->vecs + 6 * a;
        for (int d = 0; d < 2; ++d) {
            const double r0 = v[2 + d] - h * v[4 + d] - v[d]; /* source to cell 0 */
            const double r1 = r0 + v[4 + d];                   /* source to cell 1 */
            const double o = r0 / sp[d];
            origin[2 * a + d] = o;
            delta[2 * a + d] = r1 / sp[d] - o;
            src[2 * a + d] = (v[d] - off[d] + 0.5 * ext[d]) / sp[d];
        }
    }
}

void orc_plan2d(const orc_fan2d *g, double *src, double *origin, double *delta) {
    if (g->vecs) {
        plan2d_vec(g, src, origin, delta);
        return;
    }
    const double det_extent = (double)g->det_count * g->du;        /* plan_2d.hpp:116 */
    const double ext[2] = {(double)g->nx * g->sx, (double)g->ny * g->sy}; /* :112 */
    const double sp[2] = {g->sx, g->sy}, off[2] = {g->ox, g->oy};
    const double zero[2] = {0.0, 0.0};
    for (int64_t a = 0; a < g->nangles; ++a) {
        const double c = cos(g->angles[a]), s = sin(g->angles[a]);
        const double dod = g->dsd[a] - g->dso[a];
        /* init_source_fn (:76-80), rotated about the origin (:206-214) */
        const double s0[2] = {g->det_offset, g->dso[a]};
        double sw[2];
        rot2(s0, c, s, zero, sw);
        /* init_det_origin_fn / init_delta_u_fn (:50-74): cells 0 and 1 before rotation */
        const double p0[2] = {-det_extent / 2.0 + g->du * 0.5, -dod};
        const double p1[2] = {-det_extent / 2.0 + g->du * 1.5, -dod};
        double r0[2], r1[2];
        rot2(p0, c, s, sw, r0); /* precompute_det_origin_fn (:38-48): rotate, minus source */
        rot2(p1, c, s, sw, r1);
        for (int d = 0; d < 2; ++d) {
            const double o = r0[d] / sp[d];
            origin[2 * a + d] = o;
            delta[2 * a + d] = r1[d] / sp[d] - o;                        /* :226-235 */
            src[2 * a + d] = (sw[d] - off[d] + 0.5 * ext[d]) / sp[d];   /* :28-36 */
        }
    }
}

/* ------------------------------------------------------------- 2D forward (A.2) ------- */

int64_t orc_ray2d(const orc_fan2d *g, const double *src, const double *origin,
                  const double *delta, int64_t a, int64_t k, double e[2], double v[2],
                  double *scale) {
    const double ro[2] = {src[2 * a], src[2 * a + 1]};
    const double rd[2] = {origin[2 * a] + (double)k * delta[2 * a],
                          origin[2 * a + 1] + (double)k * delta[2 * a + 1]}; /* fwd:52 */
    const double hi[2] = {(double)g->nx, (double)g->ny};
    double tmin = FLT_MIN, tmax = FLT_MAX; /* intersection.h:33-34 */
    for (int d = 0; d < 2; ++d) {
        const double t1 = (-1.0 - ro[d]) / rd[d];
        const double t2 = (hi[d] - ro[d]) / rd[d];
        tmin = fmin(fmax(t1, tmin), fmax(t2, tmin));
        tmax = fmax(fmin(t1, tmax), fmin(t2, tmax));
    }
    if (tmin > tmax - 1e-6) { /* forward_2d.hpp:58 */
        e[0] = e[1] = v[0] = v[1] = 0.0;
        *scale = 0.0;
        return 0;
    }
    const double r[2] = {rd[0] * (tmax - tmin), rd[1] * (tmax - tmin)};
    const double m = fmax(fabs(r[0]), fabs(r[1]));
    e[0] = ro[0] + tmin * rd[0];
    e[1] = ro[1] + tmin * rd[1];
    v[0] = r[0] / m;
    v[1] = r[1] / m;
    *scale = hypot(v[0] * g->sx, v[1] * g->sy); /* forward_2d.hpp:77 */
    return (int64_t)nearbyint(m);              /* round half to even, forward_2d.hpp:66 */
}

static void plan2d_alloc(const orc_fan2d *g, double **src, double **org, double **del) {
    const size_t n = (size_t)g->nangles * 2;
    *src = (double *)malloc(n * sizeof(double));
    *org = (double *)malloc(n * sizeof(double));
    *del = (double *)malloc(n * sizeof(double));
    orc_plan2d(g, *src, *org, *del);
}

void orc_fwd2d(const orc_fan2d *g, const double *img, double *sino, int64_t a0, int64_t a1,
               int64_t *nsamples) {
    double *src, *org, *del;
    plan2d_alloc(g, &src, &org, &del);
    int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int64_t a = a0; a < a1; ++a) {
        for (int64_t k = 0; k < g->det_count; ++k) {
            double e[2], v[2], scale;
            const int64_t n = orc_ray2d(g, src, org, del, a, k, e, v, &scale);
            double acc = 0.0;
            for (int64_t j = 0; j < n; ++j)
                acc += orc_bilerp(img, g->nx, g->ny, e[0] + (double)j * v[0],
                                  e[1] + (double)j * v[1]);
            sino[a * g->det_count + k] = acc * scale;
            total += n;
        }
    }
    if (nsamples) *nsamples = total;
    free(src);
    free(org);
    free(del);
}

void orc_adj2d(const orc_fan2d *g, const double *sino, double *img, int64_t a0, int64_t a1) {
    double *src, *org, *del;
    plan2d_alloc(g, &src, &org, &del);
    for (int64_t a = a0; a < a1; ++a) {
        for (int64_t k = 0; k < g->det_count; ++k) {
            double e[2], v[2], scale;
            const int64_t n = orc_ray2d(g, src, org, del, a, k, e, v, &scale);
            const double y = sino[a * g->det_count + k] * scale;
            if (y == 0.0) continue;
            for (int64_t j = 0; j < n; ++j)
                bilerp_scatter(img, g->nx, g->ny, e[0] + (double)j * v[0],
                               e[1] + (double)j * v[1], y);
        }
    }
    free(src);
    free(org);
    free(del);
}

/* ------------------------------------------------------ 2D backprojection (A.3) ------- */

/* Vector form of the voxel-driven backprojection: the reference's formula (backward_2d.hpp:58-99)
 * written for an arbitrary view. With n the unit normal of the detector line pointing from the
 * source towards it and e = u / |u|: depth = n . (x - s), magnification w = L / depth with
 * L = n . (c - s), detector cell coordinate w (e . (x - s)) / |u| + e . (s - c) / |u| + D/2
 * (the reference's w yi + D/2 when s, c, u come from a centred circular view), and the view's
 * weight w^wpow / |u| (the reference's w / du). */
static void bp2d_vec(const orc_fan2d *g, const double *sino, double *img, int64_t y0,
                     int64_t y1, int wpow) {
    const int64_t A = g->nangles, D = g->det_count;
    const double half_d = (double)D / 2.0;
#pragma omp parallel for schedule(static)
    for (int64_t iy = y0; iy < y1; ++iy) {
        for (int64_t ix = 0; ix < g->nx; ++ix) {
            /* the pixel's world position as the reference's backprojector writes it (:58-59) */
            const double x = ((double)ix - (double)g->nx / 2.0) * g->sx + g->ox + 0.5;
            const double y = ((double)iy - (double)g->ny / 2.0) * g->sy + g->oy + 0.5;
            double acc = 0.0;
            for (int64_t a = 0; a < A; ++a) {
                const double *v = g->vecs + 6 * a;
                const double ul = sqrt(v[4] * v[4] + v[5] * v[5]);
                const double e0 = v[4] / ul, e1 = v[5] / ul;
                double n0 = -e1, n1 = e0;
                const double cs0 = v[2] - v[0], cs1 = v[3] - v[1]; /* source to centre */
                if (n0 * cs0 + n1 * cs1 < 0.0) {
                    n0 = -n0;
                    n1 = -n1;
                }
                const double L = n0 * cs0 + n1 * cs1;
                const double px = x - v[0], py = y - v[1];
                const double w = L / (n0 * px + n1 * py);
                const double u = w * (e0 * px + e1 * py) / ul - (e0 * cs0 + e1 * cs1) / ul + half_d;
                const double wt = wpow == 2 ? w * w : w;
                acc += orc_lerp1(sino + a * D, D, u) * wt / ul;
            }
            img[iy * g->nx + ix] += acc;
        }
    }
}

void orc_bp2d_w(const orc_fan2d *g, const double *sino, double *img, int64_t y0, int64_t y1,
                int wpow) {
    if (g->vecs) {
        bp2d_vec(g, sino, img, y0, y1, wpow);
        return;
    }
    const int64_t A = g->nangles, D = g->det_count;
    double *sn = (double *)malloc((size_t)A * sizeof(double));
    double *cs = (double *)malloc((size_t)A * sizeof(double));
    for (int64_t a = 0; a < A; ++a) { /* meta = (-sin, cos, DSD, DSO), backward_2d.hpp:45-54 */
        sn[a] = -sin(g->angles[a]);
        cs[a] = cos(g->angles[a]);
    }
    const double inv_du = 1.0 / g->du; /* :35 */
    const double half_d = (double)D / 2.0;
#pragma omp parallel for schedule(static)
    for (int64_t iy = y0; iy < y1; ++iy) {
        for (int64_t ix = 0; ix < g->nx; ++ix) {
            /* world position of the pixel, +0.5 in world units as written (:58-59) */
            const double x = ((double)ix - (double)g->nx / 2.0) * g->sx + g->ox + 0.5;
            const double y = ((double)iy - (double)g->ny / 2.0) * g->sy + g->oy + 0.5;
            const double sx = x * inv_du, sy = y * inv_du;
            double acc = 0.0;
            for (int64_t a = 0; a < A; ++a) {
                const double xi = sn[a] * x - cs[a] * y;           /* :85 */
                const double yi = cs[a] * sx + sn[a] * sy;         /* :86 */
                const double w = g->dsd[a] / (xi + g->dso[a]);     /* :89 */
                const double u = w * yi + half_d;                  /* :92 */
                const double wt = wpow == 2 ? w * w : w;
                acc += orc_lerp1(sino + a * D, D, u) * wt;         /* :96 */
            }
            img[iy * g->nx + ix] += acc * inv_du;                  /* :99 */
        }
    }
    free(sn);
    free(cs);
}

void orc_bp2d(const orc_fan2d *g, const double *sino, double *img, int64_t y0, int64_t y1) {
    orc_bp2d_w(g, sino, img, y0, y1, 1);
}

/* Transpose of orc_bp2d_w: sino[a][c] += sum over pixels of the weight with which pixel p takes
 * cell c of projection a, times img[p]. The backprojection gives pixel p, per projection, wt
 * times the linear interpolation of the row at u (orc_lerp1: cells floor(u - 1/2) and the next,
 * weights 1 - al and al; cells outside the row contribute nothing), times 1/du (circular
 * plans; the view's 1/|u| for plans from views); the transpose adds img[p] times those weights
 * to the two cells. Projections a0 .. a1 - 1. */
void orc_bp2d_T_w(const orc_fan2d *g, const double *img, double *sino, int64_t a0, int64_t a1,
                  int wpow) {
    const int64_t D = g->det_count;
    const double half_d = (double)D / 2.0, inv_du = 1.0 / g->du;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t a = a0; a < a1; ++a) {
        double *row = sino + a * D;
        for (int64_t iy = 0; iy < g->ny; ++iy) {
            for (int64_t ix = 0; ix < g->nx; ++ix) {
                const double x = ((double)ix - (double)g->nx / 2.0) * g->sx + g->ox + 0.5;
                const double y = ((double)iy - (double)g->ny / 2.0) * g->sy + g->oy + 0.5;
                double u, wt;
                if (g->vecs) { /* bp2d_vec's view formula */
                    const double *v = g->vecs + 6 * a;
                    const double ul = sqrt(v[4] * v[4] + v[5] * v[5]);
                    const double e0 = v[4] / ul, e1 = v[5] / ul;
                    double n0 = -e1, n1 = e0;
                    const double cs0 = v[2] - v[0], cs1 = v[3] - v[1];
                    if (n0 * cs0 + n1 * cs1 < 0.0) {
                        n0 = -n0;
                        n1 = -n1;
                    }
                    const double L = n0 * cs0 + n1 * cs1;
                    const double px = x - v[0], py = y - v[1];
                    const double w = L / (n0 * px + n1 * py);
                    u = w * (e0 * px + e1 * py) / ul - (e0 * cs0 + e1 * cs1) / ul + half_d;
                    wt = (wpow == 2 ? w * w : w) / ul;
                } else { /* orc_bp2d_w's circular formula */
                    const double sn = -sin(g->angles[a]), cs = cos(g->angles[a]);
                    const double xi = sn * x - cs * y;
                    const double yi = cs * (x * inv_du) + sn * (y * inv_du);
                    const double w = g->dsd[a] / (xi + g->dso[a]);
                    u = w * yi + half_d;
                    wt = (wpow == 2 ? w * w : w) * inv_du;
                }
                const double ub = u - 0.5, fl = floor(ub), al = ub - fl;
                const int64_t i = (int64_t)fl;
                const double vw = img[iy * g->nx + ix] * wt;
                if (i >= 0 && i < D) row[i] += (1.0 - al) * vw;
                if (i + 1 >= 0 && i + 1 < D) row[i + 1] += al * vw;
            }
        }
    }
}

/* ------------------------------------------------ FP32 ports for the CPU baseline ----- */

static inline float fetch2f(const float *img, int64_t nx, int64_t ny, int64_t i, int64_t j) {
    return (i >= 0 && i < nx && j >= 0 && j < ny) ? img[j * nx + i] : 0.0f;
}

static inline float bilerpf(const float *img, int64_t nx, int64_t ny, float x, float y) {
    const float xb = x - 0.5f, yb = y - 0.5f;
    const float fx = floorf(xb), fy = floorf(yb);
    const float a = xb - fx, b = yb - fy;
    const int64_t i = (int64_t)fx, j = (int64_t)fy;
    const float r0 = (1.0f - a) * fetch2f(img, nx, ny, i, j) + a * fetch2f(img, nx, ny, i + 1, j);
    const float r1 =
        (1.0f - a) * fetch2f(img, nx, ny, i, j + 1) + a * fetch2f(img, nx, ny, i + 1, j + 1);
    return (1.0f - b) * r0 + b * r1;
}

void orc_fwd2d_f32(const orc_fan2d *g, const float *img, float *sino, int64_t a0, int64_t a1) {
    double *src, *org, *del;
    plan2d_alloc(g, &src, &org, &del);
    const float hi[2] = {(float)g->nx, (float)g->ny};
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t a = a0; a < a1; ++a) {
        const float ro[2] = {(float)src[2 * a], (float)src[2 * a + 1]};
        const float og[2] = {(float)org[2 * a], (float)org[2 * a + 1]};
        const float dl[2] = {(float)del[2 * a], (float)del[2 * a + 1]};
        for (int64_t k = 0; k < g->det_count; ++k) {
            const float rd[2] = {og[0] + (float)k * dl[0], og[1] + (float)k * dl[1]};
            float tmin = FLT_MIN, tmax = FLT_MAX;
            for (int d = 0; d < 2; ++d) {
                const float t1 = (-1.0f - ro[d]) / rd[d];
                const float t2 = (hi[d] - ro[d]) / rd[d];
                tmin = fminf(fmaxf(t1, tmin), fmaxf(t2, tmin));
                tmax = fmaxf(fminf(t1, tmax), fminf(t2, tmax));
            }
            if (tmin > tmax - 1e-6) {
                sino[a * g->det_count + k] = 0.0f;
                continue;
            }
            float p[2] = {ro[0] + tmin * rd[0], ro[1] + tmin * rd[1]};
            const float r[2] = {rd[0] * (tmax - tmin), rd[1] * (tmax - tmin)};
            const float m = fmaxf(fabsf(r[0]), fabsf(r[1]));
            const int n = (int)nearbyintf(m);
            const float v[2] = {r[0] / m, r[1] / m};
            float acc = 0.0f;
            for (int j = 0; j < n; ++j) {
                acc += bilerpf(img, g->nx, g->ny, p[0], p[1]);
                p[0] += v[0];
                p[1] += v[1];
            }
            sino[a * g->det_count + k] = acc * hypotf(v[0] * (float)g->sx, v[1] * (float)g->sy);
        }
    }
    free(src);
    free(org);
    free(del);
}

void orc_bp2d_f32(const orc_fan2d *g, const float *sino, float *img, int64_t y0, int64_t y1) {
    const int64_t A = g->nangles, D = g->det_count;
    float *meta = (float *)malloc((size_t)A * 4 * sizeof(float));
    for (int64_t a = 0; a < A; ++a) {
        meta[4 * a + 0] = (float)-sin(g->angles[a]);
        meta[4 * a + 1] = (float)cos(g->angles[a]);
        meta[4 * a + 2] = (float)g->dsd[a];
        meta[4 * a + 3] = (float)g->dso[a];
    }
    const float inv_du = 1.0f / (float)g->du;
    const float half_d = (float)D / 2.0f;
#pragma omp parallel for schedule(static)
    for (int64_t iy = y0; iy < y1; ++iy) {
        for (int64_t ix = 0; ix < g->nx; ++ix) {
            const float x = ((float)ix - (float)g->nx / 2.0f) * (float)g->sx + (float)g->ox + 0.5f;
            const float y = ((float)iy - (float)g->ny / 2.0f) * (float)g->sy + (float)g->oy + 0.5f;
            const float sx = x * inv_du, sy = y * inv_du;
            float acc = 0.0f;
            for (int64_t a = 0; a < A; ++a) {
                const float *m = meta + 4 * a;
                const float xi = m[0] * x - m[1] * y;
                const float yi = m[1] * sx + m[0] * sy;
                const float w = m[2] / (xi + m[3]);
                const float ub = w * yi + half_d - 0.5f;
                const float fl = floorf(ub);
                const float al = ub - fl;
                const int64_t i = (int64_t)fl;
                const float *row = sino + a * D;
                const float t0 = (i >= 0 && i < D) ? row[i] : 0.0f;
                const float t1 = (i + 1 >= 0 && i + 1 < D) ? row[i + 1] : 0.0f;
                acc += ((1.0f - al) * t0 + al * t1) * w;
            }
            img[iy * g->nx + ix] += acc * inv_du;
        }
    }
    free(meta);
}

/* ------------------------------------------------------------ 3D rotations ------------ */

/* rotate_yzy (rotation.hpp:7-28): the matrix written in its comment, applied to v. */
void orc_rotate_yzy(const double v[3], double phi, double theta, double psi, double out[3]) {
    const double cf = cos(phi), sf = sin(phi), ct = cos(theta), st = sin(theta);
    const double cp = cos(psi), sp = sin(psi);
    const double m[3][3] = {
        {cf * cp * ct - sf * sp, -st * cf, sf * cp + sp * cf * ct},
        {st * cp, ct, sp * st},
        {-sf * cp * ct - sp * cf, sf * st, -sf * sp * ct + cp * cf},
    };
    for (int r = 0; r < 3; ++r) out[r] = m[r][0] * v[0] + m[r][1] * v[1] + m[r][2] * v[2];
}

/* rotate_roll_pitch_yaw (rotation.hpp:31-48). */
void orc_rotate_rpy(const double v[3], double roll, double pitch, double yaw, double out[3]) {
    const double cr = cos(roll), sr = sin(roll), cp = cos(pitch), sp = sin(pitch);
    const double cy = cos(yaw), sy = sin(yaw);
    const double m[3][3] = {
        {cr * cp, cr * sp * sy - sr * cy, cr * sp * cy + sr * sy},
        {sr * cp, sr * sp * sy + cr * cy, sr * sp * cy - cr * sy},
        {-sp, cp * sy, cp * cy},
    };
    for (int r = 0; r < 3; ++r) out[r] = m[r][0] * v[0] + m[r][1] * v[1] + m[r][2] * v[2];
}

/* ------------------------------------------------------------- 3D forward (A.4) ------- */

/* Views (world points and steps) into the voxel frame: w -> (w - off + n sp / 2 - sp / 2) / sp
 * (forward_3d.hpp:158-169), steps divided by the spacing. */
static void plan3d_views(const orc_cone3d *g, double *origin, double *du, double *dv,
                         double *src) {
    const double sp[3] = {g->sx, g->sy, g->sz}, off[3] = {g->ox, g->oy, g->oz};
    const double n[3] = {(double)g->nx, (double)g->ny, (double)g->nz};
    const double hu = ((double)g->nu - 1.0) / 2.0, hv = ((double)g->nv - 1.0) / 2.0;
    for (int64_t a = 0; a < g->nangles; ++a) {
        const double *v = g->vecs + 12 * a;
        for (int d = 0; d < 3; ++d) {
            const double tr = n[d] * sp[d] / 2.0 - sp[d] / 2.0;
            const double p0 = v[3 + d] - hu * v[6 + d] - hv * v[9 + d]; /* pixel (0, 0) */
            origin[3 * a + d] = (p0 - off[d] + tr) / sp[d];
            du[3 * a + d] = v[6 + d] / sp[d];
            dv[3 * a + d] = v[9 + d] / sp[d];
            src[3 * a + d] = (v[d] - off[d] + tr) / sp[d];
        }
    }
}

void orc_plan3d_fwd(const orc_cone3d *g, double *origin, double *du, double *dv, double *src) {
    if (g->vecs) {
        plan3d_views(g, origin, du, dv, src);
        return;
    }
    const double dod = g->dsd - g->dso;
    const double sp[3] = {g->sx, g->sy, g->sz}, off[3] = {g->ox, g->oy, g->oz};
    const double n[3] = {(double)g->nx, (double)g->ny, (double)g->nz};
    for (int64_t a = 0; a < g->nangles; ++a) {
        /* forward_3d.hpp:120-130: source, pixel (0,0), (1,0), (0,1) before any rotation */
        double pts[4][3] = {
            {-g->du * ((double)g->nu / 2.0) + g->du * 0.5,
             -g->dv * ((double)g->nv / 2.0) + g->dv * 0.5, 0.0},
            {0, 0, 0},
            {0, 0, 0},
            {0.0, 0.0, -g->dso},
        };
        memcpy(pts[1], pts[0], sizeof pts[0]);
        memcpy(pts[2], pts[0], sizeof pts[0]);
        pts[1][0] += g->du;
        pts[2][1] += g->dv;
        for (int q = 0; q < 3; ++q) { /* detector roll/pitch/yaw, then +DOD (:135-146) */
            double t[3];
            orc_rotate_rpy(pts[q], g->roll, g->pitch, g->yaw, t);
            t[2] += dod;
            memcpy(pts[q], t, sizeof t);
        }
        const double corx = g->cor * sin(g->phi[a]) / g->sx; /* :173-175 */
        const double corz = g->cor * cos(g->phi[a]) / g->sz;
        for (int q = 0; q < 4; ++q) {
            double t[3];
            orc_rotate_yzy(pts[q], g->phi[a], g->theta[a], g->psi[a], t); /* :149-156 */
            for (int d = 0; d < 3; ++d) {
                const double tr = n[d] * sp[d] / 2.0 - sp[d] / 2.0;         /* :159 */
                pts[q][d] = (t[d] - off[d] + tr) / sp[d];                   /* :160-169 */
            }
            pts[q][0] += corx;
            pts[q][2] += corz;
        }
        for (int d = 0; d < 3; ++d) { /* :187-190 */
            origin[3 * a + d] = pts[0][d];
            du[3 * a + d] = pts[1][d] - pts[0][d];
            dv[3 * a + d] = pts[2][d] - pts[0][d];
            src[3 * a + d] = pts[3][d];
        }
    }
}

int64_t orc_ray3d(const orc_cone3d *g, const double *origin, const double *du,
                  const double *dv, const double *src, int64_t a, int64_t iu, int64_t iv,
                  double s[3], double dir[3], double *tmin_out, double *dt_out) {
    double nrm2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double q = origin[3 * a + d] + (double)iu * du[3 * a + d] + (double)iv * dv[3 * a + d];
        s[d] = src[3 * a + d];
        dir[d] = q - s[d]; /* forward_3d.hpp:58-63 */
        nrm2 += dir[d] * dir[d];
    }
    const double nsamples = ceil(sqrt(nrm2)); /* accuracy 1 (:68, :232) */
    const double hi[3] = {(double)g->nx + 1.0, (double)g->ny + 1.0, (double)g->nz + 1.0};
    double tmin = FLT_MIN, tmax = FLT_MAX; /* intersection.h:15-16, 19-25 against [-1, n+1]^3 */
    for (int d = 0; d < 3; ++d) {
        const double t1 = (-1.0 - s[d]) / dir[d];
        const double t2 = (hi[d] - s[d]) / dir[d];
        tmin = fmin(fmax(t1, tmin), fmax(t2, tmin));
        tmax = fmax(fmin(t1, tmax), fmin(t2, tmax));
    }
    if (!(tmin <= tmax)) return -1; /* :74-76 */
    const double nst = ceil((tmax - tmin) * nsamples); /* :80 */
    *tmin_out = tmin;
    *dt_out = nst > 0.0 ? (tmax - tmin) / nst : 0.0; /* :81; nst = 0 takes one sample */
    return (int64_t)nst;
}

void orc_fwd3d(const orc_cone3d *g, const double *vol, double *sino, int64_t a0, int64_t a1,
               int64_t *nsamples) {
    const size_t n3 = (size_t)g->nangles * 3;
    double *org = malloc(n3 * sizeof(double)), *du = malloc(n3 * sizeof(double));
    double *dv = malloc(n3 * sizeof(double)), *src = malloc(n3 * sizeof(double));
    orc_plan3d_fwd(g, org, du, dv, src);
    int64_t total = 0;
#pragma omp parallel for collapse(2) schedule(dynamic, 4) reduction(+ : total)
    for (int64_t a = a0; a < a1; ++a) {
        for (int64_t iv = 0; iv < g->nv; ++iv) {
            for (int64_t iu = 0; iu < g->nu; ++iu) {
                double s[3], dir[3], tmin, dt;
                const int64_t nst = orc_ray3d(g, org, du, dv, src, a, iu, iv, s, dir, &tmin, &dt);
                double acc = 0.0;
                for (int64_t i = 0; i <= nst; ++i) {
                    const double t = tmin + (double)i * dt;
                    /* tex3D at p + 0.5 (:88-89) */
                    acc += orc_trilerp(vol, g->nx, g->ny, g->nz, s[0] + t * dir[0] + 0.5,
                                       s[1] + t * dir[1] + 0.5, s[2] + t * dir[2] + 0.5);
                }
                sino[(a * g->nv + iv) * g->nu + iu] = acc; /* unscaled (:92-93) */
                total += nst + 1;
            }
        }
    }
    if (nsamples) *nsamples = total;
    free(org);
    free(du);
    free(dv);
    free(src);
}

void orc_adj3d(const orc_cone3d *g, const double *sino, double *vol, int64_t a0, int64_t a1) {
    const size_t n3 = (size_t)g->nangles * 3;
    double *org = malloc(n3 * sizeof(double)), *du = malloc(n3 * sizeof(double));
    double *dv = malloc(n3 * sizeof(double)), *src = malloc(n3 * sizeof(double));
    orc_plan3d_fwd(g, org, du, dv, src);
    for (int64_t a = a0; a < a1; ++a)
        for (int64_t iv = 0; iv < g->nv; ++iv)
            for (int64_t iu = 0; iu < g->nu; ++iu) {
                const double y = sino[(a * g->nv + iv) * g->nu + iu];
                if (y == 0.0) continue;
                double s[3], dir[3], tmin, dt;
                const int64_t nst = orc_ray3d(g, org, du, dv, src, a, iu, iv, s, dir, &tmin, &dt);
                for (int64_t i = 0; i <= nst; ++i) {
                    const double t = tmin + (double)i * dt;
                    trilerp_scatter(vol, g->nx, g->ny, g->nz, s[0] + t * dir[0] + 0.5,
                                    s[1] + t * dir[1] + 0.5, s[2] + t * dir[2] + 0.5, y);
                }
            }
    free(org);
    free(du);
    free(dv);
    free(src);
}

/* ------------------------------------------------- 3D backprojection (A.5) ------------ */

void orc_plan3d_bp(const orc_cone3d *g, double *vorigin, double *vdx, double *vdy, double *vdz,
                   double srcout[3]) {
    const double dod = g->dsd - g->dso;
    const double sp[3] = {g->sx, g->sy, g->sz};
    const double n[3] = {(double)g->nx, (double)g->ny, (double)g->nz};
    const double off[3] = {g->ox, g->oy, g->oz};
    for (int64_t a = 0; a < g->nangles; ++a) {
        double base[3];
        for (int d = 0; d < 3; ++d) base[d] = -n[d] * sp[d] / 2.0 + sp[d] / 2.0 + off[d]; /* :124 */
        double pts[4][3];
        for (int q = 0; q < 4; ++q) {
            memcpy(pts[q], base, sizeof base);
            if (q > 0) pts[q][q - 1] += sp[q - 1]; /* :131-139 */
            /* fn (:122-139): trajectory rotation, detector offset, roll/pitch/yaw about z=DOD */
            double t[3], r[3];
            orc_rotate_yzy(pts[q], g->phi[a], g->theta[a], g->psi[a], t);
            t[0] += g->off_u;
            t[1] += g->off_v;
            t[2] -= dod;
            orc_rotate_rpy(t, g->roll, g->pitch, g->yaw, r);
            r[2] += dod;
            memcpy(pts[q], r, sizeof r);
        }
        for (int d = 0; d < 3; ++d) {
            vorigin[3 * a + d] = pts[0][d];
            vdx[3 * a + d] = pts[1][d] - pts[0][d];
            vdy[3 * a + d] = pts[2][d] - pts[0][d];
            vdz[3 * a + d] = pts[3][d] - pts[0][d];
        }
    }
    /* source (measurement_3d.hpp:122-124) with the detector rotation about z = DOD
     * (backward_3d.hpp:218-227) */
    double s0[3] = {-g->off_u, -g->off_v, -g->dso - dod}, r[3];
    orc_rotate_rpy(s0, g->roll, g->pitch, g->yaw, r);
    r[2] += dod;
    memcpy(srcout, r, sizeof r);
}

static inline double fetch_panel(const double *p, int64_t nu, int64_t nv, int64_t i, int64_t j) {
    return (i >= 0 && i < nu && j >= 0 && j < nv) ? p[j * nu + i] : 0.0;
}

/* Bilinear panel sample at texel coordinates (u, v): texel (i, j) centred at (i+.5, j+.5). */
static inline double panel_bilerp(const double *p, int64_t nu, int64_t nv, double u, double v) {
    const double ub = u - 0.5, vb = v - 0.5;
    const double fu = floor(ub), fv = floor(vb);
    const double a = ub - fu, b = vb - fv;
    const int64_t i = (int64_t)fu, j = (int64_t)fv;
    return (1.0 - b) * ((1.0 - a) * fetch_panel(p, nu, nv, i, j) + a * fetch_panel(p, nu, nv, i + 1, j)) +
           b * ((1.0 - a) * fetch_panel(p, nu, nv, i, j + 1) + a * fetch_panel(p, nu, nv, i + 1, j + 1));
}

static void bp3d_matched(const orc_cone3d *g, const double *sino, double *vol, int64_t z0,
                         int64_t z1, int fdk);

void orc_bp3d(const orc_cone3d *g, const double *sino, double *vol, int64_t z0, int64_t z1) {
    if (g->vecs) { /* views: the reference's frame exists for Euler angles only */
        bp3d_matched(g, sino, vol, z0, z1, 0);
        return;
    }
    const size_t n3 = (size_t)g->nangles * 3;
    double *O = malloc(n3 * sizeof(double)), *dx = malloc(n3 * sizeof(double));
    double *dy = malloc(n3 * sizeof(double)), *dz = malloc(n3 * sizeof(double));
    double s[3];
    orc_plan3d_bp(g, O, dx, dy, dz, s);
    const double plane = g->dso - g->dsd;             /* backward_3d.hpp:85 */
    const double hu = (double)(g->nu / 2), hv = (double)(g->nv / 2); /* integer halves, :89-90 */
    const int64_t UV = g->nu * g->nv;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t iz = z0; iz < z1; ++iz)
        for (int64_t iy = 0; iy < g->ny; ++iy)
            for (int64_t ix = 0; ix < g->nx; ++ix) {
                double acc = 0.0;
                for (int64_t a = 0; a < g->nangles; ++a) {
                    double dir[3];
                    for (int d = 0; d < 3; ++d) /* :79-82 */
                        dir[d] = O[3 * a + d] + (double)ix * dx[3 * a + d] +
                                 (double)iy * dy[3 * a + d] + (double)iz * dz[3 * a + d] - s[d];
                    const double t = (plane - s[2]) / dir[2];
                    const double u = dir[0] * t + s[0] + hu;
                    const double v = dir[1] * t + s[1] + hv;
                    acc += panel_bilerp(sino + a * UV, g->nu, g->nv, u, v); /* layer a+0.5 */
                }
                vol[(iz * g->ny + iy) * g->nx + ix] += acc;
            }
    free(O);
    free(dx);
    free(dy);
    free(dz);
}

/* Solve M x = w for the 3x3 matrix with columns c0, c1, c2 (Cramer's rule). */
static void solve3(const double c0[3], const double c1[3], const double c2[3], const double w[3],
                   double x[3]) {
#define DET3(a, b, c)                                                                         \
    ((a)[0] * ((b)[1] * (c)[2] - (b)[2] * (c)[1]) - (b)[0] * ((a)[1] * (c)[2] - (a)[2] * (c)[1]) + \
     (c)[0] * ((a)[1] * (b)[2] - (a)[2] * (b)[1]))
    const double det = DET3(c0, c1, c2);
    x[0] = DET3(w, c1, c2) / det;
    x[1] = DET3(c0, w, c2) / det;
    x[2] = DET3(c0, c1, w) / det;
#undef DET3
}

static void bp3d_matched(const orc_cone3d *g, const double *sino, double *vol, int64_t z0,
                         int64_t z1, int fdk) {
    const size_t n3 = (size_t)g->nangles * 3;
    double *org = malloc(n3 * sizeof(double)), *du = malloc(n3 * sizeof(double));
    double *dv = malloc(n3 * sizeof(double)), *src = malloc(n3 * sizeof(double));
    orc_plan3d_fwd(g, org, du, dv, src);
    const int64_t UV = g->nu * g->nv;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t iz = z0; iz < z1; ++iz)
        for (int64_t iy = 0; iy < g->ny; ++iy)
            for (int64_t ix = 0; ix < g->nx; ++ix) {
                double acc = 0.0;
                for (int64_t a = 0; a < g->nangles; ++a) {
                    /* voxel (ix,iy,iz) lies at src + t (origin + iu du + iv dv - src) */
                    double c0[3], w[3], x[3];
                    const double P[3] = {(double)ix, (double)iy, (double)iz};
                    for (int d = 0; d < 3; ++d) {
                        c0[d] = org[3 * a + d] - src[3 * a + d];
                        w[d] = P[d] - src[3 * a + d];
                    }
                    solve3(c0, du + 3 * a, dv + 3 * a, w, x);
                    if (!(x[0] > 0.0)) continue;
                    const double iu = x[1] / x[0], iv = x[2] / x[0];
                    double wt = 1.0;
                    if (fdk) {
                        const double r = g->dso / (x[0] * g->dsd);
                        wt = r * r;
                    }
                    acc += wt * panel_bilerp(sino + a * UV, g->nu, g->nv, iu + 0.5, iv + 0.5);
                }
                vol[(iz * g->ny + iy) * g->nx + ix] += acc;
            }
    free(org);
    free(du);
    free(dv);
    free(src);
}

void orc_bp3d_matched(const orc_cone3d *g, const double *sino, double *vol, int64_t z0,
                      int64_t z1) {
    bp3d_matched(g, sino, vol, z0, z1, 0);
}

void orc_bp3d_fdk(const orc_cone3d *g, const double *sino, double *vol, int64_t z0, int64_t z1) {
    bp3d_matched(g, sino, vol, z0, z1, 1);
}

/* Transpose of the 3D backprojections: sino[a][v][u] += the weight with which each voxel
 * takes panel cell (u, v) of projection a, times vol[voxel]. The backprojection gives a voxel,
 * per projection, wt times the bilinear interpolation of the panel at (u, v) (panel_bilerp:
 * cells floor(u - 1/2), floor(v - 1/2) and the next ones, cells outside the panel read zero);
 * (u, v) and wt are orc_bp3d's (mode 0; plans from views project as matched, as there),
 * bp3d_matched's (mode 1) or its FDK form (mode 2), restated below per voxel. Projections
 * a0 .. a1 - 1. */
void orc_bp3d_T(const orc_cone3d *g, int mode, const double *vol, double *sino, int64_t a0,
                int64_t a1) {
    const size_t n3 = (size_t)g->nangles * 3;
    const int ref = mode == 0 && !g->vecs;
    double *O = NULL, *dx = NULL, *dy = NULL, *dz = NULL, s[3] = {0, 0, 0};
    double *org = NULL, *du = NULL, *dv = NULL, *src = NULL;
    if (ref) {
        O = malloc(n3 * sizeof(double));
        dx = malloc(n3 * sizeof(double));
        dy = malloc(n3 * sizeof(double));
        dz = malloc(n3 * sizeof(double));
        orc_plan3d_bp(g, O, dx, dy, dz, s);
    } else {
        org = malloc(n3 * sizeof(double));
        du = malloc(n3 * sizeof(double));
        dv = malloc(n3 * sizeof(double));
        src = malloc(n3 * sizeof(double));
        orc_plan3d_fwd(g, org, du, dv, src);
    }
    const double plane = g->dso - g->dsd;
    const double hu = (double)(g->nu / 2), hv = (double)(g->nv / 2);
    const int64_t UV = g->nu * g->nv;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t a = a0; a < a1; ++a) {
        double *panel = sino + a * UV;
        for (int64_t iz = 0; iz < g->nz; ++iz)
            for (int64_t iy = 0; iy < g->ny; ++iy)
                for (int64_t ix = 0; ix < g->nx; ++ix) {
                    double u, v, wt = 1.0;
                    if (ref) { /* orc_bp3d */
                        double dir[3];
                        for (int d = 0; d < 3; ++d)
                            dir[d] = O[3 * a + d] + (double)ix * dx[3 * a + d] +
                                     (double)iy * dy[3 * a + d] + (double)iz * dz[3 * a + d] - s[d];
                        const double t = (plane - s[2]) / dir[2];
                        u = dir[0] * t + s[0] + hu;
                        v = dir[1] * t + s[1] + hv;
                    } else { /* bp3d_matched */
                        double c0[3], w[3], x[3];
                        const double P[3] = {(double)ix, (double)iy, (double)iz};
                        for (int d = 0; d < 3; ++d) {
                            c0[d] = org[3 * a + d] - src[3 * a + d];
                            w[d] = P[d] - src[3 * a + d];
                        }
                        solve3(c0, du + 3 * a, dv + 3 * a, w, x);
                        if (!(x[0] > 0.0)) continue;
                        u = x[1] / x[0] + 0.5;
                        v = x[2] / x[0] + 0.5;
                        if (mode == 2) {
                            const double r = g->dso / (x[0] * g->dsd);
                            wt = r * r;
                        }
                    }
                    const double val = wt * vol[(iz * g->ny + iy) * g->nx + ix];
                    const double ub = u - 0.5, vb = v - 0.5;
                    const double fu = floor(ub), fv = floor(vb);
                    const double al = ub - fu, be = vb - fv;
                    const int64_t i = (int64_t)fu, j = (int64_t)fv;
                    const double w4[4] = {(1.0 - be) * (1.0 - al), (1.0 - be) * al,
                                          be * (1.0 - al), be * al};
                    for (int k = 0; k < 4; ++k) {
                        const int64_t ci = i + (k & 1), cj = j + (k >> 1);
                        if (ci >= 0 && ci < g->nu && cj >= 0 && cj < g->nv)
                            panel[cj * g->nu + ci] += w4[k] * val;
                    }
                }
    }
    free(O);
    free(dx);
    free(dy);
    free(dz);
    free(org);
    free(du);
    free(dv);
    free(src);
}

/* ------------------------------------------------ 3D FP32 ports (CPU baseline only) --- */

static inline float fetch3f(const float *v, int64_t nx, int64_t ny, int64_t nz, int64_t i,
                            int64_t j, int64_t k) {
    return (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) ? v[(k * ny + j) * nx + i]
                                                                       : 0.0f;
}

static inline float trilerpf(const float *v, int64_t nx, int64_t ny, int64_t nz, float x,
                             float y, float z) {
    const float xb = x - 0.5f, yb = y - 0.5f, zb = z - 0.5f;
    const float fx = floorf(xb), fy = floorf(yb), fz = floorf(zb);
    const float a = xb - fx, b = yb - fy, c = zb - fz;
    const int64_t i = (int64_t)fx, j = (int64_t)fy, k = (int64_t)fz;
    float acc = 0.0f;
    for (int dk = 0; dk < 2; ++dk) {
        const float wz = dk ? c : 1.0f - c;
        for (int dj = 0; dj < 2; ++dj) {
            const float wy = dj ? b : 1.0f - b;
            const float r = (1.0f - a) * fetch3f(v, nx, ny, nz, i, j + dj, k + dk) +
                            a * fetch3f(v, nx, ny, nz, i + 1, j + dj, k + dk);
            acc += wz * wy * r;
        }
    }
    return acc;
}

/* kernel_forward_3d (forward_3d.hpp:57-93) in FP32 over projections [a0, a1): the per-angle
 * frame from orc_plan3d_fwd rounded to FP32, the reference's float ray set-up, slab clip and
 * float-carried march. */
void orc_fwd3d_f32(const orc_cone3d *g, const float *vol, float *sino, int64_t a0, int64_t a1) {
    const size_t n3 = (size_t)g->nangles * 3;
    double *org = malloc(n3 * sizeof(double)), *du = malloc(n3 * sizeof(double));
    double *dv = malloc(n3 * sizeof(double)), *src = malloc(n3 * sizeof(double));
    orc_plan3d_fwd(g, org, du, dv, src);
    const float hi[3] = {(float)g->nx + 1.0f, (float)g->ny + 1.0f, (float)g->nz + 1.0f};
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int64_t a = a0; a < a1; ++a) {
        for (int64_t iv = 0; iv < g->nv; ++iv) {
            for (int64_t iu = 0; iu < g->nu; ++iu) {
                float s[3], d[3], nrm2 = 0.0f;
                for (int k = 0; k < 3; ++k) {
                    const float q = (float)org[3 * a + k] + (float)iu * (float)du[3 * a + k] +
                                    (float)iv * (float)dv[3 * a + k];
                    s[k] = (float)src[3 * a + k];
                    d[k] = q - s[k];
                    nrm2 += d[k] * d[k];
                }
                const float nsamples = ceilf(sqrtf(nrm2));
                float tmin = FLT_MIN, tmax = FLT_MAX;
                for (int k = 0; k < 3; ++k) {
                    const float t1 = (-1.0f - s[k]) / d[k], t2 = (hi[k] - s[k]) / d[k];
                    tmin = fminf(fmaxf(t1, tmin), fmaxf(t2, tmin));
                    tmax = fmaxf(fminf(t1, tmax), fminf(t2, tmax));
                }
                float acc = 0.0f;
                if (tmin <= tmax) {
                    const float nst = ceilf((tmax - tmin) * nsamples);
                    const float dt = nst > 0.0f ? (tmax - tmin) / nst : 1.0f;
                    for (float t = tmin; t <= tmax; t += dt)
                        acc += trilerpf(vol, g->nx, g->ny, g->nz, s[0] + t * d[0] + 0.5f,
                                        s[1] + t * d[1] + 0.5f, s[2] + t * d[2] + 0.5f);
                }
                sino[(a * g->nv + iv) * g->nu + iu] = acc;
            }
        }
    }
    free(org);
    free(du);
    free(dv);
    free(src);
}

/* kernel_backprojection_3d (backward_3d.hpp:28-108) in FP32 over slices [z0, z1): the frame of
 * orc_plan3d_bp rounded to FP32, projection onto z = DSO - DSD, bilinear layer sample,
 * unweighted sum. */
void orc_bp3d_f32(const orc_cone3d *g, const float *sino, float *vol, int64_t z0, int64_t z1) {
    const size_t n3 = (size_t)g->nangles * 3;
    double *O = malloc(n3 * sizeof(double)), *dx = malloc(n3 * sizeof(double));
    double *dy = malloc(n3 * sizeof(double)), *dz = malloc(n3 * sizeof(double));
    double sd[3];
    orc_plan3d_bp(g, O, dx, dy, dz, sd);
    const float s[3] = {(float)sd[0], (float)sd[1], (float)sd[2]};
    const float plane = (float)(g->dso - g->dsd);
    const float hu = (float)(g->nu / 2), hv = (float)(g->nv / 2);
    const int64_t UV = g->nu * g->nv;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t iz = z0; iz < z1; ++iz)
        for (int64_t iy = 0; iy < g->ny; ++iy)
            for (int64_t ix = 0; ix < g->nx; ++ix) {
                float acc = 0.0f;
                for (int64_t a = 0; a < g->nangles; ++a) {
                    float dir[3];
                    for (int k = 0; k < 3; ++k)
                        dir[k] = (float)O[3 * a + k] + (float)ix * (float)dx[3 * a + k] +
                                 (float)iy * (float)dy[3 * a + k] +
                                 (float)iz * (float)dz[3 * a + k] - s[k];
                    const float t = (plane - s[2]) / dir[2];
                    const float u = dir[0] * t + s[0] + hu - 0.5f;
                    const float v = dir[1] * t + s[1] + hv - 0.5f;
                    const float fu = floorf(u), fv = floorf(v);
                    const float al = u - fu, be = v - fv;
                    const int64_t i = (int64_t)fu, j = (int64_t)fv;
                    const float *p = sino + a * UV;
#define PF(ii, jj) (((ii) >= 0 && (ii) < g->nu && (jj) >= 0 && (jj) < g->nv) ? p[(jj) * g->nu + (ii)] : 0.0f)
                    acc += (1.0f - be) * ((1.0f - al) * PF(i, j) + al * PF(i + 1, j)) +
                           be * ((1.0f - al) * PF(i, j + 1) + al * PF(i + 1, j + 1));
#undef PF
                }
                vol[(iz * g->ny + iy) * g->nx + ix] += acc;
            }
    free(O);
    free(dx);
    free(dy);
    free(dz);
}
