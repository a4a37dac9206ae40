/*
 * imrt_oracle.c - CPU restatement (float64) of the reference's photon
 * collapsed-cone convolution (CCC) dose path.  TEST INFRASTRUCTURE ONLY: it is
 * what the HIP photon kernels are compared against in the tests and the CPU
 * baseline of the photon benchmark; nothing in the product links or calls it.
 *
 * The reference has no JAX photon path, so the source is the CUDA extension
 * itself, with its texture semantics restated exactly (real-number trilinear
 * interpolation in place of the hardware's fixed-point filter weights):
 *
 *   orc_ph_raytrace  rayTraceKernel            CudaClasses.cu:180-228 on the
 *                    density texture (border AIR_DENSITY, IMRTClasses.cu:394)
 *   orc_ph_terma     termaKernel               IMRTClasses.cu:209-250
 *                    (offAxisFactors :152-171, headTransmission :120-150,
 *                    photon head frame with collimator :41-79)
 *   orc_ph_ccc       cccKernel                 IMRTClasses.cu:252-350
 *                    (head -> image :81-118, kernelTilt :173-203)
 *   orc_ph_head_transmission / orc_ph_tilt / orc_ph_frames  test hooks of the
 *                    helpers (head transmission, kernel tilt, head <-> image)
 *
 * Texture semantics (TextureClasses.cu:14-41, CudaClasses.cuh:90-104):
 * unnormalised coordinates tex = (p + iso) / s + 1/2, linear filtering, border
 * addressing: the continuous index c = tex - 1/2 is interpolated between
 * corners floor(c) and floor(c) + 1 per axis, a corner outside [0, n) reading
 * the border value (0 for TERMA, AIR_DENSITY for density).  The ray-trace
 * keeps a sample while every tex_a is in [0, n_a), i.e. c_a in [-1/2, n_a - 1/2).
 *
 * Float32 kept on purpose: the collapsed-cone ray length is the reference's
 * float32 counter (ray_length = L; while (ray_length >= 0) ... -= sp, with
 * sp = spacing / 10 in float32), because the number of steps depends on its
 * rounding.  Everything else is float64.
 *
 * Compile with -DORC_FLOAT for a float32 build (benchmark baseline only).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#ifdef ORC_FLOAT
typedef float real;
#define EXP expf
#define FLOOR floorf
#define SQRT sqrtf
#define ERF erff
#define SIN sinf
#define COS cosf
#else
typedef double real;
#define EXP exp
#define FLOOR floor
#define SQRT sqrt
#define ERF erf
#define SIN sin
#define COS cos
#endif

#define AIR_DENSITY 0.0012
#define PI_D 3.14159265358979323846

/* beam: isocentre minus origin, source (iso-relative, at the primary source
 * distance), sin/cos of the adjusted gantry, couch and collimator angles */
typedef struct {
    real iso[3], src[3];
    real sg, cg, sc, cc, sca, cca;
} pbeam_t;

static void load_beam(const double *g, pbeam_t *b) {
    for (int a = 0; a < 3; ++a) {
        b->iso[a] = (real)g[a];
        b->src[a] = (real)g[3 + a];
    }
    b->sg = (real)g[6]; b->cg = (real)g[7];
    b->sc = (real)g[8]; b->cc = (real)g[9];
    b->sca = (real)g[10]; b->cca = (real)g[11];
}

/* IMRTBeam::pointXYZImageToHead (IMRTClasses.cu:41-79): couch about y, gantry
 * about z, collimator about y, then (x, y, z)_head = (-x_c, z_c, y_c). */
static void img_to_head(const pbeam_t *b, const real p[3], real h[3]) {
    const real xt = p[0] * b->cc - p[2] * b->sc;
    const real yt = p[1];
    const real zt = p[0] * b->sc + p[2] * b->cc;
    const real xg = xt * b->cg + yt * b->sg;
    const real yg = -xt * b->sg + yt * b->cg;
    const real zg = zt;
    const real xc = xg * b->cca - zg * b->sca;
    const real yc = yg;
    const real zc = xg * b->sca + zg * b->cca;
    h[0] = -xc; h[1] = zc; h[2] = yc;
}

/* IMRTBeam::pointXYZHeadToImage (IMRTClasses.cu:81-118), the inverse. */
static void head_to_img(const pbeam_t *b, const real h[3], real p[3]) {
    const real xz = -h[0], yz = h[2], zz = h[1];
    const real xc = xz * b->cca + zz * b->sca;
    const real yc = yz;
    const real zc = -xz * b->sca + zz * b->cca;
    const real xg = xc * b->cg - yc * b->sg;
    const real yg = xc * b->sg + yc * b->cg;
    const real zg = zc;
    p[0] = xg * b->cc + zg * b->sc;
    p[1] = yg;
    p[2] = -xg * b->sc + zg * b->cc;
}

static void voxel_xyz(const pbeam_t *b, real s, int i, int j, int k, real v[3]) {
    v[0] = (real)i * s - b->iso[0];
    v[1] = (real)j * s - b->iso[1];
    v[2] = (real)k * s - b->iso[2];
}

static void unit_to_source(const pbeam_t *b, const real p[3], real u[3]) {
    const real dx = b->src[0] - p[0], dy = b->src[1] - p[1], dz = b->src[2] - p[2];
    const real r = 1.0 / SQRT(dx * dx + dy * dy + dz * dz);
    u[0] = dx * r; u[1] = dy * r; u[2] = dz * r;
}

/* Border-addressed trilinear sample of a float32 volume at continuous index c. */
static real sample(const float *vol, int nx, int ny, int nz, real border, real cx, real cy, real cz) {
    const real fx = FLOOR(cx), fy = FLOOR(cy), fz = FLOOR(cz);
    const long ix = (long)fx, iy = (long)fy, iz = (long)fz;
    const real ax = cx - fx, ay = cy - fy, az = cz - fz;
    real acc = 0.0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const long x = ix + dx, y = iy + dy, z = iz + dz;
                const real w = (dx ? ax : 1.0 - ax) * (dy ? ay : 1.0 - ay) * (dz ? az : 1.0 - az);
                const real val = (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz)
                                     ? (real)vol[x + (long)nx * (y + (long)ny * z)]
                                     : border;
                acc += w * val;
            }
    return acc;
}

static void voxel_ijk(int64_t v, int nx, int ny, int *i, int *j, int *k) {
    *i = (int)(v % nx);
    *j = (int)((v / nx) % ny);
    *k = (int)(v / ((int64_t)nx * ny));
}

/* rayTraceKernel (CudaClasses.cu:180-228) on the density texture: samples at
 * p_n = v + n u (1 mm steps toward the source) while tex is inside the grid;
 * WET = -0.05 + 0.1 sum max(rho, 0) (g/cm^2). */
static real raytrace_voxel(const float *rho, int nx, int ny, int nz, real s, const pbeam_t *b, int i, int j, int k) {
    real v[3], u[3];
    voxel_xyz(b, s, i, j, k, v);
    unit_to_source(b, v, u);
    real wet = -0.05;
    for (long n = 0;; ++n) {
        const real cx = (v[0] + u[0] * (real)n + b->iso[0]) / s;
        const real cy = (v[1] + u[1] * (real)n + b->iso[1]) / s;
        const real cz = (v[2] + u[2] * (real)n + b->iso[2]) / s;
        if (!(cx >= -0.5 && cx < nx - 0.5 && cy >= -0.5 && cy < ny - 0.5 && cz >= -0.5 && cz < nz - 0.5)) break;
        const real d = sample(rho, nx, ny, nz, AIR_DENSITY, cx, cy, cz);
        wet += 0.1 * (d > 0.0 ? d : 0.0);
    }
    return wet;
}

int orc_ph_raytrace(const float *rho, int nx, int ny, int nz, double s, const double *g, const int64_t *idx,
                    int64_t n_idx, double *out) {
    pbeam_t b;
    load_beam(g, &b);
    const int64_t n = idx ? n_idx : (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t t = 0; t < n; ++t) {
        int i, j, k;
        voxel_ijk(idx ? idx[t] : t, nx, ny, &i, &j, &k);
        out[t] = (double)raytrace_voxel(rho, nx, ny, nz, (real)s, &b, i, j, k);
    }
    return 0;
}

/* np.interp semantics of the lowerBound interpolation (IMRTClasses.cu:5-17,
 * :152-171): clamped at both ends, linear inside; knots ascending. */
static real interp(real x, const double *xd, const double *yd, int n) {
    if (x <= xd[0]) return (real)yd[0];
    if (x >= xd[n - 1]) return (real)yd[n - 1];
    int i = 1;
    while (i < n - 1 && x > xd[i]) ++i;
    const real f = (x - xd[i - 1]) / (xd[i] - xd[i - 1]);
    return (real)yd[i - 1] + f * ((real)yd[i] - (real)yd[i - 1]);
}

/* headTransmission (IMRTClasses.cu:120-150): fraction of a Gaussian source of
 * width sigma at distance iso_to_source seen from head-frame point h through
 * the leaf-pair apertures; mlc rows (x1, x2, y_offset, y_width) at the
 * isocentre plane, scaled to the MLC plane by (SAD - mlc_distance) / SAD. */
static real head_transmission(const real h[3], real sad_p, real mlc_dist, real iso_to_source, real sigma,
                              const double *mlc, int n_pairs) {
    const real mlc_scale = (sad_p - mlc_dist) / sad_p;
    const real div = (iso_to_source - h[2]) / (mlc_dist - h[2]);
    const real inv = 1.0 / (sigma * SQRT(2.0));
    real t = 0.0;
    for (int p = 0; p < n_pairs; ++p) {
        const real x1 = (real)mlc[4 * p], x2 = (real)mlc[4 * p + 1];
        const real yo = (real)mlc[4 * p + 2], yw = (real)mlc[4 * p + 3];
        const real yb = (yo - 0.5 * yw) * mlc_scale, yt = (yo + 0.5 * yw) * mlc_scale;
        const real xl = x1 * mlc_scale, xr = x2 * mlc_scale;
        const real tip1 = (xl - h[0]) * div + h[0], tip2 = (xr - h[0]) * div + h[0];
        const real e1 = (yb - h[1]) * div + h[1], e2 = (yt - h[1]) * div + h[1];
        const real ex = 0.5 * (ERF(tip2 * inv) - ERF(tip1 * inv));
        const real ey = 0.5 * (ERF(e2 * inv) - ERF(e1 * inv));
        t += ex * ey;
    }
    return t;
}

int orc_ph_head_transmission(const double *pts, int64_t n, double sad_p, double mlc_dist, double iso_to_source,
                             double sigma, const double *mlc, int n_pairs, double *out) {
    for (int64_t t = 0; t < n; ++t) {
        const real h[3] = {(real)pts[3 * t], (real)pts[3 * t + 1], (real)pts[3 * t + 2]};
        out[t] = (double)head_transmission(h, (real)sad_p, (real)mlc_dist, (real)iso_to_source, (real)sigma, mlc,
                                           n_pairs);
    }
    return 0;
}

/* Model scalars, in this order:
 *  0 primary_source_distance   1 scatter_source_distance  2 primary_source_size
 *  3 scatter_source_size       4 mlc_distance             5 scatter_source_weight
 *  6 electron_attenuation      7 electron_source_weight   8 electron_fitted_dmax
 *  9 mu_calibration
 * prof: radius[n_prof], intensities[n_prof], softening[n_prof];
 * spec: attenuation[n_spec], primary weights[n_spec], scatter weights[n_spec]. */
enum { M_SADP, M_SADS, M_SIGP, M_SIGS, M_MLCD, M_SSW, M_EA, M_ESW, M_EDMAX, M_MUCAL };

/* termaKernel (IMRTClasses.cu:209-250) at one voxel of WET wet (g/cm^2). */
static void terma_voxel(const pbeam_t *b, real s, int i, int j, int k, real wet, const double *m,
                        const double *prof, int n_prof, const double *spec, int n_spec, const double *mlc,
                        int n_pairs, real *terma_out, real *electron_out) {
    real v[3], h[3];
    voxel_xyz(b, s, i, j, k, v);
    img_to_head(b, v, h);
    const real sad_p = (real)m[M_SADP], sad_s = (real)m[M_SADS];
    const real d_p = sad_p - h[2], d_s = sad_s - h[2];
    /* offAxisFactors (IMRTClasses.cu:152-171) */
    const real r_cax = SQRT(h[0] * h[0] + h[1] * h[1]) * d_p / sad_p;
    const real oaf = interp(r_cax, prof, prof + n_prof, n_prof);
    const real soft = interp(r_cax, prof, prof + 2 * n_prof, n_prof);
    const real tp = head_transmission(h, sad_p, (real)m[M_MLCD], sad_p, (real)m[M_SIGP], mlc, n_pairs);
    const real ts = head_transmission(h, sad_p, (real)m[M_MLCD], sad_s, (real)m[M_SIGS], mlc, n_pairs);
    /* fminf(1, ts / tp): a NaN quotient (0 / 0) gives 1, as fminf returns its other operand */
    const real q = ts / tp;
    const real tr = (q == q && q < 1.0) ? q : 1.0;
    real terma = 0.0;
    for (int e = 0; e < n_spec; ++e) {
        const real att = EXP(-(real)spec[e] * wet * soft);
        terma += (1.0 - tr) * ((real)spec[n_spec + e] * att * (sad_p / d_p) * (sad_p / d_p)) +
                 tr * (real)spec[2 * n_spec + e] * att * (sad_s / d_s);
    }
    const real ea = (real)m[M_EA], dmax = (real)m[M_EDMAX];
    real el = (EXP(-ea * wet) - EXP(-ea * dmax)) / (1.0 - EXP(-ea * dmax));
    if (!(el > 0.0)) el = 0.0;
    const real ssw = (real)m[M_SSW];
    *terma_out = oaf * ((1.0 - ssw) * tp * terma + ssw * ts * terma);
    *electron_out = (real)m[M_ESW] * (0.4 + 0.3 * tr) * el * tp;
}

int orc_ph_terma(const double *wet, int nx, int ny, int nz, double s, const double *g, const double *m,
                 const double *prof, int n_prof, const double *spec, int n_spec, const double *mlc, int n_pairs,
                 const int64_t *idx, int64_t n_idx, double *terma, double *electron) {
    pbeam_t b;
    load_beam(g, &b);
    const int64_t n = idx ? n_idx : (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < n; ++t) {
        int i, j, k;
        voxel_ijk(idx ? idx[t] : t, nx, ny, &i, &j, &k);
        real te, el;
        terma_voxel(&b, (real)s, i, j, k, (real)wet[t], m, prof, n_prof, spec, n_spec, mlc, n_pairs, &te, &el);
        terma[t] = (double)te;
        electron[t] = (double)el;
    }
    return 0;
}

/* kernelTilt (IMRTClasses.cu:173-203): rotate an image-space vector by the
 * rotation about axis x uvec that takes the isocentre's unit vector to the
 * source onto the voxel's (Rodrigues' formula). */
static void tilt(const pbeam_t *b, const real v[3], real vec[3]) {
    real u[3], a[3];
    const real zero[3] = {0.0, 0.0, 0.0};
    unit_to_source(b, v, u);
    unit_to_source(b, zero, a);
    const real c = a[0] * u[0] + a[1] * u[1] + a[2] * u[2];
    const real k[3] = {a[1] * u[2] - a[2] * u[1], a[2] * u[0] - a[0] * u[2], a[0] * u[1] - a[1] * u[0]};
    real r[3] = {k[1] * vec[2] - k[2] * vec[1], k[2] * vec[0] - k[0] * vec[2], k[0] * vec[1] - k[1] * vec[0]};
    const real scal = (k[0] * vec[0] + k[1] * vec[1] + k[2] * vec[2]) / (1.0 + c);
    for (int a2 = 0; a2 < 3; ++a2) r[a2] += vec[a2] * c + k[a2] * scal;
    vec[0] = r[0]; vec[1] = r[1]; vec[2] = r[2];
}

int orc_ph_tilt(const double *g, const double *vox, const double *vec_in, int64_t n, double *vec_out) {
    pbeam_t b;
    load_beam(g, &b);
    for (int64_t t = 0; t < n; ++t) {
        const real v[3] = {(real)vox[3 * t], (real)vox[3 * t + 1], (real)vox[3 * t + 2]};
        real w[3] = {(real)vec_in[3 * t], (real)vec_in[3 * t + 1], (real)vec_in[3 * t + 2]};
        tilt(&b, v, w);
        for (int a = 0; a < 3; ++a) vec_out[3 * t + a] = (double)w[a];
    }
    return 0;
}

int orc_ph_frames(const double *g, const double *pts, int64_t n, int to_head, double *out) {
    pbeam_t b;
    load_beam(g, &b);
    for (int64_t t = 0; t < n; ++t) {
        const real p[3] = {(real)pts[3 * t], (real)pts[3 * t + 1], (real)pts[3 * t + 2]};
        real q[3];
        if (to_head)
            img_to_head(&b, p, q);
        else
            head_to_img(&b, p, q);
        for (int a = 0; a < 3; ++a) out[3 * t + a] = (double)q[a];
    }
    return 0;
}

/* cccKernel (IMRTClasses.cu:252-350) at one voxel.  kernel rows (one per
 * cone): theta (deg), Am, am, Bm, bm, ray_length (cm), as float32. */
static real ccc_voxel(const float *terma, const float *rho, int nx, int ny, int nz, float s32, const pbeam_t *b,
                      const float *kernel, int i, int j, int k, real electron, real mu_cal, real mu) {
    const real s = (real)s32;
    const int64_t v_idx = i + (int64_t)nx * (j + (int64_t)ny * k);
    if (!((real)terma[v_idx] > 0.01)) return 0.0;
    real v[3];
    voxel_xyz(b, s, i, j, k, v);
    const float sp32 = s32 / 10.0f;
    const real sp = (real)sp32;
    real dose = 0.0;
    for (int c = 0; c < 6; ++c) {
        const real th = (real)kernel[6 * c] * (PI_D / 180.0);
        const real Am = kernel[6 * c + 1], am = kernel[6 * c + 2];
        const real Bm = kernel[6 * c + 3], bm = kernel[6 * c + 4];
        const real sth = SIN(th), cth = COS(th);
        for (int az = 0; az < 12; ++az) {
            const real phi = (real)az * (PI_D / 6.0);
            const real th_h[3] = {sth * COS(phi), sth * SIN(phi), cth};
            real t[3];
            head_to_img(b, th_h, t);
            tilt(b, v, t);
            real rp = 0.0, rs = 0.0;
            /* the reference's float32 ray-length counter */
            for (float rl = kernel[6 * c + 5]; rl >= 0.0f; rl = rl - sp32) {
                const real len = (real)rl * 10.0;
                const real cx = (v[0] + t[0] * len + b->iso[0]) / s;
                const real cy = (v[1] + t[1] * len + b->iso[1]) / s;
                const real cz = (v[2] + t[2] * len + b->iso[2]) / s;
                const real ti = sample(terma, nx, ny, nz, 0.0, cx, cy, cz);
                real di = sample(rho, nx, ny, nz, AIR_DENSITY, cx, cy, cz);
                di = (di > AIR_DENSITY ? di : AIR_DENSITY) * sp;
                const real e = EXP(-am * di);
                rp = rp * e + ti * sth * (Am / (am * am)) * (1.0 - e);
                rs = rs * (1.0 - bm * di) + ti * di * sth * (Bm / bm);
            }
            dose += am * rp + bm * rs;
        }
    }
    dose += electron;
    if (dose != dose || dose < 0.0) return 0.0;
    return mu_cal * dose * mu;
}

/* terma: float32 volume (the TERMA texture's data), rho: float32 density,
 * electron: per listed voxel (or per voxel without idx). */
int orc_ph_ccc(const float *terma, const float *rho, const double *electron, int nx, int ny, int nz, float s,
               const double *g, const float *kernel, double mu_cal, double mu, const int64_t *idx, int64_t n_idx,
               double *out) {
    pbeam_t b;
    load_beam(g, &b);
    const int64_t n = idx ? n_idx : (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < n; ++t) {
        int i, j, k;
        voxel_ijk(idx ? idx[t] : t, nx, ny, &i, &j, &k);
        out[t] = (double)ccc_voxel(terma, rho, nx, ny, nz, s, &b, kernel, i, j, k, (real)electron[t], (real)mu_cal,
                                   (real)mu);
    }
    return 0;
}
