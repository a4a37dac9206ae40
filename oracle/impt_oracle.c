/*
 * impt_oracle.c - CPU restatement (float64) of the reference's proton
 * pencil-beam dose path.  TEST INFRASTRUCTURE ONLY: it is the reference the
 * HIP kernels are compared against in the tests and the CPU baseline of the
 * benchmark; nothing in the product links or calls it.
 *
 * Primary source: DoseCUDA/Jax/impt_jax.py (the JAX path), evaluated in
 * float64 and in the axis convention of the CUDA extension (numpy axis 2 = x,
 * CudaClasses.cu:168-178), which impt_jax.py itself does not use
 * (impt_jax.py:889-897 treats axis 0 as x).
 *
 *   orc_raymarch  _raytrace_kernel_optimized  impt_jax.py:235-284 (+ :548-575)
 *   orc_smooth    _smooth_wet_kernel          impt_jax.py:317-418
 *   orc_wet_at    raymarch + smooth at chosen voxels (neighbours marched on demand)
 *   orc_pb_dose   _pencil_beam_single_layer   impt_jax.py:425-541, summed over
 *                 layers as compute_dose_pure impt_jax.py:602-682
 *   orc_pb_grad   analytic dL/dmu, dL/dx, dL/dy of orc_pb_dose (no reference
 *                 counterpart: the reference differentiates with jax.grad)
 *   orc_halo      the nuclear-halo fit alone (test hook)
 *
 * Decisions where the JAX code leaves float-level behaviour open:
 *  - trilinear sampling with zero padding per corner (scipy 'grid-constant'),
 *    samples whose continuous index leaves [0, n) on any axis dropped;
 *  - neighbour rounding half to even (jnp.round) on the real-number
 *    coordinate (half-voxel ties detected to 1e-6), exact 60-degree table;
 *  - a Gaussian whose sigma is not positive contributes 0.
 *
 * Compile with -DORC_FLOAT for a float32 build (benchmark baseline only).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef ORC_FLOAT
typedef float real;
#define R(x) ((float)(x))
#define SQRT sqrtf
#define EXP expf
#define LOG logf
#define FLOOR floorf
#define RINT rintf
#else
typedef double real;
#define R(x) ((double)(x))
#define SQRT sqrt
#define EXP exp
#define LOG log
#define FLOOR floor
#define RINT rint
#endif

typedef struct {
    real iso[3], src[3];
    real sg, cg, sc, cc;
    real vsadx, vsady;
} beam_t;

static void load_beam(const double *g, beam_t *b) {
    for (int a = 0; a < 3; ++a) {
        b->iso[a] = R(g[a]);
        b->src[a] = R(g[3 + a]);
    }
    b->sg = R(g[6]); b->cg = R(g[7]); b->sc = R(g[8]); b->cc = R(g[9]);
    b->vsadx = R(g[10]); b->vsady = R(g[11]);
}

/* voxel centre relative to the isocentre (impt_jax.py:186-188, CUDA axes) */
static void voxel_xyz(const beam_t *b, real s, int i, int j, int k, real p[3]) {
    p[0] = (real)i * s - b->iso[0];
    p[1] = (real)j * s - b->iso[1];
    p[2] = (real)k * s - b->iso[2];
}

/* image -> beam's-eye-view (impt_jax.py:190-201) */
static void image_to_head(const beam_t *b, const real p[3], real h[3]) {
    real xt = p[0] * b->cc + p[2] * (-b->sc);
    real yt = p[1];
    real zt = -p[0] * (-b->sc) + p[2] * b->cc;
    real xg = xt * b->cg - yt * (-b->sg);
    real yg = xt * (-b->sg) + yt * b->cg;
    h[0] = -xg;
    h[1] = zt;
    h[2] = yg;
}

/* beam's-eye-view -> image (impt_jax.py:291-314) */
static void head_to_image(const beam_t *b, const real h[3], real p[3]) {
    real xz = -h[0], yz = h[2], zz = h[1];
    real xg = xz * b->cg - yz * b->sg;
    real yg = xz * b->sg + yz * b->cg;
    real zg = zz;
    p[0] = xg * b->cc + zg * b->sc;
    p[1] = yg;
    p[2] = -xg * b->sc + zg * b->cc;
}

static inline int64_t vidx(int nx, int ny, int i, int j, int k) {
    return (int64_t)i + (int64_t)nx * ((int64_t)j + (int64_t)ny * k);
}

/* order-1 interpolation with zero padding per corner */
static real trilinear(const float *rho, int nx, int ny, int nz, real cx, real cy, real cz) {
    real fx = FLOOR(cx), fy = FLOOR(cy), fz = FLOOR(cz);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    real wx[2] = {R(1) - (cx - fx), cx - fx};
    real wy[2] = {R(1) - (cy - fy), cy - fy};
    real wz[2] = {R(1) - (cz - fz), cz - fz};
    real acc = 0;
    for (int c = 0; c < 2; ++c) {
        int z = iz + c;
        if (z < 0 || z >= nz) continue;
        for (int bb = 0; bb < 2; ++bb) {
            int y = iy + bb;
            if (y < 0 || y >= ny) continue;
            for (int a = 0; a < 2; ++a) {
                int x = ix + a;
                if (x < 0 || x >= nx) continue;
                acc += wx[a] * wy[bb] * wz[c] * (real)rho[vidx(nx, ny, x, y, z)];
            }
        }
    }
    return acc;
}

static real raymarch_one(const float *rho, int nx, int ny, int nz, real s, const beam_t *b, int i, int j, int k) {
    real v[3];
    voxel_xyz(b, s, i, j, k, v);
    real d[3] = {b->src[0] - v[0], b->src[1] - v[1], b->src[2] - v[2]};
    real len = SQRT(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    real u[3] = {d[0] / len, d[1] / len, d[2] / len};
    /* impt_jax.py:569-575 */
    double md = sqrt((double)nx * s * nx * s + (double)ny * s * ny * s + (double)nz * s * nz * s) + 500.0;
    long max_steps = (long)md + 10;
    real wet = R(-0.05);
    int entered = 0;
    for (long n = 0; n < max_steps; ++n) {
        real rl = (real)n;
        real c[3];
        for (int a = 0; a < 3; ++a) c[a] = ((v[a] + u[a] * rl) + b->iso[a]) / s;
        int inside = c[0] >= 0 && c[0] < nx && c[1] >= 0 && c[1] < ny && c[2] >= 0 && c[2] < nz;
        if (inside) {
            real rv = trilinear(rho, nx, ny, nz, c[0], c[1], c[2]);
            wet += (rv > 0 ? rv : 0) * R(0.1);
            entered = 1;
        } else if (entered) {
            break; /* a convex grid is left once; later samples are all outside */
        }
    }
    return wet;
}

int orc_raymarch(const float *rho, int nx, int ny, int nz, double spacing, const double *geom, const int64_t *idx,
                 int64_t n, double *out) {
    beam_t b;
    load_beam(geom, &b);
    const real s = R(spacing);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t t = 0; t < n; ++t) {
        int64_t v = idx ? idx[t] : t;
        int i = (int)(v % nx);
        int64_t r = v / nx;
        int j = (int)(r % ny), k = (int)(r / ny);
        out[t] = (double)raymarch_one(rho, nx, ny, nz, s, &b, i, j, k);
    }
    return 0;
}

static const double k_cos6[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
static const double k_sin6[6] = {0.0, 0.86602540378443864676, 0.86602540378443864676,
                                 0.0, -0.86602540378443864676, -0.86602540378443864676};

/* Nearest integer, ties to even, on the real-number coordinate: a value within
 * 1e-6 of a half-integer is an exact tie of the exact arithmetic (offsets built
 * from 60-degree directions and 30-degree-multiple gantry angles are rational),
 * which float evaluation would otherwise break by rounding noise. */
static long round_half_even_exact(real c) {
    real fl = FLOOR(c);
    real fr = c - fl;
    long f = (long)fl;
    if (fabs((double)fr - 0.5) < 1e-6) return (f & 1) ? f + 1 : f;
    return (long)RINT(c);
}

/* neighbour list of one voxel for smoothing; returns count (<= 54) */
static int smooth_neighbours(const beam_t *b, int nx, int ny, int nz, real s, int i, int j, int k, real wc,
                             int64_t *nb) {
    real v[3], h[3];
    voxel_xyz(b, s, i, j, k, v);
    image_to_head(b, v, h);
    real max_dr = wc * R(10);
    if (max_dr > R(10)) max_dr = R(10);
    int cnt = 0;
    for (int a = 0; a < 6; ++a) {
        for (int dd = 1; dd <= 10; ++dd) {
            real dr = (real)dd;
            if (!(dr < max_dr)) continue;
            real hp[3] = {h[0] + dr * R(k_cos6[a]), h[1] + dr * R(k_sin6[a]), h[2]};
            real ip[3];
            head_to_image(b, hp, ip);
            long ni = round_half_even_exact((ip[0] + b->iso[0]) / s);
            long nj = round_half_even_exact((ip[1] + b->iso[1]) / s);
            long nk = round_half_even_exact((ip[2] + b->iso[2]) / s);
            if (ni >= 0 && ni < nx && nj >= 0 && nj < ny && nk >= 0 && nk < nz)
                nb[cnt++] = vidx(nx, ny, (int)ni, (int)nj, (int)nk);
        }
    }
    return cnt;
}

int orc_smooth(const double *raw, int nx, int ny, int nz, double spacing, const double *geom, const int64_t *idx,
               int64_t n, double *out) {
    beam_t b;
    load_beam(geom, &b);
    const real s = R(spacing);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t t = 0; t < n; ++t) {
        int64_t v = idx ? idx[t] : t;
        int i = (int)(v % nx);
        int64_t r = v / nx;
        int j = (int)(r % ny), k = (int)(r / ny);
        int64_t nb[64];
        real wc = (real)raw[v];
        int cnt = smooth_neighbours(&b, nx, ny, nz, s, i, j, k, wc, nb);
        real sum = wc;
        for (int q = 0; q < cnt; ++q) sum += (real)raw[nb[q]];
        out[t] = (double)(sum / (real)(cnt + 1));
    }
    return 0;
}

int orc_wet_at(const float *rho, int nx, int ny, int nz, double spacing, const double *geom, const int64_t *idx,
               int64_t n, double *out) {
    beam_t b;
    load_beam(geom, &b);
    const real s = R(spacing);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < n; ++t) {
        int64_t v = idx[t];
        int i = (int)(v % nx);
        int64_t r = v / nx;
        int j = (int)(r % ny), k = (int)(r / ny);
        real wc = raymarch_one(rho, nx, ny, nz, s, &b, i, j, k);
        int64_t nb[64];
        int cnt = smooth_neighbours(&b, nx, ny, nz, s, i, j, k, wc, nb);
        real sum = wc;
        for (int q = 0; q < cnt; ++q) {
            int64_t w = nb[q];
            int wi = (int)(w % nx);
            int64_t wr = w / nx;
            sum += raymarch_one(rho, nx, ny, nz, s, &b, wi, (int)(wr % ny), (int)(wr / ny));
        }
        out[t] = (double)(sum / (real)(cnt + 1));
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
typedef struct {
    const double *spots; /* [S][3] x, y, mu (sorted by energy) */
    const int32_t *layers; /* [L][3] spot_start, n_spots, energy_id */
    const double *r80; /* [L] */
    int n_layers;
    const float *dvp;
    int dvp_len;
    const float *depths, *sigmas, *idds;
    int lut_len;
} pb_t;

/* searchsorted(side='left') lookup with end clamps (impt_jax.py:452-462) */
static void lut_interp(const float *dep, const float *sig, const float *idd, int L, real wet, real *o_idd,
                       real *o_sig) {
    int lo = 0, hi = L;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if ((real)dep[mid] < wet) lo = mid + 1; else hi = mid;
    }
    int i = lo;
    if (i >= L) { *o_idd = idd[L - 1]; *o_sig = sig[L - 1]; return; }
    if (i <= 0) { *o_idd = idd[0]; *o_sig = sig[0]; return; }
    real den = (real)dep[i] - (real)dep[i - 1];
    real f = den > 0 ? (wet - (real)dep[i - 1]) / den : 0;
    *o_idd = (real)idd[i - 1] + f * ((real)idd[i] - (real)idd[i - 1]);
    *o_sig = (real)sig[i - 1] + f * ((real)sig[i] - (real)sig[i - 1]);
}

typedef struct {
    real Pf, Hf, ps, hs; /* amplitudes; -1/(2 sigma^2) or 0 if the Gaussian is absent */
    int p_on, h_on;
} vterm_t;

/* nuclear-halo width sigma_n (mm) and weight u_n at WET wet_mm of an energy
 * with range r80 (impt_jax.py:469-483; nuclearHalo IMPTClasses.cu:104-121):
 * the depth clamped to [0.1, r80 - 0.1] (fmin(fmax()) order, so r80 < 0.2
 * gives r80 - 0.1), sigma_n >= 0, 0 <= u_n <= 0.9 */
static void halo_fit(real r80, real wet_mm, real *o_sn, real *o_un) {
    real w = wet_mm;
    if (w < R(0.1)) w = R(0.1);
    if (w > r80 - R(0.1)) w = r80 - R(0.1);
    real sn = R(2.85) + R(0.0014) * r80 * LOG(w + R(3)) + R(0.06) * w - R(7.4e-5) * w * w -
              (R(0.22) * r80) / ((w - r80 - R(5)) * (w - r80 - R(5)));
    real un = R(0.052) * LOG(R(1.13) + w / (R(11.2) - R(0.023) * r80)) +
              (R(0.35) * (R(0.0017) * r80 * r80 - r80)) / ((r80 + R(3)) * (r80 + R(3)) - w * w) -
              R(1.61e-9) * w * (r80 + R(3)) * (r80 + R(3));
    if (sn < 0) sn = 0;
    if (un < 0) un = 0;
    if (un > R(0.9)) un = R(0.9);
    *o_sn = sn;
    *o_un = un;
}

/* test hook: halo_fit at n depths */
int orc_halo(double r80, const double *wet_mm, int64_t n, double *sn, double *un) {
    for (int64_t t = 0; t < n; ++t) {
        real a, b;
        halo_fit(R(r80), R(wet_mm[t]), &a, &b);
        sn[t] = (double)a;
        un[t] = (double)b;
    }
    return 0;
}

/* per (voxel, layer) terms (impt_jax.py:445-496) */
static void voxel_terms(const pb_t *P, int l, real wet_mm, real dsrc, vterm_t *T) {
    const int eid = P->layers[3 * l + 2];
    const real r80 = (real)P->r80[l];
    real idd, sms;
    lut_interp(P->depths + (int64_t)eid * P->lut_len, P->sigmas + (int64_t)eid * P->lut_len,
               P->idds + (int64_t)eid * P->lut_len, P->lut_len, wet_mm, &idd, &sms);
    const float *c = P->dvp + (int64_t)eid * P->dvp_len;
    real d = dsrc - wet_mm + R(0.7) * r80;
    real st = (real)c[2] * d * d + (real)c[3] * d + (real)c[4] + sms;
    real sn, un;
    halo_fit(r80, wet_mm, &sn, &un);
    real st2 = st * st;
    real sh = SQRT(st2 + sn * sn);
    real sh2 = sh * sh;
    const real two_pi = R(6.28318530717958647692);
    T->Pf = (R(1) - un) * idd / (two_pi * st2 + R(1e-8));
    T->Hf = un * idd / (two_pi * sh2 + R(1e-8));
    T->p_on = st > 0;
    T->h_on = sh > 0;
    T->ps = T->p_on ? R(-0.5) / st2 : 0;
    T->hs = T->h_on ? R(-0.5) / sh2 : 0;
}

/* squared distance to a spot axis and its pieces (impt_jax.py:510-523) */
static real cax_d2(const beam_t *b, const real h[3], real xs, real ys, real *q, real *rx, real *ry, real *tx,
                   real *ty) {
    *tx = xs / b->vsadx;
    *ty = ys / b->vsady;
    const real tz = R(-1);
    *rx = h[0] - xs;
    *ry = h[1] - ys;
    const real rz = h[2];
    real tt = *tx * *tx + *ty * *ty + tz * tz;
    real rr = *rx * *rx + *ry * *ry + rz * rz;
    real tr = *tx * *rx + *ty * *ry + tz * rz;
    *q = tr / tt;
    return rr - tr * tr / tt;
}

static void pb_setup(pb_t *P, const double *spots, const int32_t *layers, const double *r80, int n_layers,
                     const float *dvp, int dvp_len, const float *depths, const float *sigmas, const float *idds,
                     int lut_len) {
    P->spots = spots; P->layers = layers; P->r80 = r80; P->n_layers = n_layers;
    P->dvp = dvp; P->dvp_len = dvp_len;
    P->depths = depths; P->sigmas = sigmas; P->idds = idds; P->lut_len = lut_len;
}

int orc_pb_dose(const double *wet, int nx, int ny, int nz, double spacing, const double *geom, const double *spots,
                const int32_t *layers, const double *r80, int n_layers, const float *dvp, int dvp_len,
                const float *depths, const float *sigmas, const float *idds, int lut_len, const int64_t *idx,
                int64_t n, double *out) {
    beam_t b;
    load_beam(geom, &b);
    pb_t P;
    pb_setup(&P, spots, layers, r80, n_layers, dvp, dvp_len, depths, sigmas, idds, lut_len);
    const real s = R(spacing);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t t = 0; t < n; ++t) {
        int64_t v = idx ? idx[t] : t;
        int i = (int)(v % nx);
        int64_t r = v / nx;
        int j = (int)(r % ny), k = (int)(r / ny);
        real p[3], h[3];
        voxel_xyz(&b, s, i, j, k, p);
        image_to_head(&b, p, h);
        real dx = b.src[0] - p[0], dy = b.src[1] - p[1], dz = b.src[2] - p[2];
        real dsrc = SQRT(dx * dx + dy * dy + dz * dz);
        real wet_mm = (real)wet[v] * R(10);
        real dose = 0;
        for (int l = 0; l < n_layers; ++l) {
            if (!(wet_mm <= R(1.1) * (real)r80[l])) continue; /* mask, impt_jax.py:448-449, 537 */
            vterm_t T;
            voxel_terms(&P, l, wet_mm, dsrc, &T);
            const int s0 = layers[3 * l], ns = layers[3 * l + 1];
            real acc = 0;
            for (int q = s0; q < s0 + ns; ++q) {
                real qq, rx, ry, tx, ty;
                real d2 = cax_d2(&b, h, (real)spots[3 * q], (real)spots[3 * q + 1], &qq, &rx, &ry, &tx, &ty);
                real pd = T.p_on ? T.Pf * EXP(T.ps * d2) : 0;
                real hd = T.h_on ? T.Hf * EXP(T.hs * d2) : 0;
                acc += (real)spots[3 * q + 2] * (pd + hd);
            }
            dose += acc;
        }
        out[t] = (double)dose;
    }
    return 0;
}

/* dL/dmu_s = sum_v g_v K_vs; dL/dx_s = sum_v g_v mu_s dK/dd2 * dd2/dx_s with
 * dd2/dx_s = -2 (1 + q/VSADX)(r_x - q t_x), q = (t.r)/|t|^2 (and y alike). */
int orc_pb_grad(const double *wet, const double *gvol, int nx, int ny, int nz, double spacing, const double *geom,
                const double *spots, int n_spots, const int32_t *layers, const double *r80, int n_layers,
                const float *dvp, int dvp_len, const float *depths, const float *sigmas, const float *idds,
                int lut_len, double *grad_mu, double *grad_xy) {
    beam_t b;
    load_beam(geom, &b);
    pb_t P;
    pb_setup(&P, spots, layers, r80, n_layers, dvp, dvp_len, depths, sigmas, idds, lut_len);
    const real s = R(spacing);
    const int64_t nvox = (int64_t)nx * ny * nz;
    int nthr = 1;
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    nthr = omp_get_max_threads();
#endif
    double *acc = (double *)calloc((size_t)nthr * n_spots * 3, sizeof(double));
    if (!acc) return -1;
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        extern int omp_get_thread_num(void);
        tid = omp_get_thread_num();
#endif
        double *my = acc + (size_t)tid * n_spots * 3;
#pragma omp for schedule(static)
        for (int64_t v = 0; v < nvox; ++v) {
            const real g = (real)gvol[v];
            if (g == 0) continue;
            int i = (int)(v % nx);
            int64_t r = v / nx;
            int j = (int)(r % ny), k = (int)(r / ny);
            real p[3], h[3];
            voxel_xyz(&b, s, i, j, k, p);
            image_to_head(&b, p, h);
            real dx = b.src[0] - p[0], dy = b.src[1] - p[1], dz = b.src[2] - p[2];
            real dsrc = SQRT(dx * dx + dy * dy + dz * dz);
            real wet_mm = (real)wet[v] * R(10);
            for (int l = 0; l < n_layers; ++l) {
                if (!(wet_mm <= R(1.1) * (real)r80[l])) continue;
                vterm_t T;
                voxel_terms(&P, l, wet_mm, dsrc, &T);
                const int s0 = layers[3 * l], ns = layers[3 * l + 1];
                for (int q = s0; q < s0 + ns; ++q) {
                    real qq, rx, ry, tx, ty;
                    real mu = (real)spots[3 * q + 2];
                    real d2 = cax_d2(&b, h, (real)spots[3 * q], (real)spots[3 * q + 1], &qq, &rx, &ry, &tx, &ty);
                    real pd = T.p_on ? T.Pf * EXP(T.ps * d2) : 0;
                    real hd = T.h_on ? T.Hf * EXP(T.hs * d2) : 0;
                    real dK = pd * T.ps + hd * T.hs; /* dK/d(d2) */
                    real ddx = R(-2) * (R(1) + qq / b.vsadx) * (rx - qq * tx);
                    real ddy = R(-2) * (R(1) + qq / b.vsady) * (ry - qq * ty);
                    my[3 * q] += (double)(g * (pd + hd));
                    my[3 * q + 1] += (double)(g * mu * dK * ddx);
                    my[3 * q + 2] += (double)(g * mu * dK * ddy);
                }
            }
        }
    }
    for (int q = 0; q < n_spots; ++q) {
        double a = 0, bx = 0, by = 0;
        for (int t = 0; t < nthr; ++t) {
            a += acc[((size_t)t * n_spots + q) * 3];
            bx += acc[((size_t)t * n_spots + q) * 3 + 1];
            by += acc[((size_t)t * n_spots + q) * 3 + 2];
        }
        grad_mu[q] = a;
        grad_xy[2 * q] = bx;
        grad_xy[2 * q + 1] = by;
    }
    free(acc);
    return 0;
}
