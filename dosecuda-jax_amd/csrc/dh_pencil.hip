// Double-Gaussian pencil-beam spot x voxel kernels (gfx950, wave64).
//
//  k_spot_prep      spot records + per-spot constants of the beam-axis geometry
//  k_layer_prep     per-layer record: constants, reach of its spots
//  k_pb_forward     dose: lanes = voxels of a 4x4x4 brick, spots wave-uniform
//  k_brick_summary  per-brick WET range of the voxels with dL/dD != 0
//  k_pb_adjoint     dL/dmu, dL/dx, dL/dy partials: lanes = spots, voxel terms
//                   broadcast from LDS
//  k_pb_reduce      fixed-order sum of the adjoint partials
//
// Per (voxel, layer) terms follow pencilBeamKernel (IMPTClasses.cu:204-271) and
// _pencil_beam_single_layer (impt_jax.py:425-541):
//   wet  = 10 * WET(cm), skip when wet > 1.1 R80
//   (idd, sigma_ms) = LUT lerp at wet, clamped at both ends
//   d    = |src - v| - wet + 0.7 R80,  sigma_air = A d^2 + B d + C
//   st   = sigma_air + sigma_ms
//   w    = clamp(wet, 0.1, R80 - 0.1) -> halo sigma sn, halo weight un
//   sh   = hypot(st, sn)
//   D   += sum_s mu_s [ (1-un) idd/(2 pi st^2 + 1e-8) exp(-d2_s/(2 st^2))
//                     +    un  idd/(2 pi sh^2 + 1e-8) exp(-d2_s/(2 sh^2)) ]
// with d2_s the squared distance from the voxel (beam's-eye view) to spot s's
// axis, through (x_s, y_s, 0) along t = (x_s/VSADX, y_s/VSADY, -1)
// (caxDistance, IMPTClasses.cu:123-133).  A Gaussian whose sigma is not
// positive contributes nothing (the -inf exponent scale of impt_jax.py:491-492).
//
// Skipping work (cutoff_sigma > 0): a spot whose axis stays further than
// cutoff_sigma halo sigmas from every voxel of a brick contributes less than
// exp(-cutoff^2/2) of its peak there and is not evaluated.  Tests run
// lane-parallel: the 64 lanes of a wave judge 64 candidates (layers, spots or
// bricks) at once and the wave then walks only the survivors.
#include "dh_common.h"

namespace dh {

// Spot record with the constants of its beam-axis geometry.  With
// t = (tx, ty, -1) the spot axis direction, a head-frame point h sits at
// in-plane offset e = (hx - x + hz tx, hy - y + hz ty) from where the axis
// crosses the plane z = hz, and its squared distance to the axis is the
// quadratic form  d2 = m11 e_x^2 + 2 m12 e_x e_y + m22 e_y^2  of
// M = I - t_xy t_xy^T / |t|^2.  Unlike |r|^2 - (t.r)^2/|t|^2 this has no
// cancellation far from the isocentre plane (|r| ~ 150 mm vs d ~ 1 mm).
struct SpotExt {
    float x, y, tx, ty, mu, m11, m12x2, m22;
};

struct Model {
    const float2 *idd_sigma;
    const float *depths;
    const float *dvp;
    int n_energies, lut_len, dvp_len, uniform;
    float depth0, inv_ddepth;
};

// Layer record (64 bytes): the layer's spot range and energy, the constants of
// its sigma_air polynomial, and the reach of its spots: bounding box of the
// positions at the isocentre plane, the smallest eigenvalue 1/(1 + |t_xy|^2)
// of their distance forms and max |sigma_ms| of the energy's lookup row.
struct LayerC {
    int start, n, eid, ok;
    float r80, A, B, C;
    float xmin, xmax, ymin, ymax;
    float lam_min, sms_max, lim, pad;
};

constexpr float kLog2e = 1.4426950408889634f;
constexpr float kTwoPi = 6.283185307179586f;

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// natural log through v_log_f32 (log2)
__device__ __forceinline__ float fast_log(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994531f; }

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off));
    return v;
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_spot_prep(const dh_spot *__restrict__ in, SpotExt *__restrict__ out, int n,
                                                  float inv_vsadx, float inv_vsady) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    dh_spot p = in[s];
    SpotExt e;
    e.x = p.x;
    e.y = p.y;
    e.mu = p.mu;
    e.tx = p.x * inv_vsadx;
    e.ty = p.y * inv_vsady;
    const float itt = 1.0f / (e.tx * e.tx + e.ty * e.ty + 1.0f);
    e.m11 = 1.0f - e.tx * e.tx * itt;
    e.m22 = 1.0f - e.ty * e.ty * itt;
    e.m12x2 = -2.0f * e.tx * e.ty * itt;
    out[s] = e;
}

// Reach of spots [start, start + n) (lane-parallel) into c.
__device__ __forceinline__ void reach_of(const SpotExt *__restrict__ sp, int start, int n, const Model &m, int eid,
                                         int lane, LayerC &c) {
    float xmin = 3.0e38f, xmax = -3.0e38f, ymin = 3.0e38f, ymax = -3.0e38f, t2max = 0.0f, smax = 0.0f;
    for (int s = lane; s < n; s += WAVE) {
        const SpotExt q = sp[start + s];
        xmin = fminf(xmin, q.x);
        xmax = fmaxf(xmax, q.x);
        ymin = fminf(ymin, q.y);
        ymax = fmaxf(ymax, q.y);
        t2max = fmaxf(t2max, q.tx * q.tx + q.ty * q.ty);
    }
    const float2 *row = m.idd_sigma + (long long)eid * m.lut_len;
    for (int t = lane; t < m.lut_len; t += WAVE) smax = fmaxf(smax, fabsf(row[t].y));
    c.xmin = wave_min(xmin);
    c.xmax = wave_max(xmax);
    c.ymin = wave_min(ymin);
    c.ymax = wave_max(ymax);
    c.lam_min = 1.0f / (1.0f + wave_max(t2max));
    c.sms_max = wave_max(smax);
}

// One wave per layer.  A record the kernels cannot use safely (energy id
// outside the tables, spot range outside the spot array) gets ok = 0 and is
// skipped rather than read out of bounds.
__global__ __launch_bounds__(64) void k_layer_prep(const SpotExt *__restrict__ sp, int n_spots,
                                                  const dh_layer *__restrict__ layers, Model m,
                                                  LayerC *__restrict__ out) {
    const int l = blockIdx.x, lane = threadIdx.x;
    const dh_layer L = layers[l];
    LayerC c;
    c.ok = (unsigned)L.energy_id < (unsigned)m.n_energies && L.spot_start >= 0 && L.n_spots > 0 &&
           L.spot_start <= n_spots - L.n_spots;
    c.start = c.ok ? L.spot_start : 0;
    c.n = c.ok ? L.n_spots : 0;
    c.eid = c.ok ? L.energy_id : 0;
    c.r80 = L.r80;
    c.lim = 1.1f * L.r80;
    const float *d = m.dvp + (long long)c.eid * m.dvp_len;
    c.A = d[2];
    c.B = d[3];
    c.C = d[4];
    c.pad = 0.0f;
    if (c.ok) {
        reach_of(sp, c.start, c.n, m, c.eid, lane, c);
    } else {
        c.xmin = c.ymin = -3.0e38f;
        c.xmax = c.ymax = 3.0e38f;
        c.lam_min = 0.0f;
        c.sms_max = 0.0f;
    }
    if (lane == 0) out[l] = c;
}

// LUT lookup (interpolateProtonLUT, IMPTClasses.cu:69-92): i = first depth >= wet;
// i == 0 -> row 0, i == len -> last row, else lerp between i-1 and i.
__device__ __forceinline__ void lut_lookup(const Model &m, int eid, float wet, float &idd, float &sig) {
    const float2 *row = m.idd_sigma + (long long)eid * m.lut_len;
    const int L = m.lut_len;
    int lo;
    float f;
    if (m.uniform) {
        const float q = (wet - m.depth0) * m.inv_ddepth;
        if (!(q > 0.0f)) {
            const float2 r = row[0];
            idd = r.x;
            sig = r.y;
            return;
        }
        if (q > (float)(L - 1)) {
            const float2 r = row[L - 1];
            idd = r.x;
            sig = r.y;
            return;
        }
        lo = min((int)q, L - 2);
        f = q - (float)lo;
    } else {
        const float *dep = m.depths + (long long)eid * L;
        int l = 0, r = L;
        while (l < r) {
            const int mid = (l + r) >> 1;
            if (wet <= dep[mid]) r = mid; else l = mid + 1;
        }
        if (l == 0 || l == L) {
            const float2 v = row[l == 0 ? 0 : L - 1];
            idd = v.x;
            sig = v.y;
            return;
        }
        lo = l - 1;
        f = (wet - dep[lo]) / (dep[l] - dep[lo]);
    }
    const float2 a = row[lo], c = row[lo + 1];
    idd = fmaf(c.x - a.x, f, a.x);
    sig = fmaf(c.y - a.y, f, a.y);
}

// Brick summary for the reach test: centre in the beam's-eye view, distance
// of the centre to the source, WET range (mm) of the voxels that matter.
struct BrickInfo {
    float3 hc;
    float dsrc, wlo, whi;
};

__device__ __forceinline__ float3 brick_centre(const Grid &g, const Geom &b, int bi, int bj, int bk) {
    const float half = 0.5f * (BRICK - 1);
    return make_float3(fmaf((float)bi * BRICK + half, g.s, -b.iso[0]), fmaf((float)bj * BRICK + half, g.s, -b.iso[1]),
                       fmaf((float)bk * BRICK + half, g.s, -b.iso[2]));
}

__device__ __forceinline__ BrickInfo brick_info_at(const Geom &b, float3 vc, float wlo, float whi) {
    BrickInfo B;
    B.hc = image_to_head(b, vc);
    const float sx = b.src[0] - vc.x, sy = b.src[1] - vc.y, sz = b.src[2] - vc.z;
    B.dsrc = fast_sqrt(sx * sx + sy * sy + sz * sz);
    B.wlo = wlo;
    B.whi = whi;
    return B;
}

// True when every spot of the reach record c is negligible for the whole
// brick: the brick centre's distance to the nearest spot axis, less the brick
// radius, exceeds cut * (an upper bound of the halo sigma over the brick).  The
// bound takes sigma_air over the range of d = |src - v| - wet + 0.7 R80 (a
// quadratic: ends and vertex) plus max |sigma_ms|, and the halo width without
// its negative terms, over the brick's WET range.
__device__ __forceinline__ bool reach_skip(const LayerC &c, const BrickInfo &B, float ivx, float ivy, float cut,
                                           float rbrick) {
    if (!(c.r80 >= 5.0f)) return false;
    const float fx = fmaf(-B.hc.z, ivx, 1.0f), fy = fmaf(-B.hc.z, ivy, 1.0f);
    const float ex = fmaxf(0.0f, fmaxf(c.xmin * fx - B.hc.x, B.hc.x - c.xmax * fx));
    const float ey = fmaxf(0.0f, fmaxf(c.ymin * fy - B.hc.y, B.hc.y - c.ymax * fy));
    const float dlb = 0.9999f * fast_sqrt(c.lam_min * (ex * ex + ey * ey)) - rbrick;
    if (!(dlb > 0.0f)) return false;
    const float whi = fminf(B.whi, c.lim);
    const float w_ub = fminf(fmaxf(whi, 0.1f), c.r80 - 0.1f);
    const float sn_ub = 2.85f + 0.0014f * c.r80 * fast_log(w_ub + 3.0f) + 0.06f * w_ub;
    const float d07 = 0.7f * c.r80;
    const float dlo = B.dsrc - rbrick - whi + d07;
    const float dhi = B.dsrc + rbrick - B.wlo + d07;
    float sa = fmaxf(fabsf(fmaf(fmaf(c.A, dlo, c.B), dlo, c.C)), fabsf(fmaf(fmaf(c.A, dhi, c.B), dhi, c.C)));
    if (c.A != 0.0f) {
        const float vx = -c.B / (2.0f * c.A);
        if (vx > dlo && vx < dhi) sa = fmaxf(sa, fabsf(fmaf(fmaf(c.A, vx, c.B), vx, c.C)));
    }
    const float st_ub = sa + c.sms_max;
    const float sh_ub = 1.001f * fast_sqrt(st_ub * st_ub + sn_ub * sn_ub);
    return dlb > cut * sh_ub;
}

// Per (voxel, layer) Gaussian parameters.
struct VoxTerms {
    float Af, Bf;    // primary / halo amplitude (idd * weight / (2 pi sigma^2 + 1e-8))
    float a1, a2;    // exponent scales in the exp2 domain: -0.5 log2(e) / sigma^2 (0 if sigma <= 0)
    float sh;        // halo sigma (culling bound)
};

__device__ __forceinline__ VoxTerms vox_terms(const Model &m, const LayerC &c, float wet, float dsrc) {
    float idd, sms;
    lut_lookup(m, c.eid, wet, idd, sms);
    const float r80 = c.r80;
    const float d = dsrc - wet + 0.7f * r80;
    const float st = fmaf(fmaf(c.A, d, c.B), d, c.C) + sms;

    // nuclear halo (nuclearHalo, IMPTClasses.cu:104-121)
    const float w = fminf(fmaxf(wet, 0.1f), r80 - 0.1f);
    const float e5 = w - r80 - 5.0f;
    const float r3 = r80 + 3.0f;
    float sn = 2.85f + 0.0014f * r80 * fast_log(w + 3.0f) + 0.06f * w - 7.4e-5f * w * w -
               0.22f * r80 * fast_rcp(e5 * e5);
    float un = 0.052f * fast_log(1.13f + w * fast_rcp(11.2f - 0.023f * r80)) +
               0.35f * (0.0017f * r80 * r80 - r80) * fast_rcp(r3 * r3 - w * w) - 1.61e-9f * w * r3 * r3;
    sn = fmaxf(sn, 0.0f);
    un = fminf(fmaxf(un, 0.0f), 0.9f);

    const float st2 = st * st;
    const float sh2 = fmaf(sn, sn, st2);
    VoxTerms t;
    const bool pos1 = st > 0.0f, pos2 = sh2 > 0.0f;
    t.Af = pos1 ? (1.0f - un) * idd * fast_rcp(fmaf(kTwoPi, st2, 1e-8f)) : 0.0f;
    t.Bf = pos2 ? un * idd * fast_rcp(fmaf(kTwoPi, sh2, 1e-8f)) : 0.0f;
    t.a1 = pos1 ? (-0.5f * kLog2e) * fast_rcp(st2) : 0.0f;
    t.a2 = pos2 ? (-0.5f * kLog2e) * fast_rcp(sh2) : 0.0f;
    t.sh = fast_sqrt(sh2);
    return t;
}

// Squared distance from head-frame point (hx, hy, hz) to a spot axis.
__device__ __forceinline__ float axis_d2(float hx, float hy, float hz, const SpotExt &q) {
    const float u = fmaf(hz, q.tx, hx - q.x);
    const float v = fmaf(hz, q.ty, hy - q.y);
    const float w = fmaf(q.m12x2, v, q.m11 * u);
    return fmaf(u, w, (q.m22 * v) * v);
}

__device__ __forceinline__ void brick_of(const BrickGrid &bg, long long brick, int &bi, int &bj, int &bk) {
    bi = (int)(brick % bg.bx);
    const long long t = brick / bg.bx;
    bj = (int)(t % bg.by);
    bk = (int)(t / bg.by);
}

// ---------------------------------------------------------------------------
// Forward.  One wave per 4x4x4 brick, one voxel per lane; a lane owns its
// voxel's sum (deterministic, no atomics).  The lanes first judge 64 layers at
// a time (reaches the brick's WET range, and with culling: some spot of the
// layer within reach of the brick); the wave then walks the surviving layers.
// Per layer, each 64-spot chunk is judged lane-parallel against the brick
// (distance from the brick centre to the spot axis vs cutoff * max halo sigma
// + brick radius) and the surviving spots are evaluated four at a time, their
// records coming in through scalar loads.
//
// COUNT: no dose; counts[0] += in-range voxels x layer spots (the reference's
// work), counts[1] += in-range voxels x spots evaluated here.
__device__ __forceinline__ void spot_eval(const SpotExt &q, float hx, float hy, float hz, float a1, float a2,
                                          float &P, float &H) {
    const float d2 = axis_d2(hx, hy, hz, q);
    P = fmaf(q.mu, __builtin_amdgcn_exp2f(a1 * d2), P);
    H = fmaf(q.mu, __builtin_amdgcn_exp2f(a2 * d2), H);
}

template <bool CULL, bool COUNT = false>
__global__ __launch_bounds__(256) void k_pb_forward(const float *__restrict__ wetv, float *__restrict__ dose, Grid g,
                                                   Geom b, BrickGrid bg, unsigned nblocks, Model m,
                                                   const LayerC *__restrict__ lc, int n_layers,
                                                   const SpotExt *__restrict__ sp, float cut, float rbrick,
                                                   int accumulate, unsigned long long *__restrict__ counts = nullptr) {
    // plain dispatch order: consecutive bricks go to different XCDs, which
    // spreads the heavy bricks of the beam path over the whole chip
    const long long brick = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (brick >= bg.count()) return;  // whole wave leaves together
    const int lane = threadIdx.x & 63;
    int bi, bj, bk;
    brick_of(bg, brick, bi, bj, bk);
    const int i = bi * BRICK + (lane & 3), j = bj * BRICK + ((lane >> 2) & 3), k = bk * BRICK + (lane >> 4);
    const bool valid = i < g.nx && j < g.ny && k < g.nz;
    const long long idx = valid ? i + (long long)g.nx * (j + (long long)g.ny * k) : 0;
    const float wet = valid ? wetv[idx] * 10.0f : 3.0e38f;

    const float3 v = voxel_xyz(g, b, i, j, k);
    const float3 h = image_to_head(b, v);
    const float sx = b.src[0] - v.x, sy = b.src[1] - v.y, sz = b.src[2] - v.z;
    const float dsrc = sqrtf(sx * sx + sy * sy + sz * sz);
    const float ivx = 1.0f / b.vsadx, ivy = 1.0f / b.vsady;
    const BrickInfo B = brick_info_at(b, brick_centre(g, b, bi, bj, bk), wave_min(wet),
                                      wave_max(valid ? wet : -3.0e38f));

    float acc = 0.0f;
    unsigned long long n_dense = 0, n_exec = 0;
    for (int l0 = 0; l0 < n_layers; l0 += WAVE) {
        bool lrel = false;
        if (l0 + lane < n_layers) {
            const LayerC c = lc[l0 + lane];
            lrel = c.ok && B.wlo <= c.lim;
            if (CULL && lrel) lrel = !reach_skip(c, B, ivx, ivy, cut, rbrick);
        }
        unsigned long long lmask = __ballot(lrel);
        while (lmask) {
            const int l = l0 + __builtin_ctzll(lmask);
            lmask &= lmask - 1ull;
            const LayerC c = lc[l];
            const bool act = wet <= c.lim;
            const unsigned long long amask = __ballot(act);
            if (!amask) continue;
            const VoxTerms T = vox_terms(m, c, wet, dsrc);
            float rc2 = 0.0f;
            if (CULL) {
                const float rc = fmaf(cut, wave_max(act ? T.sh : 0.0f), rbrick);
                rc2 = rc * rc;
            }
            float P = 0.0f, H = 0.0f;
            const SpotExt *lsp = sp + c.start;
            for (int c0 = 0; c0 < c.n; c0 += WAVE) {
                const int nin = min(WAVE, c.n - c0);
                unsigned long long mask;
                if (CULL) {
                    bool rel = false;
                    if (lane < nin) rel = axis_d2(B.hc.x, B.hc.y, B.hc.z, lsp[c0 + lane]) < rc2;
                    mask = __ballot(rel);
                } else {
                    mask = nin == WAVE ? ~0ull : ((1ull << nin) - 1ull);
                }
                if (COUNT) {
                    n_exec += (unsigned long long)__popcll(amask) * (unsigned)__popcll(mask);
                    continue;
                }
                const SpotExt *csp = lsp + c0;
                while (__popcll(mask) >= 4) {
                    const int s0 = __builtin_ctzll(mask);
                    mask &= mask - 1ull;
                    const int s1 = __builtin_ctzll(mask);
                    mask &= mask - 1ull;
                    const int s2 = __builtin_ctzll(mask);
                    mask &= mask - 1ull;
                    const int s3 = __builtin_ctzll(mask);
                    mask &= mask - 1ull;
                    const SpotExt q0 = csp[s0], q1 = csp[s1], q2 = csp[s2], q3 = csp[s3];
                    spot_eval(q0, h.x, h.y, h.z, T.a1, T.a2, P, H);
                    spot_eval(q1, h.x, h.y, h.z, T.a1, T.a2, P, H);
                    spot_eval(q2, h.x, h.y, h.z, T.a1, T.a2, P, H);
                    spot_eval(q3, h.x, h.y, h.z, T.a1, T.a2, P, H);
                }
                while (mask) {
                    const int s0 = __builtin_ctzll(mask);
                    mask &= mask - 1ull;
                    spot_eval(csp[s0], h.x, h.y, h.z, T.a1, T.a2, P, H);
                }
            }
            if (act) acc = fmaf(T.Af, P, fmaf(T.Bf, H, acc));
        }
    }
    if (COUNT) {
        for (int l = 0; l < n_layers; ++l) {
            const LayerC c = lc[l];
            if (c.ok) n_dense += (unsigned long long)__popcll(__ballot(wet <= c.lim)) * (unsigned)c.n;
        }
        if (lane == 0) {
            atomicAdd(&counts[0], n_dense);
            atomicAdd(&counts[1], n_exec);
        }
        return;
    }
    if (valid) dose[idx] = accumulate ? dose[idx] + acc : acc;
}

// ---------------------------------------------------------------------------
// One wave per brick: WET range (mm) of the voxels with dL/dD != 0; an empty
// brick gets (+big, -big).
__global__ __launch_bounds__(256) void k_brick_summary(const float *__restrict__ wetv,
                                                      const float *__restrict__ gdose, Grid g, BrickGrid bg,
                                                      float2 *__restrict__ out) {
    const long long brick = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (brick >= bg.count()) return;
    const int lane = threadIdx.x & 63;
    int bi, bj, bk;
    brick_of(bg, brick, bi, bj, bk);
    const int i = bi * BRICK + (lane & 3), j = bj * BRICK + ((lane >> 2) & 3), k = bk * BRICK + (lane >> 4);
    float lo = 3.0e38f, hi = -3.0e38f;
    if (i < g.nx && j < g.ny && k < g.nz) {
        const long long idx = i + (long long)g.nx * (j + (long long)g.ny * k);
        if (gdose[idx] != 0.0f) lo = hi = wetv[idx] * 10.0f;
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    if (lane == 0) out[brick] = make_float2(lo, hi);
}

// ---------------------------------------------------------------------------
// Adjoint.  Workgroup = one wave = one 64-spot chunk of one layer x one slab of
// bricks.  Lane = spot: its three sums stay in registers across the whole
// slab.  The lanes judge 64 bricks at a time (holds voxels with dL/dD != 0
// within the layer's range, and with culling: within reach of the chunk);
// per surviving brick the wave acts as lanes = voxels to compute the voxel
// terms, compacts the voxels that matter into LDS, then every spot lane walks
// them with broadcast LDS reads.
//   dL/dmu_s = sum_v g_v [A e1 + B e2]
//   dL/dx_s  = -2 mu_s sum_v g_v [A a1 e1 + B a2 e2] (1 + q/VSADX) p_x
//   dL/dy_s  = -2 mu_s sum_v g_v [A a1 e1 + B a2 e2] (1 + q/VSADY) p_y
// with r = (hx - x_s, hy - y_s, hz) the voxel relative to the spot,
// q = (t . r) / |t|^2, p = r - q t (its part perpendicular to the axis) and
// a1, a2 = -1/(2 sigma^2) (natural-log scale).
//
// COUNT: no gradient; counts[2] += voxels (dL/dD != 0, in range) x chunk spots
// (the reference's work), counts[3] += those voxels x spots evaluated here.
struct LdsVox {
    float hx, hy, hz, a1;
    float a2, GA, GB, GAa;
    float GBb, pad0, pad1, pad2;
};

template <bool CULL, int CS, bool COUNT = false>
__global__ __launch_bounds__(64) void k_pb_adjoint(const float *__restrict__ wetv, const float *__restrict__ gdose,
                                                  Grid g, Geom b, BrickGrid bg, Model m,
                                                  const LayerC *__restrict__ lc, int n_layers,
                                                  const SpotExt *__restrict__ sp, int n_spots, float cut,
                                                  float rbrick, int nslab, const float2 *__restrict__ bsum,
                                                  float4 *__restrict__ partial,
                                                  unsigned long long *__restrict__ counts = nullptr) {
    __shared__ LdsVox s_vox[WAVE];
    constexpr int G = WAVE / CS;  // voxel subsets per wave: lane = group * CS + slot
    const int lane = threadIdx.x;
    const int slot = lane % CS, group = lane / CS;
    // chunk -> (layer, first spot of the chunk within the layer)
    int cid = blockIdx.x, l = 0, c0 = -1;
    for (; l < n_layers; ++l) {
        const int nch = (lc[l].n + CS - 1) / CS;
        if (cid < nch) {
            c0 = cid * CS;
            break;
        }
        cid -= nch;
    }
    if (c0 < 0) return;  // surplus workgroup of the upper-bound grid
    const LayerC L = lc[l];
    if (!L.ok) return;
    const int slab = blockIdx.y;
    const bool sval = c0 + slot < L.n;
    const int spot = L.start + c0 + slot;
    SpotExt q;
    if (sval) q = sp[spot];
    else q = SpotExt{0.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 1.f};
    const float itt = 1.0f / (q.tx * q.tx + q.ty * q.ty + 1.0f);
    const float ivx = 1.0f / b.vsadx, ivy = 1.0f / b.vsady;
    const unsigned long long chunk_n = (unsigned long long)min(CS, L.n - c0);
    LayerC R = L;  // reach of this chunk's spots only
    if (CULL) reach_of(sp, L.start + c0, (int)chunk_n, m, L.eid, lane, R);

    const long long nb = bg.count();
    const long long b0 = nb * slab / nslab, b1 = nb * (slab + 1) / nslab;
    float K = 0.0f, FX = 0.0f, FY = 0.0f;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    unsigned long long n_dense = 0, n_exec = 0;

    for (long long base = b0; base < b1; base += WAVE) {
        bool brel = false;
        if (base + lane < b1) {
            const float2 ws = bsum[base + lane];
            brel = ws.x <= ws.y && ws.x <= L.lim;
            if (CULL && brel && !COUNT) {
                int ti, tj, tk;
                brick_of(bg, base + lane, ti, tj, tk);
                brel = !reach_skip(R, brick_info_at(b, brick_centre(g, b, ti, tj, tk), ws.x, ws.y), ivx, ivy, cut,
                                   rbrick);
            }
        }
        unsigned long long fm = __ballot(brel);
        while (fm) {
            const long long brick = base + __builtin_ctzll(fm);
            fm &= fm - 1ull;
            int bi, bj, bk;
            brick_of(bg, brick, bi, bj, bk);
            const int i = bi * BRICK + (lane & 3), j = bj * BRICK + ((lane >> 2) & 3), k = bk * BRICK + (lane >> 4);
            const bool valid = i < g.nx && j < g.ny && k < g.nz;
            const long long idx = valid ? i + (long long)g.nx * (j + (long long)g.ny * k) : 0;
            const float gv = valid ? gdose[idx] : 0.0f;
            const float wet = valid ? wetv[idx] * 10.0f : 3.0e38f;
            const bool act = valid && gv != 0.0f && wet <= L.lim;
            const unsigned long long am = __ballot(act);
            if (!am) continue;
            if (COUNT) n_dense += (unsigned long long)__popcll(am) * chunk_n;

            const float3 v = voxel_xyz(g, b, i, j, k);
            const float3 h = image_to_head(b, v);
            const float sx = b.src[0] - v.x, sy = b.src[1] - v.y, sz = b.src[2] - v.z;
            const float dsrc = sqrtf(sx * sx + sy * sy + sz * sz);
            const VoxTerms T = vox_terms(m, L, wet, dsrc);

            bool rel = sval;
            if (CULL) {
                const float rc = fmaf(cut, wave_max(act ? T.sh : 0.0f), rbrick);
                const float3 hc = image_to_head(b, brick_centre(g, b, bi, bj, bk));
                rel = sval && axis_d2(hc.x, hc.y, hc.z, q) < rc * rc;
            }
            if (COUNT) {
                n_exec += (unsigned long long)__popcll(am) * (unsigned)__popcll(__ballot(rel && group == 0));
                continue;
            }
            if (!__ballot(rel)) continue;
            if (act) {
                const int pos = __popcll(am & lt_mask);
                LdsVox e;
                e.hx = h.x;
                e.hy = h.y;
                e.hz = h.z;
                e.a1 = T.a1;
                e.a2 = T.a2;
                e.GA = gv * T.Af;
                e.GB = gv * T.Bf;
                // derivative scales in natural-log units: a_nat = a_exp2 / log2(e)
                e.GAa = e.GA * (T.a1 * (1.0f / kLog2e));
                e.GBb = e.GB * (T.a2 * (1.0f / kLog2e));
                e.pad0 = e.pad1 = e.pad2 = 0.0f;
                s_vox[pos] = e;
            }
            __syncthreads();
            const int na = __popcll(am);
            if (rel) {
                for (int t = group; t < na; t += G) {
                    const LdsVox e = s_vox[t];
                    // in-plane offset e_xy, its part perpendicular to the axis p,
                    // q_e = (e.t)/|t|^2; d2 = e_xy . p (no cancellation)
                    const float u = fmaf(e.hz, q.tx, e.hx - q.x);
                    const float w = fmaf(e.hz, q.ty, e.hy - q.y);
                    const float qe = fmaf(u, q.tx, w * q.ty) * itt;
                    const float px = fmaf(-qe, q.tx, u), py = fmaf(-qe, q.ty, w);
                    const float d2 = fmaf(u, px, w * py);
                    const float e1 = __builtin_amdgcn_exp2f(e.a1 * d2);
                    const float e2 = __builtin_amdgcn_exp2f(e.a2 * d2);
                    K = fmaf(e.GA, e1, K);
                    K = fmaf(e.GB, e2, K);
                    const float F = fmaf(e.GAa, e1, e.GBb * e2);
                    // q of the ray r = e - hz t is qe - hz
                    const float qr = qe - e.hz;
                    FX = fmaf(F * fmaf(qr, ivx, 1.0f), px, FX);
                    FY = fmaf(F * fmaf(qr, ivy, 1.0f), py, FY);
                }
            }
            __syncthreads();
        }
    }
    if (COUNT) {
        if (lane == 0) {
            atomicAdd(&counts[2], n_dense);
            atomicAdd(&counts[3], n_exec);
        }
        return;
    }
    // fold the voxel subsets (fixed order: deterministic)
#pragma unroll
    for (int off = CS; off < WAVE; off <<= 1) {
        K += __shfl_xor(K, off);
        FX += __shfl_xor(FX, off);
        FY += __shfl_xor(FY, off);
    }
    if (sval && group == 0) partial[(long long)slab * n_spots + spot] = make_float4(K, FX, FY, 0.0f);
}

// Spots that no chunk covers keep the zeroed partials (k_zero4).
__global__ __launch_bounds__(256) void k_zero4(float4 *__restrict__ p, long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// 256 threads = 64 spots x 4 slab groups; each group sums every 4th slab,
// then the 4 group sums are added in a fixed order.
__global__ __launch_bounds__(256) void k_pb_reduce(const float4 *__restrict__ partial, const SpotExt *__restrict__ sp,
                                                  int n_spots, int nslab, float *__restrict__ grad_mu,
                                                  float *__restrict__ grad_xy) {
    __shared__ float4 s_part[4][64];
    const int sl = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const int s = blockIdx.x * 64 + sl;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    if (s < n_spots) {
        for (int d = grp; d < nslab; d += 4) {
            const float4 p = partial[(long long)d * n_spots + s];
            acc.x += p.x;
            acc.y += p.y;
            acc.z += p.z;
        }
    }
    s_part[grp][sl] = acc;
    __syncthreads();
    if (grp == 0 && s < n_spots) {
        float K = 0.f, FX = 0.f, FY = 0.f;
        for (int q = 0; q < 4; ++q) {
            K += s_part[q][sl].x;
            FX += s_part[q][sl].y;
            FY += s_part[q][sl].z;
        }
        const float mu2 = -2.0f * sp[s].mu;
        grad_mu[s] = K;
        grad_xy[2 * s] = mu2 * FX;
        grad_xy[2 * s + 1] = mu2 * FY;
    }
}

// ---------------------------------------------------------------------------
static constexpr long long kAlign = 256;
static long long align_up(long long x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Slabs of bricks per spot chunk: enough (chunk, slab) workgroups (~32k) that
// the few heavy ones (chunks and slabs in the beam path) do not leave the chip
// idle at the end, with at least 64 bricks per slab.
// Spots per adjoint workgroup: 64, or 32 (with two voxel subsets per wave) when
// the average layer is small enough that 64-spot chunks would leave many lanes
// idle.  Scores: lane use / (1 + per-brick overhead share), the overhead share
// growing as chunks shrink.
static int adjoint_chunk(int n_spots, int n_layers) {
    if (n_layers <= 0) return 64;
    const double n = (double)n_spots / n_layers;
    const double u64 = n / (ceil(n / 64.0) * 64.0), u32 = n / (ceil(n / 32.0) * 32.0);
    return (u32 / 1.38 > u64 / 1.23) ? 32 : 64;
}

static int adjoint_slabs(const dh_grid &g, int n_spots, int n_layers) {
    const long long nb = (long long)((g.nx + 3) / 4) * ((g.ny + 3) / 4) * ((g.nz + 3) / 4);
    const int cs = adjoint_chunk(n_spots, n_layers);
    const long long nch = ((long long)n_spots + (cs - 1LL) * n_layers) / cs + 1;
    long long d = (32768 + nch - 1) / nch;
    if (d > 1024) d = 1024;
    if (d > nb / 64) d = nb / 64;
    if (d > nb) d = nb;
    if (d < 1) d = 1;
    return (int)d;
}

// Workspace: [SpotExt x S][LayerC x L][adjoint partials float4 x nslab*S][brick summaries float2]
struct Workspace {
    SpotExt *ext;
    LayerC *layer;
    float4 *partial;
    float2 *bsum;
    int nslab;
    long long bytes;
};

static Workspace carve(void *base, const dh_grid &g, int S, int L) {
    Workspace w;
    char *p = reinterpret_cast<char *>(base);
    w.nslab = adjoint_slabs(g, S, L);
    const long long nb = (long long)((g.nx + 3) / 4) * ((g.ny + 3) / 4) * ((g.nz + 3) / 4);
    long long off = 0;
    w.ext = reinterpret_cast<SpotExt *>(p + off);
    off += align_up((long long)S * (long long)sizeof(SpotExt));
    w.layer = reinterpret_cast<LayerC *>(p + off);
    off += align_up((long long)(L > 0 ? L : 1) * (long long)sizeof(LayerC));
    w.partial = reinterpret_cast<float4 *>(p + off);
    off += align_up((long long)w.nslab * S * 16LL);
    w.bsum = reinterpret_cast<float2 *>(p + off);
    off += align_up(nb * 8LL);
    w.bytes = off;
    return w;
}

static bool fill_model(const dh_beam_model *bm, Model &m) {
    if (!bm || !bm->idd_sigma || !bm->dvp || bm->n_energies <= 0 || bm->lut_len < 2 || bm->dvp_len < 5 ||
        (!bm->depth_uniform && !bm->depths) || (bm->depth_uniform && !(bm->ddepth > 0.0f))) {
        set_error("invalid beam model");
        return false;
    }
    m.idd_sigma = reinterpret_cast<const float2 *>(bm->idd_sigma);
    m.depths = bm->depths;
    m.dvp = bm->dvp;
    m.n_energies = bm->n_energies;
    m.lut_len = bm->lut_len;
    m.dvp_len = bm->dvp_len;
    m.uniform = bm->depth_uniform ? 1 : 0;
    m.depth0 = bm->depth0;
    m.inv_ddepth = bm->depth_uniform ? 1.0f / bm->ddepth : 0.0f;
    return true;
}

static bool common_args(const dh_grid &g, const dh_beam_geom *geom, const dh_spot_set *ss, void *ws,
                        long long ws_bytes) {
    if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0 || !(g.spacing > 0.0f) || !geom || !ss) {
        set_error("invalid grid, geometry or spot set");
        return false;
    }
    if ((double)g.nx * g.ny * g.nz >= 2147483647.0) {
        set_error("grid %d x %d x %d too large (below 2^31 voxels required)", g.nx, g.ny, g.nz);
        return false;
    }
    if (ss->n_spots < 0 || ss->n_layers < 0 || (ss->n_spots > 0 && (!ss->spots || !ss->layers))) {
        set_error("invalid spot set (n_spots=%d, n_layers=%d)", ss->n_spots, ss->n_layers);
        return false;
    }
    const long long need = dh_pb_workspace_size(ss->n_spots, ss->n_layers, g);
    if (ss->n_spots > 0 && (!ws || ws_bytes < need || ((uintptr_t)ws % kAlign) != 0)) {
        set_error("workspace too small or misaligned (%lld bytes given, %lld needed, 256-byte alignment)",
                  (long long)ws_bytes, need);
        return false;
    }
    return true;
}

// Spot records and per-layer reach into the workspace.
static int prepare(hipStream_t st, const dh_spot_set *spots, const dh_beam_geom *geom, const Model &m,
                   const Workspace &w) {
    const int S = spots->n_spots, L = spots->n_layers;
    hipLaunchKernelGGL(k_spot_prep, dim3((S + 255) / 256), dim3(256), 0, st, spots->spots, w.ext, S,
                       1.0f / geom->vsadx, 1.0f / geom->vsady);
    int rc = check_launch("k_spot_prep");
    if (rc || L == 0) return rc;
    hipLaunchKernelGGL(k_layer_prep, dim3(L), dim3(64), 0, st, w.ext, S, spots->layers, m, w.layer);
    return check_launch("k_layer_prep");
}

static float brick_radius(const dh_grid &g) { return 0.5f * (BRICK - 1) * 1.7320508075688772f * g.spacing * 1.0001f; }

template <bool COUNT>
static int launch_forward(hipStream_t st, const float *wet, float *dose, const dh_grid &g, const dh_beam_geom *geom,
                          const Model &m, const dh_spot_set *spots, const Workspace &w, float cutoff_sigma,
                          int accumulate, unsigned long long *counts) {
    const Grid gg = make_grid(g);
    const Geom gb = make_geom(*geom);
    const BrickGrid bg = brick_grid(gg);
    const unsigned nb = (unsigned)((bg.count() + 3) / 4);
    const int S = spots->n_spots;
    const int nl = S > 0 ? spots->n_layers : 0;
    if (cutoff_sigma > 0.0f)
        hipLaunchKernelGGL((k_pb_forward<true, COUNT>), dim3(nb), dim3(256), 0, st, wet, dose, gg, gb, bg, nb, m,
                           w.layer, nl, w.ext, cutoff_sigma, brick_radius(g), accumulate, counts);
    else
        hipLaunchKernelGGL((k_pb_forward<false, COUNT>), dim3(nb), dim3(256), 0, st, wet, dose, gg, gb, bg, nb, m,
                           w.layer, nl, w.ext, 0.0f, brick_radius(g), accumulate, counts);
    return check_launch(COUNT ? "k_pb_forward(count)" : "k_pb_forward");
}

template <bool COUNT>
static int launch_adjoint(hipStream_t st, const float *wet, const float *grad_dose, const dh_grid &g,
                          const dh_beam_geom *geom, const Model &m, const dh_spot_set *spots, const Workspace &w,
                          float cutoff_sigma, unsigned long long *counts) {
    const Grid gg = make_grid(g);
    const Geom gb = make_geom(*geom);
    const BrickGrid bg = brick_grid(gg);
    const int S = spots->n_spots;
    hipLaunchKernelGGL(k_brick_summary, dim3((unsigned)((bg.count() + 3) / 4)), dim3(256), 0, st, wet, grad_dose,
                       gg, bg, w.bsum);
    int rc = check_launch("k_brick_summary");
    if (rc) return rc;
    const int cs = adjoint_chunk(S, spots->n_layers);
    const unsigned nch = (unsigned)(((long long)S + (cs - 1LL) * spots->n_layers) / cs + 1);
    const dim3 grid(nch, w.nslab);
    const float rb = brick_radius(g);
    const bool cull = cutoff_sigma > 0.0f;
    if (cs == 32) {
        if (cull)
            hipLaunchKernelGGL((k_pb_adjoint<true, 32, COUNT>), grid, dim3(64), 0, st, wet, grad_dose, gg, gb, bg, m,
                               w.layer, spots->n_layers, w.ext, S, cutoff_sigma, rb, w.nslab, w.bsum, w.partial,
                               counts);
        else
            hipLaunchKernelGGL((k_pb_adjoint<false, 32, COUNT>), grid, dim3(64), 0, st, wet, grad_dose, gg, gb, bg, m,
                               w.layer, spots->n_layers, w.ext, S, 0.0f, rb, w.nslab, w.bsum, w.partial, counts);
    } else {
        if (cull)
            hipLaunchKernelGGL((k_pb_adjoint<true, 64, COUNT>), grid, dim3(64), 0, st, wet, grad_dose, gg, gb, bg, m,
                               w.layer, spots->n_layers, w.ext, S, cutoff_sigma, rb, w.nslab, w.bsum, w.partial,
                               counts);
        else
            hipLaunchKernelGGL((k_pb_adjoint<false, 64, COUNT>), grid, dim3(64), 0, st, wet, grad_dose, gg, gb, bg, m,
                               w.layer, spots->n_layers, w.ext, S, 0.0f, rb, w.nslab, w.bsum, w.partial, counts);
    }
    return check_launch(COUNT ? "k_pb_adjoint(count)" : "k_pb_adjoint");
}

}  // namespace dh

using namespace dh;

extern "C" int64_t dh_pb_workspace_size(int32_t n_spots, int32_t n_layers, dh_grid g) {
    if (n_spots <= 0) return 0;
    return carve(nullptr, g, n_spots, n_layers).bytes;
}

extern "C" int dh_pb_forward(const float *wet, float *dose, dh_grid g, const dh_beam_geom *geom,
                             const dh_beam_model *model, const dh_spot_set *spots, float cutoff_sigma,
                             int32_t accumulate, void *workspace, int64_t workspace_bytes, void *stream) {
    if (!common_args(g, geom, spots, workspace, workspace_bytes)) return DH_ERR_ARG;
    if (!wet || !dose) {
        set_error("dh_pb_forward: null volume pointer");
        return DH_ERR_ARG;
    }
    Model m;
    if (!fill_model(model, m)) return DH_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const Workspace w = carve(workspace, g, spots->n_spots, spots->n_layers);
    if (spots->n_spots > 0) {
        const int rc = prepare(st, spots, geom, m, w);
        if (rc) return rc;
    }
    return launch_forward<false>(st, wet, dose, g, geom, m, spots, w, cutoff_sigma, accumulate, nullptr);
}

extern "C" int dh_pb_adjoint(const float *wet, const float *grad_dose, dh_grid g, const dh_beam_geom *geom,
                             const dh_beam_model *model, const dh_spot_set *spots, float cutoff_sigma,
                             float *grad_mu, float *grad_xy, void *workspace, int64_t workspace_bytes,
                             void *stream) {
    if (!common_args(g, geom, spots, workspace, workspace_bytes)) return DH_ERR_ARG;
    const int S = spots->n_spots;
    if (S == 0) return DH_OK;
    if (!wet || !grad_dose || !grad_mu || !grad_xy) {
        set_error("dh_pb_adjoint: null pointer");
        return DH_ERR_ARG;
    }
    Model m;
    if (!fill_model(model, m)) return DH_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const Workspace w = carve(workspace, g, S, spots->n_layers);
    int rc = prepare(st, spots, geom, m, w);
    if (rc) return rc;
    const long long np = (long long)w.nslab * S;
    hipLaunchKernelGGL(k_zero4, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, w.partial, np);
    rc = check_launch("k_zero4");
    if (rc) return rc;
    rc = launch_adjoint<false>(st, wet, grad_dose, g, geom, m, spots, w, cutoff_sigma, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pb_reduce, dim3((S + 63) / 64), dim3(256), 0, st, w.partial, w.ext, S, w.nslab, grad_mu,
                       grad_xy);
    return check_launch("k_pb_reduce");
}

extern "C" int dh_pb_count_pairs(const float *wet, const float *grad_dose, dh_grid g, const dh_beam_geom *geom,
                                 const dh_beam_model *model, const dh_spot_set *spots, float cutoff_sigma,
                                 unsigned long long *counts, void *workspace, int64_t workspace_bytes,
                                 void *stream) {
    if (!common_args(g, geom, spots, workspace, workspace_bytes)) return DH_ERR_ARG;
    if (!wet || !counts) {
        set_error("dh_pb_count_pairs: null pointer");
        return DH_ERR_ARG;
    }
    const int S = spots->n_spots;
    if (S == 0) return DH_OK;  // counts are accumulated into; nothing to add
    Model m;
    if (!fill_model(model, m)) return DH_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const Workspace w = carve(workspace, g, S, spots->n_layers);
    int rc = prepare(st, spots, geom, m, w);
    if (rc) return rc;
    rc = launch_forward<true>(st, wet, nullptr, g, geom, m, spots, w, cutoff_sigma, 0, counts);
    if (rc || !grad_dose) return rc;
    return launch_adjoint<true>(st, wet, grad_dose, g, geom, m, spots, w, cutoff_sigma, counts);
}
