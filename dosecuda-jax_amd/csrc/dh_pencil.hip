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
// Each kernel has a batched twin (*_batch) that walks the items of many beams
// (and plans) in one launch, reading per-beam contexts from a device table;
// both run the same per-item device function, so their results are identical.
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
// lane-parallel: the 64 lanes of a wave test 64 candidates (layers, spots or
// bricks) at once and the wave then walks only the survivors.
#include "dh_common.h"

#include <math.h>

#include <mutex>

namespace dh {

constexpr float kLog2e = 1.4426950408889634f;
constexpr float kTwoPi = 6.283185307179586f;

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// natural log through v_log_f32 (log2)
__device__ __forceinline__ float fast_log(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994531f; }

// Largest value over the wave of a non-negative float, wave-uniform.  The bit
// patterns of non-negative floats order like integers, so an integer max is
// exact; the reduction runs on DPP lane moves (quad swaps, the row mirrors,
// then the row broadcasts), which leaves the maximum in lane 63.  Needs all 64
// lanes active.
__device__ __forceinline__ float wave_max_nonneg(float x) {
    int v = __float_as_int(x);
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false));  // row_mirror
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false));  // row_bcast15 -> rows 1, 3
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false));  // row_bcast31 -> rows 2, 3
    return __int_as_float(__builtin_amdgcn_readlane(v, 63));
}

// Exact wave minimum / maximum of any floats (no NaN), wave-uniform: the bit
// pattern mapped to an int whose signed order is the float order, reduced with
// DPP lane moves as wave_max_nonneg (all 64 lanes active).
__device__ __forceinline__ int float_key(float x) {
    const int b = __float_as_int(x);
    return b ^ ((b >> 31) & 0x7fffffff);
}
__device__ __forceinline__ float key_float(int k) { return __int_as_float(k ^ ((k >> 31) & 0x7fffffff)); }

__device__ __forceinline__ float wave_min_dpp(float x) {
    int v = float_key(x);
    constexpr int kMax = 0x7fffffff;
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0xB1, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x4E, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x141, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x140, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x142, 0xA, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(kMax, v, 0x143, 0xC, 0xF, false));
    return key_float(__builtin_amdgcn_readlane(v, 63));
}

__device__ __forceinline__ float wave_max_dpp(float x) {
    int v = float_key(x);
    constexpr int kMin = (int)0x80000000;
    v = max(v, __builtin_amdgcn_update_dpp(kMin, v, 0xB1, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(kMin, v, 0x4E, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(kMin, v, 0x141, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(kMin, v, 0x140, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(kMin, v, 0x142, 0xA, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(kMin, v, 0x143, 0xC, 0xF, false));
    return key_float(__builtin_amdgcn_readlane(v, 63));
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off));
    return v;
}

// Spot part of a reach record: bounding box and smallest distance-form
// eigenvalue of spots [start, start + n) (lane-parallel) into c.
__device__ __forceinline__ void reach_spots(const SpotExt *__restrict__ sp, int start, int n, int lane, LayerC &c) {
    float xmin = 3.0e38f, xmax = -3.0e38f, ymin = 3.0e38f, ymax = -3.0e38f, t2max = 0.0f;
    for (int s = lane; s < n; s += WAVE) {
        const SpotExt q = gld(sp, start + s);
        xmin = fminf(xmin, q.x);
        xmax = fmaxf(xmax, q.x);
        ymin = fminf(ymin, q.y);
        ymax = fmaxf(ymax, q.y);
        t2max = fmaxf(t2max, q.tx * q.tx + q.ty * q.ty);
    }
    c.xmin = wave_min(xmin);
    c.xmax = wave_max(xmax);
    c.ymin = wave_min(ymin);
    c.ymax = wave_max(ymax);
    c.lam_min = 1.0f / (1.0f + wave_max(t2max));
}

// Reach of spots [start, start + n) of energy eid (lane-parallel) into c.
__device__ __forceinline__ void reach_of(const SpotExt *__restrict__ sp, int start, int n, const Model &m, int eid,
                                         int lane, LayerC &c) {
    reach_spots(sp, start, n, lane, c);
    float smax = 0.0f;
    const float2 *row = gptr(m.idd_sigma, (long long)eid * m.lut_len);
    for (int t = lane; t < m.lut_len; t += WAVE) smax = fmaxf(smax, fabsf(row[t].y));
    c.sms_max = wave_max(smax);
}

// Reach record of a chunk [c0, c0 + n) of usable layer L: the layer's own record
// when the chunk is the whole layer (the same computation over the same spots),
// else the chunk's spot part with the layer's max |sigma_ms| (it depends on the
// energy only).
__device__ __forceinline__ LayerC chunk_reach(const SpotExt *__restrict__ sp, const LayerC &L, int c0, int n,
                                              int lane) {
    LayerC R = L;
    if (c0 != 0 || n != L.n) reach_spots(sp, L.start + c0, n, lane, R);
    return R;
}

// LUT lookup (interpolateProtonLUT, IMPTClasses.cu:69-92): i = first depth >= wet;
// i == 0 -> row 0, i == len -> last row, else lerp between i-1 and i.
__device__ __forceinline__ void lut_lookup(const Model &m, int eid, float wet, float &idd, float &sig) {
    const float2 *row = gptr(m.idd_sigma, (long long)eid * m.lut_len);
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
        const float *dep = gptr(m.depths, (long long)eid * L);
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

// The lookup above on a uniform depth grid with rows [i0, i0 + 64) of the
// row staged in LDS (tab) by the wave: one coalesced 512-byte load per wave
// and visit instead of two scattered 8-byte gathers per lane.  i0 is the row
// of the smallest WET of the voxels that use the result, so their
// interpolations read staged rows unless the wave's WET spans 63 rows or more
// (those read the row from global memory).  Same values, same arithmetic:
// bit-identical to lut_lookup.
__device__ __forceinline__ int lut_stage(const Model &m, int eid, float wlo, int lane, float2 *tab) {
    const float q = (wlo - m.depth0) * m.inv_ddepth;
    const int i0 = (int)fminf(fmaxf(q, 0.0f), (float)(m.lut_len - 1));  // clamped before truncation
    const float2 *row = gptr(m.idd_sigma, (long long)eid * m.lut_len);
    tab[lane] = row[min(i0 + lane, m.lut_len - 1)];
    // the wave's lanes read each other's rows
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return i0;
}

__device__ __forceinline__ void lut_lookup_staged(const Model &m, int eid, float wet, const float2 *tab, int i0,
                                                  float &idd, float &sig) {
    const float2 *row = gptr(m.idd_sigma, (long long)eid * m.lut_len);
    const int L = m.lut_len;
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
    const int lo = min((int)q, L - 2);
    const float f = q - (float)lo;
    const int t = lo - i0;
    float2 a, c;
    if (t >= 0 && t < WAVE - 1) {
        a = tab[t];
        c = tab[t + 1];
    } else {
        a = row[lo];
        c = row[lo + 1];
    }
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

// Upper bound of |sigma_ms| from layer c's lookup row for lookups at WET in
// [wlo, whi] (mm): the largest |sigma| of the rows lut_lookup interpolates
// between there (on a uniform depth grid rows floor(q) and floor(q) + 1 of
// q = (wet - depth0) / ddepth, clamped to the row), else over the whole row
// (also when the range spans more than SMS_ROWS rows).  The lerp stays within
// its two rows up to rounding, which the 1.001 margin of reach_skip covers.
constexpr int SMS_ROWS = 24;
__device__ __forceinline__ float sms_bound(const Model &m, const LayerC &c, float wlo, float whi) {
    if (!m.uniform) return c.sms_max;
    const float qlo = (wlo - m.depth0) * m.inv_ddepth, qhi = (whi - m.depth0) * m.inv_ddepth;
    if (!(qhi - qlo < (float)SMS_ROWS)) return c.sms_max;
    // clamped before truncation (positive: truncation is floor); qhi <= 0 reads row 0 only, rows 0..1
    // are taken
    const int i0 = (int)fminf(fmaxf(qlo, 0.0f), (float)(m.lut_len - 2));
    const int i1 = (int)fminf(fmaxf(qhi, 0.0f), (float)(m.lut_len - 2)) + 1;
    const float2 *row = gptr(m.idd_sigma, (long long)c.eid * m.lut_len);
    float s = 0.0f;
    for (int t = i0; t <= i1; ++t) s = fmaxf(s, fabsf(row[t].y));
    return s;
}

// True when every spot of the reach record c is negligible for the whole
// brick: the brick centre's distance to the nearest spot axis, less the brick
// radius, exceeds cut * (an upper bound of the halo sigma over the brick).  The
// bound takes sigma_air over the range of d = |src - v| - wet + 0.7 R80 (a
// quadratic: ends and vertex) plus max |sigma_ms| over the lookup rows of the
// brick's WET range, and the halo width's increasing terms at the top of that
// range plus its decreasing terms at the bottom.
__device__ __forceinline__ bool reach_skip(const LayerC &c, const BrickInfo &B, const Model &m, float ivx, float ivy,
                                           float cut, float rbrick, float *sh_bound = nullptr) {
    if (sh_bound) *sh_bound = 3.0e38f;
    if (!(c.r80 >= 5.0f)) return false;
    const float fx = fmaf(-B.hc.z, ivx, 1.0f), fy = fmaf(-B.hc.z, ivy, 1.0f);
    const float ex = fmaxf(0.0f, fmaxf(c.xmin * fx - B.hc.x, B.hc.x - c.xmax * fx));
    const float ey = fmaxf(0.0f, fmaxf(c.ymin * fy - B.hc.y, B.hc.y - c.ymax * fy));
    const float dlb = 0.9999f * fast_sqrt(c.lam_min * (ex * ex + ey * ey)) - rbrick;
    if (!(dlb > 0.0f) && !sh_bound) return false;
    const float whi = fminf(B.whi, c.lim);
    // sigma_n = (increasing terms) + (decreasing terms) in the clamped WET w (R80 >= 5 here: the log term
    // grows; -7.4e-5 w^2 and -0.22 R80 / (w - R80 - 5)^2 fall for 0 < w < R80 + 5), so over the brick's
    // range it is at most the increasing terms at the largest WET plus the decreasing ones at the
    // smallest: a bound that only grows when the range widens
    const float w_ub = fminf(fmaxf(whi, 0.1f), c.w_top);
    const float w_lb = fminf(fmaxf(B.wlo, 0.1f), c.w_top);
    const float e5 = w_lb - c.r80 - 5.0f;
    const float sn_ub = fmaxf(2.85f + c.sn_log * fast_log(w_ub + 3.0f) + 0.06f * w_ub - 7.4e-5f * w_lb * w_lb -
                                  c.sn_pole * fast_rcp(e5 * e5),
                              0.0f);
    const float d07 = c.d07;
    const float dlo = B.dsrc - rbrick - whi + d07;
    const float dhi = B.dsrc + rbrick - B.wlo + d07;
    float sa = fmaxf(fabsf(fmaf(fmaf(c.A, dlo, c.B), dlo, c.C)), fabsf(fmaf(fmaf(c.A, dhi, c.B), dhi, c.C)));
    if (c.A != 0.0f && c.vx > dlo && c.vx < dhi) sa = fmaxf(sa, fabsf(fmaf(fmaf(c.A, c.vx, c.B), c.vx, c.C)));
    const float st_ub = sa + sms_bound(m, c, B.wlo, whi);
    const float sh_ub = 1.001f * fast_sqrt(st_ub * st_ub + sn_ub * sn_ub);
    if (sh_bound) *sh_bound = sh_ub;
    return dlb > cut * sh_ub;
}

// Per (voxel, layer) Gaussian parameters.
struct VoxTerms {
    float Af, Bf;    // primary / halo amplitude (idd * weight / (2 pi sigma^2 + 1e-8))
    float a1, a2;    // exponent scales in the exp2 domain: -0.5 log2(e) / sigma^2 (0 if sigma <= 0)
    float sh;        // halo sigma (culling bound)
    float st;        // primary sigma where positive, else 0 (culling bound of the primary term)
    float s2;        // sqrt(-a2)
    float r12;       // a1 / a2 = sigma_h^2 / sigma_t^2 where both are positive, else 0
};

// From the lookup's (IDD, sigma_ms) at wet.
__device__ __forceinline__ VoxTerms vox_terms(const Model &m, const LayerC &c, float wet, float dsrc, float idd,
                                              float sms) {
    const float r80 = c.r80;
    const float d = dsrc - wet + c.d07;
    const float st = fmaf(fmaf(c.A, d, c.B), d, c.C) + sms;

    // nuclear halo (nuclearHalo, IMPTClasses.cu:104-121); the R80-only factors come from the layer record
    const float w = fminf(fmaxf(wet, 0.1f), c.w_top);
    const float e5 = w - r80 - 5.0f;
    const float r3 = c.r3;
    float sn = 2.85f + c.sn_log * fast_log(w + 3.0f) + 0.06f * w - 7.4e-5f * w * w - c.sn_pole * fast_rcp(e5 * e5);
    float un = 0.052f * fast_log(1.13f + w * c.un_lin) + c.un_pole * fast_rcp(r3 * r3 - w * w) -
               1.61e-9f * w * r3 * r3;
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
    t.st = pos1 ? st : 0.0f;
    t.s2 = fast_sqrt(-t.a2);
    t.r12 = (pos1 && pos2) ? sh2 * fast_rcp(st2) : 0.0f;
    return t;
}

__device__ __forceinline__ VoxTerms vox_terms(const Model &m, const LayerC &c, float wet, float dsrc) {
    float idd, sms;
    lut_lookup(m, c.eid, wet, idd, sms);
    return vox_terms(m, c, wet, dsrc, idd, sms);
}

// Distance from a spot axis beyond which the voxel's halo term is below
// exp(-cut^2 / 2) of the spot's peak dose at the voxel's depth (primary plus
// halo on the axis, Af + Bf): cut_h sigma_h with cut_h^2 = cut^2 + 2 ln(Bf /
// (Af + Bf)) <= cut^2 (a margin of 1e-3 in cut_h^2 covers the rounding of the
// logarithm); 0 without a halo term.  A primary term is dropped beyond cut
// sigma_t, below exp(-cut^2 / 2) of the same peak, so every skipped term keeps
// that bound.
__device__ __forceinline__ float halo_reach(const VoxTerms &T, float cut) {
    if (!(T.Bf > 0.0f)) return 0.0f;
    const float ch2 = fmaf(2.0f, fast_log(T.Bf * fast_rcp(T.Af + T.Bf)), fmaf(cut, cut, 1e-3f));
    return T.sh * fast_sqrt(fmaxf(ch2, 0.0f));
}

// Squared distance from head-frame point (hx, hy, hz) to a spot axis.
__device__ __forceinline__ float axis_d2(float hx, float hy, float hz, const SpotExt &q) {
    const float u = fmaf(hz, q.tx, hx - q.x);
    const float v = fmaf(hz, q.ty, hy - q.y);
    const float w = fmaf(q.m12x2, v, q.m11 * u);
    return fmaf(u, w, (q.m22 * v) * v);
}

__device__ __forceinline__ void brick_of(const BrickGrid &bg, long long brick, int &bi, int &bj, int &bk) {
    brick_coords(bg, brick, bi, bj, bk);
}

// ---------------------------------------------------------------------------
// Preparation: spot records, layer records.
__device__ __forceinline__ void spot_prep_one(const dh_spot *__restrict__ in, SpotExt *__restrict__ out, int s,
                                              float inv_vsadx, float inv_vsady) {
    const dh_spot p = gld(in, s);
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
    gst(out, s, e);
}

__global__ __launch_bounds__(256) void k_spot_prep(BeamCtx c) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c.n_spots) return;
    spot_prep_one(c.spots_in, c.sp, s, c.b.ivsadx, c.b.ivsady);
}

// A layer record whose energy id and spot range the kernels can use.
__device__ __forceinline__ bool layer_in_range(const dh_layer &L, int n_energies, int n_spots) {
    return (unsigned)L.energy_id < (unsigned)n_energies && L.spot_start >= 0 && L.n_spots > 0 &&
           L.spot_start <= n_spots - L.n_spots;
}

// One wave per layer.  A record the kernels cannot use safely (energy id
// outside the tables, spot range outside the spot array, or a spot range that
// intersects an earlier in-range layer's) gets ok = 0 and is skipped rather
// than read out of bounds.  The usable layers are then disjoint runs, so their
// spot counts sum to at most n_spots: the adjoint's chunk grid (adjoint_chunks)
// covers them, and no spot's partials are written by two layers.
__device__ __forceinline__ void layer_prep_one(const BeamCtx &cx, int l, int lane) {
    const dh_layer L = gld(cx.layers_in, l);
    LayerC c;
    bool overlap = false;
    if (layer_in_range(L, cx.m.n_energies, cx.n_spots)) {
        for (int k0 = 0; k0 < l; k0 += WAVE) {
            const int k = k0 + lane;
            bool hit = false;
            if (k < l) {
                const dh_layer P = gld(cx.layers_in, k);
                hit = layer_in_range(P, cx.m.n_energies, cx.n_spots) && P.spot_start < L.spot_start + L.n_spots &&
                      L.spot_start < P.spot_start + P.n_spots;
            }
            overlap = overlap || __ballot(hit) != 0;
        }
    }
    c.ok = layer_in_range(L, cx.m.n_energies, cx.n_spots) && !overlap;
    c.start = c.ok ? L.spot_start : 0;
    c.n = c.ok ? L.n_spots : 0;
    c.eid = c.ok ? L.energy_id : 0;
    c.r80 = L.r80;
    c.lim = 1.1f * L.r80;
    const float *d = gptr(cx.m.dvp, (long long)c.eid * cx.m.dvp_len);
    c.A = d[2];
    c.B = d[3];
    c.C = d[4];
    c.vx = c.A != 0.0f ? -c.B / (2.0f * c.A) : 0.0f;
    const float r80 = c.r80;
    c.sn_log = 0.0014f * r80;
    c.sn_pole = 0.22f * r80;
    c.un_lin = fast_rcp(11.2f - 0.023f * r80);
    c.un_pole = 0.35f * (0.0017f * r80 * r80 - r80);
    c.r3 = r80 + 3.0f;
    c.d07 = 0.7f * r80;
    c.w_top = r80 - 0.1f;
    c.pad = 0.0f;
    if (c.ok) {
        reach_of(cx.sp, c.start, c.n, cx.m, c.eid, lane, c);
    } else {
        c.xmin = c.ymin = -3.0e38f;
        c.xmax = c.ymax = 3.0e38f;
        c.lam_min = 0.0f;
        c.sms_max = 0.0f;
    }
    if (lane == 0) gst(cx.lc, l, c);
}

__global__ __launch_bounds__(64) void k_layer_prep(BeamCtx c) { layer_prep_one(c, blockIdx.x, threadIdx.x); }

// Spot weights / positions from input-order arrays (one thread per sorted spot).
__global__ __launch_bounds__(256) void k_spot_params(dh_spot *__restrict__ spots, const long long *__restrict__ order,
                                                    const float *__restrict__ mu, const float *__restrict__ xy, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long r = gld(order, i);
    if ((unsigned long long)r >= (unsigned long long)n) return;
    dh_spot s = gld(spots, i);
    if (mu) s.mu = gld(mu, r);
    if (xy) {
        s.x = gld(xy, 2 * r);
        s.y = gld(xy, 2 * r + 1);
    }
    gst(spots, i, s);
}

// Batched preparation: one thread per spot of every beam, then one wave per
// layer of every beam.
__global__ __launch_bounds__(256) void k_spot_prep_batch(const BeamCtx *__restrict__ ctx,
                                                        const long long *__restrict__ off, int n_beams,
                                                        long long total) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int bm = find_owner(off, n_beams, t);
    const BeamCtx &c = load_ctx(ctx, bm);
    spot_prep_one(c.spots_in, c.sp, (int)(t - off[bm]), c.b.ivsadx, c.b.ivsady);
}

__global__ __launch_bounds__(64) void k_layer_prep_batch(const BeamCtx *__restrict__ ctx,
                                                        const long long *__restrict__ off, int n_beams) {
    const long long item = blockIdx.x;
    const int bm = find_owner(off, n_beams, item);
    const BeamCtx &c = load_ctx(ctx, bm);
    layer_prep_one(c, (int)(item - off[bm]), threadIdx.x);
}

// ---------------------------------------------------------------------------
// Forward.  One wave per 4x4x4 brick, one voxel per lane; a lane owns its
// voxel's sum (deterministic, no atomics).  The lanes first test 64 layers at
// a time (reaches the brick's WET range, and with culling: some spot of the
// layer within reach of the brick); the wave then walks the surviving layers.
// Per layer, each 64-spot chunk is tested lane-parallel against the brick
// (distance from the brick centre to the spot axis vs cutoff * max sigma +
// brick radius, once with the halo sigma and once with the narrower primary
// sigma) and the surviving spots are evaluated four at a time, their records
// coming in through scalar loads: spots within reach of the primary Gaussian
// with both terms, the rest of the halo's reach with the halo term alone (the
// primary term there is below exp(-cutoff^2/2) of its peak, as for any culled
// spot).
//
// COUNT: no dose; *n_dense += in-range voxels x layer spots (the reference's
// work), *n_exec += in-range voxels x spots evaluated here, a spot evaluated
// with the halo term alone counting one half.
//
// The exponents are evaluated on the in-plane offset scaled by s = sqrt(-a2)
// (a2 = -0.5 log2(e) / sigma_h^2, the halo exponent scale): with
// fx = 1 - hz / VSADX the offset from the spot axis is u = hx - x fx (as
// hx - x + hz tx, tx = x / VSADX), so s u = s hx - x (s fx) is one fma with
// per-(voxel, layer) factors, likewise s v, and the quadratic form of (s u, s v)
// is qs = -a2 d2 directly.  The halo term is exp2(-qs), the primary term
// exp2(-r12 qs) with r12 = a1 / a2 = sigma_h^2 / sigma_t^2 (0 without a
// primary Gaussian).
struct ScaledVox {
    float shx, shy, sfx, sfy, r12;
};

__device__ __forceinline__ ScaledVox scaled_vox(float hx, float hy, float hz, float ivx, float ivy, float s,
                                               float r12) {
    ScaledVox z;
    z.shx = s * hx;
    z.shy = s * hy;
    z.sfx = s * fmaf(-hz, ivx, 1.0f);
    z.sfy = s * fmaf(-hz, ivy, 1.0f);
    z.r12 = r12;
    return z;
}

// -a2 d2 for the spot q
__device__ __forceinline__ float scaled_d2(const SpotExt &q, const ScaledVox &z) {
    const float us = fmaf(-q.x, z.sfx, z.shx);
    const float vs = fmaf(-q.y, z.sfy, z.shy);
    const float w = fmaf(q.m12x2, vs, q.m11 * us);
    return fmaf(us, w, (q.m22 * vs) * vs);
}

__device__ __forceinline__ void spot_eval(const SpotExt &q, const ScaledVox &z, float &P, float &H) {
    const float qs = scaled_d2(q, z);
    P = fmaf(q.mu, __builtin_amdgcn_exp2f(-(z.r12 * qs)), P);
    H = fmaf(q.mu, __builtin_amdgcn_exp2f(-qs), H);
}

// Halo term only: the spot is beyond reach of the (narrower) primary Gaussian.
__device__ __forceinline__ void spot_eval_h(const SpotExt &q, const ScaledVox &z, float &H) {
    H = fmaf(q.mu, __builtin_amdgcn_exp2f(-scaled_d2(q, z)), H);
}

struct VoxelOf {
    bool valid;
    long long idx;
};

__device__ __forceinline__ VoxelOf brick_voxel(const Grid &g, const BrickGrid &bg, long long brick, int lane, int &bi,
                                               int &bj, int &bk, int &i, int &j, int &k) {
    brick_of(bg, brick, bi, bj, bk);
    i = bi * BRICK + (lane & 3);
    j = bj * BRICK + ((lane >> 2) & 3);
    k = bk * BRICK + (lane >> 4);
    VoxelOf r;
    r.valid = i < g.nx && j < g.ny && k < g.nz;
    r.idx = r.valid ? i + (long long)g.nx * (j + (long long)g.ny * k) : 0;
    return r;
}

template <bool CULL, bool COUNT>
__device__ __forceinline__ float forward_brick(const BeamCtx &c, long long brick, int lane, VoxelOf &vo,
                                               unsigned long long &n_dense, unsigned long long &n_exec) {
    int bi, bj, bk, i, j, k;
    vo = brick_voxel(c.g, c.bg, brick, lane, bi, bj, bk, i, j, k);
    const float wet = vo.valid ? gld(c.wet, vo.idx) * 10.0f : 3.0e38f;
    const LayerC *__restrict__ lc = c.lc;
    const SpotExt *__restrict__ sp = c.sp;
    float acc = 0.0f;
    unsigned long long n_exec2 = 0;  // COUNT: twice the evaluated pairs

    // Bricks no layer can use (every voxel deeper than 1.1 R80 of every layer:
    // outside the spots' reach in reach-limited WET, or beyond the range) leave
    // before the geometry: the dose there is 0.
    const float wlo = wave_min_dpp(wet);
    bool usable = false;
    for (int l0 = 0; l0 < c.n_layers && !usable; l0 += WAVE) {
        bool u = false;
        if (l0 + lane < c.n_layers) {
            const LayerC L = gld(lc, l0 + lane);
            u = L.ok && wlo <= L.lim;
        }
        usable = __ballot(u) != 0ull;
    }
    if (!usable) return acc;  // (COUNT: no voxel is in range of any layer, nothing to count)

    const float3 v = voxel_xyz(c.g, c.b, i, j, k);
    const float3 h = image_to_head(c.b, v);
    const float sx = c.b.src[0] - v.x, sy = c.b.src[1] - v.y, sz = c.b.src[2] - v.z;
    const float dsrc = sqrtf(sx * sx + sy * sy + sz * sz);
    const float ivx = c.b.ivsadx, ivy = c.b.ivsady;
    const BrickInfo B = brick_info_at(c.b, brick_centre(c.g, c.b, bi, bj, bk), wlo,
                                      wave_max_dpp(vo.valid ? wet : -3.0e38f));
    for (int l0 = 0; l0 < c.n_layers; l0 += WAVE) {
        bool lrel = false;
        float shb = 3.0e38f;  // CULL: upper bound of the halo sigma over the brick for layer l0 + lane
        if (l0 + lane < c.n_layers) {
            const LayerC L = gld(lc, l0 + lane);
            lrel = L.ok && B.wlo <= L.lim;
            if (CULL && lrel) lrel = !reach_skip(L, B, c.m, ivx, ivy, c.cut, c.rbrick, &shb);
        }
        unsigned long long lmask = __ballot(lrel);
        while (lmask) {
            const int lo = __builtin_ctzll(lmask);
            const int l = l0 + lo;
            lmask &= lmask - 1ull;
            const LayerC L = gld(lc, l);
            const bool act = wet <= L.lim;
            const unsigned long long amask = __ballot(act);
            if (!amask) continue;
            const SpotExt *lsp = gptr(sp, L.start);
            // CULL: spots of the first chunk within the bound radius of the brick centre (cut times the
            // halo-sigma bound of the brick plus the brick radius, at least the exact radius below); a
            // visit with none there, and none in later chunks, evaluates nothing and stops before the
            // voxel terms
            float d2c0 = 3.0e38f;
            if (CULL) {
                const float rb = fmaf(c.cut, __int_as_float(__builtin_amdgcn_readlane(__float_as_int(shb), lo)),
                                      c.rbrick);
                const float rb2 = rb * rb;
                bool any = false;
                for (int c0 = 0; c0 < L.n && !any; c0 += WAVE) {
                    float d2 = 3.0e38f;
                    if (lane < min(WAVE, L.n - c0)) d2 = axis_d2(B.hc.x, B.hc.y, B.hc.z, lsp[c0 + lane]);
                    if (c0 == 0) d2c0 = d2;
                    any = __ballot(d2 < rb2) != 0ull;
                }
                if (!any) continue;
            }
            float idd, sms;
            if (c.m.uniform) {
                // the layer's LUT rows from the brick's smallest WET on, in this wave's LDS slot
                __shared__ float2 s_lut[4][WAVE];
                float2 *tab = s_lut[(threadIdx.x >> 6) & 3];
                const int i0 = lut_stage(c.m, L.eid, B.wlo, lane, tab);
                lut_lookup_staged(c.m, L.eid, wet, tab, i0, idd, sms);
            } else {
                lut_lookup(c.m, L.eid, wet, idd, sms);
            }
            const VoxTerms T = vox_terms(c.m, L, wet, dsrc, idd, sms);
            float rcH2 = 0.0f, rcP2 = 0.0f;
            if (CULL) {
                // sigmas are non-negative (sh a square root, st 0 unless positive)
                const float rh = wave_max_nonneg(act ? halo_reach(T, c.cut) : 0.0f) + c.rbrick;
                const float rp = fmaf(c.cut, wave_max_nonneg(act ? T.st : 0.0f), c.rbrick);
                rcH2 = rh * rh;
                rcP2 = rp * rp;
            }
            float P = 0.0f, H = 0.0f;
            const ScaledVox z = scaled_vox(h.x, h.y, h.z, ivx, ivy, T.s2, T.r12);
            for (int c0 = 0; c0 < L.n; c0 += WAVE) {
                const int nin = min(WAVE, L.n - c0);
                unsigned long long mask, hmask;
                if (CULL) {
                    bool relH = false, relP = false;
                    if (lane < nin) {
                        const float d2c = c0 == 0 ? d2c0 : axis_d2(B.hc.x, B.hc.y, B.hc.z, lsp[c0 + lane]);
                        relH = d2c < rcH2;
                        relP = d2c < rcP2;  // both terms; relH alone: the halo term only
                    }
                    mask = __ballot(relP);
                    hmask = __ballot(relH) & ~mask;
                } else {
                    mask = nin == WAVE ? ~0ull : ((1ull << nin) - 1ull);
                    hmask = 0ull;
                }
                if (COUNT) {
                    n_exec2 += (unsigned long long)__popcll(amask) * (2u * (unsigned)__popcll(mask) +
                                                                      (unsigned)__popcll(hmask));
                    continue;
                }
                const SpotExt *csp = lsp + c0;
                // trip counts from scalar popcounts (a 64-bit compare would go to the VALU)
                for (int nq = __builtin_amdgcn_readfirstlane(__popcll(mask)) >> 2; nq > 0; --nq) {
                    const int s0 = __builtin_ctzll(mask);
                    mask &= mask - 1ull;
                    const int s1 = __builtin_ctzll(mask);
                    mask &= mask - 1ull;
                    const int s2 = __builtin_ctzll(mask);
                    mask &= mask - 1ull;
                    const int s3 = __builtin_ctzll(mask);
                    mask &= mask - 1ull;
                    const SpotExt q0 = csp[s0], q1 = csp[s1], q2 = csp[s2], q3 = csp[s3];
                    spot_eval(q0, z, P, H);
                    spot_eval(q1, z, P, H);
                    spot_eval(q2, z, P, H);
                    spot_eval(q3, z, P, H);
                }
                while (mask) {
                    const int s0 = __builtin_ctzll(mask);
                    mask &= mask - 1ull;
                    spot_eval(csp[s0], z, P, H);
                }
                for (int nq = __builtin_amdgcn_readfirstlane(__popcll(hmask)) >> 2; nq > 0; --nq) {
                    const int s0 = __builtin_ctzll(hmask);
                    hmask &= hmask - 1ull;
                    const int s1 = __builtin_ctzll(hmask);
                    hmask &= hmask - 1ull;
                    const int s2 = __builtin_ctzll(hmask);
                    hmask &= hmask - 1ull;
                    const int s3 = __builtin_ctzll(hmask);
                    hmask &= hmask - 1ull;
                    const SpotExt q0 = csp[s0], q1 = csp[s1], q2 = csp[s2], q3 = csp[s3];
                    spot_eval_h(q0, z, H);
                    spot_eval_h(q1, z, H);
                    spot_eval_h(q2, z, H);
                    spot_eval_h(q3, z, H);
                }
                while (hmask) {
                    const int s0 = __builtin_ctzll(hmask);
                    hmask &= hmask - 1ull;
                    spot_eval_h(csp[s0], z, H);
                }
            }
            if (act) acc = fmaf(T.Af, P, fmaf(T.Bf, H, acc));
        }
    }
    if (COUNT) {
        n_exec += n_exec2 / 2;
        for (int l = 0; l < c.n_layers; ++l) {
            const LayerC L = gld(lc, l);
            if (L.ok) n_dense += (unsigned long long)__popcll(__ballot(wet <= L.lim)) * (unsigned)L.n;
        }
    }
    return acc;
}

template <bool CULL, bool COUNT = false>
__global__ __launch_bounds__(256) void k_pb_forward(BeamCtx c, int accumulate,
                                                   unsigned long long *__restrict__ counts = nullptr) {
    // plain dispatch order: consecutive bricks go to different XCDs, which
    // spreads the heavy bricks of the beam path over the whole chip
    const long long brick = (long long)blockIdx.x * 4 + wave_in_block();
    if (brick >= c.bg.count()) return;  // whole wave leaves together
    const int lane = threadIdx.x & 63;
    VoxelOf vo;
    unsigned long long n_dense = 0, n_exec = 0;
    const float acc = forward_brick<CULL, COUNT>(c, brick, lane, vo, n_dense, n_exec);
    if (COUNT) {
        if (lane == 0) {
            atomicAdd(&counts[0], n_dense);
            atomicAdd(&counts[1], n_exec);
        }
        return;
    }
    if (vo.valid) gst(c.dose, vo.idx, accumulate ? gld(c.dose, vo.idx) + acc : acc);
}

// Batched forward: one wave per brick of every beam; beam b owns items
// [off[b], off[b + 1]) and writes its own dose volume (the plan's dose when the
// plan has one beam, else a per-beam volume that k_plan_sum adds up).  A wave
// takes item base + its wave index, then steps by `stride`: stride >= total
// gives one brick per wave (hardware dispatch), stride = 4 * gridDim.x a
// persistent grid that strides over all bricks.  Either way each brick is
// computed by exactly one wave, so the results are the same.
// The persistent form is its own instantiation: the loop around the brick
// body costs registers (80 VGPRs instead of 62), which the one-shot form
// should not pay.
__device__ __forceinline__ long long first_item(long long base) {
    return base + (long long)blockIdx.x * 4 + wave_in_block();
}

template <bool CULL>
__device__ __forceinline__ void forward_item(const BeamCtx *__restrict__ ctx, const long long *__restrict__ off,
                                             int n_beams, long long item) {
    const int bm = find_owner(off, n_beams, item);
    const BeamCtx &c = load_ctx(ctx, bm);
    VoxelOf vo;
    unsigned long long nd = 0, ne = 0;
    const float acc = forward_brick<CULL, false>(c, item - off[bm], threadIdx.x & 63, vo, nd, ne);
    if (vo.valid) gst(c.dose, vo.idx, acc);
}

template <bool CULL, bool PERSIST>
__global__ __launch_bounds__(256) void k_pb_forward_batch(const BeamCtx *__restrict__ ctx,
                                                         const long long *__restrict__ off, int n_beams,
                                                         long long base, long long total, long long stride) {
    if (PERSIST) {
        for (long long item = first_item(base); item < total; item += stride) forward_item<CULL>(ctx, off, n_beams, item);
    } else {
        const long long item = first_item(base);
        if (item < total) forward_item<CULL>(ctx, off, n_beams, item);
    }
}

// Dose of a plan whose beams write volumes of their own: the sum of the beam
// volumes, added in beam order (the order of single-beam calls with
// accumulate = 1); zero for a plan without beams.  Plan p owns blocks
// [off[p], off[p + 1]) of PLAN_SUM_VOXELS voxels; a thread takes four voxels
// 256 apart, so each beam's four loads are in flight together.
__global__ __launch_bounds__(256) void k_plan_sum(const PlanCtx *__restrict__ plans, const BeamCtx *__restrict__ ctx,
                                                 const long long *__restrict__ off, int n_plans) {
    const long long blk = blockIdx.x;
    const int p = find_owner(off, n_plans, blk);
    const PlanCtx P = plans[p];
    const long long t0 = (blk - off[p]) * PLAN_SUM_VOXELS + threadIdx.x;
    float d[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // 0 + D_0 == D_0: the same sums as beam-by-beam accumulation
    for (int b = P.beam0; b < P.beam0 + P.n_beams; ++b) {
        const float *bd = ctx[b].dose;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const long long t = t0 + 256 * q;
            if (t < P.nvox) d[q] += gld(bd, t);
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const long long t = t0 + 256 * q;
        if (t < P.nvox) gst(P.dose, t, d[q]);
    }
}

// ---------------------------------------------------------------------------
// One wave per brick: WET range (mm) of the voxels with dL/dD != 0; an empty
// brick gets (+big, -big).
__device__ __forceinline__ void brick_summary_one(const BeamCtx &c, long long brick, int lane) {
    int bi, bj, bk, i, j, k;
    const VoxelOf vo = brick_voxel(c.g, c.bg, brick, lane, bi, bj, bk, i, j, k);
    float lo = 3.0e38f, hi = -3.0e38f;
    if (vo.valid && gld(c.gdose, vo.idx) != 0.0f) lo = hi = gld(c.wet, vo.idx) * 10.0f;
    lo = wave_min(lo);
    hi = wave_max(hi);
    if (lane == 0) gst(c.bsum, brick, make_float2(lo, hi));
}

__global__ __launch_bounds__(256) void k_brick_summary(BeamCtx c) {
    const long long brick = (long long)blockIdx.x * 4 + wave_in_block();
    if (brick >= c.bg.count()) return;
    brick_summary_one(c, brick, threadIdx.x & 63);
}

// Batched: one workgroup per row of bricks along x (4 voxel rows in y x 4
// slices in z, the whole x extent) of a beam with spots.  Thread t reads voxel
// x0 + t of the 16 voxel rows, so every load is one contiguous run of the
// row; then the four lanes of each brick fold their ranges.  min / max are
// exact, so the result equals the one-wave-per-brick summary's.
__global__ __launch_bounds__(256) void k_brick_summary_batch(const BeamCtx *__restrict__ ctx,
                                                            const long long *__restrict__ off, int n_beams,
                                                            long long base) {
    const long long row = base + blockIdx.x;
    const int bm = find_owner(off, n_beams, row);
    const BeamCtx &c = load_ctx(ctx, bm);
    const unsigned r = (unsigned)(row - off[bm]), by = (unsigned)c.bg.by;
    const int bj = (int)(r % by), bk = (int)(r / by);
    const int nx = c.g.nx, ny = c.g.ny, nz = c.g.nz;
    for (int x0 = 0; x0 < BRICK * c.bg.bx; x0 += 256) {
        const int x = x0 + (int)threadIdx.x;
        float gv[16], wv[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int j = BRICK * bj + (q & 3), k = BRICK * bk + (q >> 2);
            const bool ok = x < nx && j < ny && k < nz;
            // voxel indices fit in 32 bits; an unused lane loads voxel 0
            const unsigned idx = ok ? (unsigned)x + (unsigned)nx * ((unsigned)j + (unsigned)ny * (unsigned)k) : 0u;
            const float g = gld(c.gdose, idx), w = gld(c.wet, idx);
            gv[q] = ok ? g : 0.0f;
            wv[q] = w;
        }
        float lo = 3.0e38f, hi = -3.0e38f;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            if (gv[q] != 0.0f) {
                const float v = wv[q] * 10.0f;
                lo = fminf(lo, v);
                hi = fmaxf(hi, v);
            }
        }
        // x0 is a multiple of 256: lanes 4m..4m+3 hold the four columns of one brick
        lo = fminf(lo, __shfl_xor(lo, 1));
        hi = fmaxf(hi, __shfl_xor(hi, 1));
        lo = fminf(lo, __shfl_xor(lo, 2));
        hi = fmaxf(hi, __shfl_xor(hi, 2));
        const int bi = x >> 2;
        if ((threadIdx.x & 3) == 0 && bi < c.bg.bx)
            gst(c.bsum, bi + (long long)c.bg.bx * (bj + (long long)c.bg.by * bk), make_float2(lo, hi));
    }
}

// ---------------------------------------------------------------------------
// Reach-limited WET.  With culling, a brick that every layer skips for the
// widest WET range possible ([-1 mm, 1.1 R80]) gets no dose and no gradient
// term whatever its WET, so its WET need not be computed.  reach_skip is
// monotone in the WET range, so every brick the dose kernels evaluate (and
// every brick whose voxels the adjoint visits) passes this test.
//   pass 1 (k_wet_need_batch): bit 0 of need[brick] = some layer reaches it
//           (smoothed WET needed);
//   pass 2 (k_wet_dilate_batch): bit 1 = some brick that smoothing reads raw
//           WET for, i.e. b - o for an offset o of the beam's touch set, has
//           bit 0 (raw WET needed).
// Every WET value the dose kernels then read equals the full computation's.
// Both passes run one lane per brick: a wave takes 64 consecutive items of the
// batch's brick space and walks the beams they belong to (normally one), so
// one context load serves 64 bricks.  `mine`: the lane's item belongs to the
// beam in hand; the other lanes count as decided and store nothing.
__device__ __forceinline__ void wet_need_lanes(const BeamCtx &c, long long brick, bool mine) {
    const bool all = !(c.cut > 0.0f);
    bool need = mine && all, done = !mine || all;
    if (__ballot(!done)) {
        int bi, bj, bk;
        brick_of(c.bg, mine ? brick : 0, bi, bj, bk);
        const float3 vc = brick_centre(c.g, c.b, bi, bj, bk);
        const float ivx = c.b.ivsadx, ivy = c.b.ivsady;
        for (int l = 0; l < c.n_layers; ++l) {
            if (!__ballot(!done)) break;
            const LayerC L = cld(c.lc, l);
            if (!done && L.ok && !reach_skip(L, brick_info_at(c.b, vc, -1.0f, L.lim), c.m, ivx, ivy, c.cut, c.rbrick))
                need = done = true;
        }
    }
    if (mine) gst(c.need, brick, (unsigned char)(need ? 1u : 0u));
}

// Offsets of the touch set go eight at a time (their loads in flight
// together), with a vote after each group; only bit 1 of this brick's own
// byte changes, and other waves read bit 0 only.
__device__ __forceinline__ void wet_dilate_lanes(const BeamCtx &c, long long brick, bool mine) {
    int bi = 0, bj = 0, bk = 0;
    unsigned own = 0;
    if (mine) {
        brick_of(c.bg, brick, bi, bj, bk);
        own = gld(c.need, brick);
    }
    bool raw = !mine || (own & 1u);
    const int bx = c.bg.bx, by = c.bg.by, bz = c.bg.bz;
    for (int t0 = 0; t0 < c.n_touch; t0 += 8) {
        if (!__ballot(!raw)) break;
        const bool open = !raw;
        // the table holds MAX_TOUCH (a multiple of 8) entries per beam: the
        // eight keys load together, entries past n_touch are not used
        int key[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) key[u] = cld(c.touch, t0 + u);
        unsigned hit = 0;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int si = bi - ((key[u] & 255) - 128), sj = bj - (((key[u] >> 8) & 255) - 128),
                      sk = bk - (((key[u] >> 16) & 255) - 128);
            const bool in = open && t0 + u < c.n_touch && (unsigned)si < (unsigned)bx && (unsigned)sj < (unsigned)by &&
                            (unsigned)sk < (unsigned)bz;
            // brick indices fit in 32 bits (the grid check keeps nx*ny*(nz+1)
            // below 2^28); every lane loads (brick 0 where the offset is not
            // used), so the eight loads need no branches and fly together
            const unsigned at = in ? (unsigned)si + (unsigned)bx * ((unsigned)sj + (unsigned)by * (unsigned)sk) : 0u;
            const unsigned v = gld(c.need, at);
            hit |= in ? v : 0u;
        }
        raw = raw || (hit & 1u);
    }
    if (mine && raw) gst(c.need, brick, (unsigned char)(own | 2u));
}

template <bool DILATE>
__device__ __forceinline__ void need_wave(const BeamCtx *__restrict__ ctx, const long long *__restrict__ off,
                                          int n_beams, long long base, long long total) {
    const long long w0 = base + ((long long)blockIdx.x * 4 + wave_in_block()) * WAVE;
    if (w0 >= total) return;  // whole wave leaves together
    const long long item = w0 + (threadIdx.x & 63);
    const long long last = w0 + WAVE < total ? w0 + WAVE : total;
    for (long long first = w0; first < last;) {
        const int bm = find_owner(off, n_beams, first);
        const long long end = bm + 1 < n_beams && off[bm + 1] < last ? off[bm + 1] : last;
        const BeamCtx &c = load_ctx(ctx, bm);
        const bool mine = item >= first && item < end;
        if (DILATE) wet_dilate_lanes(c, item - off[bm], mine);
        else wet_need_lanes(c, item - off[bm], mine);
        first = end;
    }
}

__global__ __launch_bounds__(256) void k_wet_need_batch(const BeamCtx *__restrict__ ctx,
                                                       const long long *__restrict__ off, int n_beams,
                                                       long long base, long long total) {
    need_wave<false>(ctx, off, n_beams, base, total);
}

__global__ __launch_bounds__(256) void k_wet_dilate_batch(const BeamCtx *__restrict__ ctx,
                                                         const long long *__restrict__ off, int n_beams,
                                                         long long base, long long total) {
    need_wave<true>(ctx, off, n_beams, base, total);
}

// ---------------------------------------------------------------------------
// Adjoint.  Workgroup = one wave = one CS-spot chunk of one layer x one slab
// of bricks.  Lane = (voxel subset, spot); a spot's three sums stay in
// registers across the whole slab.  The lanes test 64 bricks at a time
// (holds voxels with dL/dD != 0 within the layer's range, and with culling:
// within reach of the chunk); per surviving brick the wave acts as lanes =
// voxels to compute the voxel terms, compacts the voxels that matter into LDS,
// then every spot lane walks its subset of them with broadcast LDS reads.
//   dL/dmu_s = sum_v g_v [A e1 + B e2]
//   dL/dx_s  = -2 mu_s sum_v g_v [A a1 e1 + B a2 e2] (1 + q/VSADX) p_x
//   dL/dy_s  = -2 mu_s sum_v g_v [A a1 e1 + B a2 e2] (1 + q/VSADY) p_y
// with r = (hx - x_s, hy - y_s, hz) the voxel relative to the spot,
// q = (t . r) / |t|^2, p = r - q t (its part perpendicular to the axis) and
// a1, a2 = -1/(2 sigma^2) (natural-log scale).
//
// COUNT: no gradient; *n_dense += voxels (dL/dD != 0, in range) x chunk spots
// (the reference's work), *n_exec += those voxels x spots evaluated here.
struct LdsVox {
    float hx, hy, hz, a1;
    float a2, GA, GB, GAa;
    float GBb, pad0, pad1, pad2;
};

template <bool CULL, int CS, bool COUNT>
__device__ __forceinline__ void adjoint_item(const BeamCtx &c, int cid, int slab, int lane, LdsVox *s_vox,
                                             unsigned long long &n_dense, unsigned long long &n_exec) {
    constexpr int G = WAVE / CS;  // voxel subsets per wave: lane = group * CS + slot
    const int slot = lane % CS, group = lane / CS;
    const LayerC *__restrict__ lc = c.lc;
    // chunk -> (layer, first spot of the chunk within the layer)
    int l = 0, c0 = -1;
    for (; l < c.n_layers; ++l) {
        const int nch = (gld(lc, l).n + CS - 1) / CS;
        if (cid < nch) {
            c0 = cid * CS;
            break;
        }
        cid -= nch;
    }
    if (c0 < 0) return;  // surplus item of the upper-bound grid
    const LayerC L = gld(lc, l);
    if (!L.ok) return;
    const bool sval = c0 + slot < L.n;
    const int spot = L.start + c0 + slot;
    SpotExt q;
    if (sval) q = gld(c.sp, spot);
    else q = SpotExt{0.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 1.f};
    const float itt = 1.0f / (q.tx * q.tx + q.ty * q.ty + 1.0f);
    const float ttx = q.tx * itt, tty = q.ty * itt;  // t_xy / |t|^2
    const float ivx = c.b.ivsadx, ivy = c.b.ivsady;
    const unsigned long long chunk_n = (unsigned long long)min(CS, L.n - c0);
    // reach of this chunk's spots only
    const LayerC R = CULL ? chunk_reach(c.sp, L, c0, (int)chunk_n, lane) : L;

    const long long nb = c.bg.count();
    const long long b0 = nb * slab / c.nslab, b1 = nb * (slab + 1) / c.nslab;
    // dL/dx = FX0 + FX1 / VSADX with FX0 = sum F p_x, FX1 = sum F q_r p_x (same for y)
    float K = 0.0f, FX0 = 0.0f, FX1 = 0.0f, FY0 = 0.0f, FY1 = 0.0f;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;

    for (long long base = b0; base < b1; base += WAVE) {
        bool brel = false;
        if (base + lane < b1) {
            const float2 ws = gld(c.bsum, base + lane);
            brel = ws.x <= ws.y && ws.x <= L.lim;
            if (CULL && brel && !COUNT) {
                int ti, tj, tk;
                brick_of(c.bg, base + lane, ti, tj, tk);
                brel = !reach_skip(R, brick_info_at(c.b, brick_centre(c.g, c.b, ti, tj, tk), ws.x, ws.y), c.m, ivx, ivy,
                                   c.cut, c.rbrick);
            }
        }
        unsigned long long fm = __ballot(brel);
        while (fm) {
            const long long brick = base + __builtin_ctzll(fm);
            fm &= fm - 1ull;
            int bi, bj, bk, i, j, k;
            const VoxelOf vo = brick_voxel(c.g, c.bg, brick, lane, bi, bj, bk, i, j, k);
            const float gv = vo.valid ? gld(c.gdose, vo.idx) : 0.0f;
            const float wet = vo.valid ? gld(c.wet, vo.idx) * 10.0f : 3.0e38f;
            const bool act = vo.valid && gv != 0.0f && wet <= L.lim;
            const unsigned long long am = __ballot(act);
            if (!am) continue;
            if (COUNT) n_dense += (unsigned long long)__popcll(am) * chunk_n;

            const float3 v = voxel_xyz(c.g, c.b, i, j, k);
            const float3 h = image_to_head(c.b, v);
            const float sx = c.b.src[0] - v.x, sy = c.b.src[1] - v.y, sz = c.b.src[2] - v.z;
            const float dsrc = sqrtf(sx * sx + sy * sy + sz * sz);
            const VoxTerms T = vox_terms(c.m, L, wet, dsrc);

            bool rel = sval;
            if (CULL) {
                const float rc = wave_max_nonneg(act ? fmaxf(c.cut * T.st, halo_reach(T, c.cut)) : 0.0f) + c.rbrick;
                const float3 hc = image_to_head(c.b, brick_centre(c.g, c.b, bi, bj, bk));
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
            // pad the list to whole rounds of G voxels with zero voxels: every
            // term of a zero voxel is an exact +0 (GA = GB = 0, e = exp2(0)),
            // so the sums are unchanged and all lanes run the same trip count
            const int na = __popcll(am), nit = (na + G - 1) / G;
            if (G > 1 && lane >= na && lane < nit * G) s_vox[lane] = LdsVox{};
            __syncthreads();
            if (rel) {
                for (int it = 0; it < nit; ++it) {
                    const LdsVox e = s_vox[it * G + group];
                    // r = (a, b, hz) the voxel relative to the spot, (u, w) its
                    // in-plane offset e_xy = (a, b) + hz t_xy; q = (t.r)/|t|^2
                    // = (e_xy . t_xy)/|t|^2 - hz and p = (a, b) - q t_xy, the
                    // part of e_xy perpendicular to the axis; d2 = e_xy . p
                    // (no cancellation)
                    const float a = e.hx - q.x, b = e.hy - q.y;
                    const float u = fmaf(e.hz, q.tx, a), w = fmaf(e.hz, q.ty, b);
                    const float qr = fmaf(u, ttx, fmaf(w, tty, -e.hz));
                    const float px = fmaf(-qr, q.tx, a), py = fmaf(-qr, q.ty, b);
                    const float d2 = fmaf(u, px, w * py);
                    const float e1 = __builtin_amdgcn_exp2f(e.a1 * d2);
                    const float e2 = __builtin_amdgcn_exp2f(e.a2 * d2);
                    K = fmaf(e.GA, e1, K);
                    K = fmaf(e.GB, e2, K);
                    const float F = fmaf(e.GAa, e1, e.GBb * e2);
                    const float Fq = F * qr;
                    FX0 = fmaf(F, px, FX0);
                    FX1 = fmaf(Fq, px, FX1);
                    FY0 = fmaf(F, py, FY0);
                    FY1 = fmaf(Fq, py, FY1);
                }
            }
            __syncthreads();
        }
    }
    if (COUNT) return;
    // fold the voxel subsets (fixed order: deterministic)
    float FX = fmaf(FX1, ivx, FX0), FY = fmaf(FY1, ivy, FY0);
#pragma unroll
    for (int off = CS; off < WAVE; off <<= 1) {
        K += __shfl_xor(K, off);
        FX += __shfl_xor(FX, off);
        FY += __shfl_xor(FY, off);
    }
    if (sval && group == 0) gst(c.partial, (long long)slab * c.n_spots + spot, make_float4(K, FX, FY, 0.0f));
}

// Packed adjoint (CS = PACK_CS): a chunk holds up to 128 spots of one layer
// (a layer of n spots splits into ceil(n / 128) near-equal chunks), two per
// lane.  Per brick, the chunk's spots within reach are listed and run in
// passes of np spots: lane = group * np + slot, G = min(64 / np, 16) voxel
// subsets per pass, so the lanes carry pairs of reachable spots only.  The
// pass count minimises the loop trips (plus a fixed cost per pass).  Each
// pass folds its G subsets through LDS into the spot's owner lane in a fixed
// order (deterministic).  With culling, spots beyond the primary radius of the
// brick come last in the list, and a pass holding only such spots evaluates
// the halo term alone (the forward drops the same primary terms).
constexpr int PACK_CS = 128;
constexpr int PACK_GMAX = 16;
constexpr int PACK_VOX = WAVE + PACK_GMAX;  // voxel list incl. the zero tail a pass may read

struct PackLds {
    LdsVox vox[PACK_VOX];
    float4 part[WAVE];
    int list[PACK_CS];
};

// floor(a / b) for 0 <= a < 256, 1 <= b <= 128: a + 0.5 keeps the quotient
// at least 0.5 / b from an integer (2^-16 relative), far beyond the 2^-22
// relative error of the hardware reciprocal and the product.
__device__ __forceinline__ int small_div(int a, int b) {
    return (int)(((float)a + 0.5f) * __builtin_amdgcn_rcpf((float)b));
}

// Squared distance from (hx, hy, hz) to the axis of the spot at (x, y) with
// slopes (tx, ty) and itt = 1 / |t|^2: the same terms as the adjoint loop.
__device__ __forceinline__ float axis_d2_t(float hx, float hy, float hz, float x, float y, float tx, float ty,
                                           float itt) {
    const float a = hx - x, b = hy - y;
    const float u = fmaf(hz, tx, a), w = fmaf(hz, ty, b);
    const float qr = fmaf(u, tx * itt, fmaf(w, ty * itt, -hz));
    return fmaf(u, fmaf(-qr, tx, a), w * fmaf(-qr, ty, b));
}

// Pass plans of the packed adjoint, fixed at compile time (the wave-uniform
// counts then index a constant table by scalar loads instead of taking
// reciprocals on the VALU per visit and pass):
//   groups[np]            G = min(64 / np, 16) voxel subsets for np spots per pass;
//   trips[na][np]         ceil(na / G) loop trips over na voxels (indexed by np, so
//                         that both lookups of a pass issue together);
//   passes[nrel-1][na-1]  the pass count P in [ceil(nrel / 64), +3] (and <= nrel) with the
//                         fewest P (ceil(na / G(ceil(nrel / P))) + PACK_PASS_COST), the first on a tie;
//   spots[nrel-1][na-1]   ceil(nrel / P), the spots of a full pass.
// PACK_PASS_COST is the work outside the loop per pass (spot exchange,
// subset fold) in loop trips.
constexpr int PACK_PASS_COST = 3;

struct PackPlanTables {
    unsigned char groups[WAVE + 1];
    unsigned char trips[WAVE + 1][WAVE + 1];
    unsigned char passes[PACK_CS][WAVE];
    unsigned char spots[PACK_CS][WAVE];
};

constexpr int pack_groups_of(int np) { return WAVE / np < PACK_GMAX ? WAVE / np : PACK_GMAX; }

constexpr int pack_cost_of(int nrel, int na, int P) {
    return P * ((na + pack_groups_of((nrel + P - 1) / P) - 1) / pack_groups_of((nrel + P - 1) / P) + PACK_PASS_COST);
}

constexpr PackPlanTables make_pack_plan_tables() {
    PackPlanTables t{};
    for (int np = 1; np <= WAVE; ++np) t.groups[np] = (unsigned char)pack_groups_of(np);
    for (int na = 0; na <= WAVE; ++na)
        for (int np = 1; np <= WAVE; ++np)
            t.trips[na][np] = (unsigned char)((na + pack_groups_of(np) - 1) / pack_groups_of(np));
    for (int nrel = 1; nrel <= PACK_CS; ++nrel) {
        for (int na = 1; na <= WAVE; ++na) {
            const int p_min = (nrel + WAVE - 1) / WAVE;
            int P = p_min, best = pack_cost_of(nrel, na, P);
            for (int Q = p_min + 1; Q <= p_min + 3 && Q <= nrel; ++Q) {
                const int cq = pack_cost_of(nrel, na, Q);
                if (cq < best) {
                    best = cq;
                    P = Q;
                }
            }
            t.passes[nrel - 1][na - 1] = (unsigned char)P;
            t.spots[nrel - 1][na - 1] = (unsigned char)((nrel + P - 1) / P);
        }
    }
    return t;
}

__constant__ PackPlanTables kPackPlan = make_pack_plan_tables();

template <bool CULL, bool COUNT>
__device__ __forceinline__ void adjoint_item_packed(const BeamCtx &c, int cid, int slab, int lane, PackLds &S,
                                                    unsigned long long &n_dense, unsigned long long &n_exec) {
    const LayerC *__restrict__ lc = c.lc;
    int l = 0, c0 = -1, cnt = 0;
    for (; l < c.n_layers; ++l) {
        const int n = gld(lc, l).n;
        const int nch = (n + PACK_CS - 1) / PACK_CS;
        if (cid < nch) {
            c0 = (int)((long long)cid * n / nch);
            cnt = (int)((long long)(cid + 1) * n / nch) - c0;
            break;
        }
        cid -= nch;
    }
    if (c0 < 0) return;  // surplus item of the upper-bound grid
    const LayerC L = gld(lc, l);
    if (!L.ok) return;
    // lane owns chunk spots lane and lane + 64
    const bool v0 = lane < cnt, v1 = lane + WAVE < cnt;
    float x0 = 0.0f, y0 = 0.0f, tx0 = 0.0f, ty0 = 0.0f, x1 = 0.0f, y1 = 0.0f, tx1 = 0.0f, ty1 = 0.0f;
    if (v0) {
        const SpotExt q = gld(c.sp, L.start + c0 + lane);
        x0 = q.x;
        y0 = q.y;
        tx0 = q.tx;
        ty0 = q.ty;
    }
    if (v1) {
        const SpotExt q = gld(c.sp, L.start + c0 + lane + WAVE);
        x1 = q.x;
        y1 = q.y;
        tx1 = q.tx;
        ty1 = q.ty;
    }
    const float it0 = 1.0f / (tx0 * tx0 + ty0 * ty0 + 1.0f);
    const float it1 = 1.0f / (tx1 * tx1 + ty1 * ty1 + 1.0f);
    const float ivx = c.b.ivsadx, ivy = c.b.ivsady;
    const LayerC R = CULL ? chunk_reach(c.sp, L, c0, cnt, lane) : L;
    if (lane < PACK_VOX - WAVE) S.vox[WAVE + lane] = LdsVox{};  // permanent zero tail

    const long long nb = c.bg.count();
    const long long b0 = nb * slab / c.nslab, b1 = nb * (slab + 1) / c.nslab;
    float K0 = 0.0f, FX0 = 0.0f, FY0 = 0.0f, K1 = 0.0f, FX1 = 0.0f, FY1 = 0.0f;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;

    for (long long base = b0; base < b1; base += WAVE) {
        bool brel = false;
        float shb = 3.0e38f;  // CULL: upper bound of the halo sigma over brick base + lane
        if (base + lane < b1) {
            const float2 ws = gld(c.bsum, base + lane);
            brel = ws.x <= ws.y && ws.x <= L.lim;
            if (CULL && brel && !COUNT) {
                int ti, tj, tk;
                brick_of(c.bg, base + lane, ti, tj, tk);
                brel = !reach_skip(R, brick_info_at(c.b, brick_centre(c.g, c.b, ti, tj, tk), ws.x, ws.y), c.m, ivx, ivy,
                                   c.cut, c.rbrick, &shb);
            }
        }
        unsigned long long fm = __ballot(brel);
        while (fm) {
            const int bl = __builtin_ctzll(fm);
            const long long brick = base + bl;
            fm &= fm - 1ull;
            int bi, bj, bk, i, j, k;
            const VoxelOf vo = brick_voxel(c.g, c.bg, brick, lane, bi, bj, bk, i, j, k);
            const float gv = vo.valid ? gld(c.gdose, vo.idx) : 0.0f;
            const float wet = vo.valid ? gld(c.wet, vo.idx) * 10.0f : 3.0e38f;
            const bool act = vo.valid && gv != 0.0f && wet <= L.lim;
            const unsigned long long am = __ballot(act);
            if (!am) continue;
            if (COUNT) n_dense += (unsigned long long)__popcll(am) * (unsigned)cnt;

            // CULL: the chunk's spots within the bound radius of the brick centre (cut times the brick's
            // halo-sigma bound plus the brick radius, at least the exact radius below); none: nothing to
            // evaluate, leave before the voxel terms
            float d2s0 = 3.0e38f, d2s1 = 3.0e38f;
            if (CULL) {
                const float3 hc = image_to_head(c.b, brick_centre(c.g, c.b, bi, bj, bk));
                if (v0) d2s0 = axis_d2_t(hc.x, hc.y, hc.z, x0, y0, tx0, ty0, it0);
                if (v1) d2s1 = axis_d2_t(hc.x, hc.y, hc.z, x1, y1, tx1, ty1, it1);
                const float rb = fmaf(c.cut, __int_as_float(__builtin_amdgcn_readlane(__float_as_int(shb), bl)),
                                      c.rbrick);
                if (!__ballot(d2s0 < rb * rb || d2s1 < rb * rb)) continue;
            }
            const float3 v = voxel_xyz(c.g, c.b, i, j, k);
            const float3 h = image_to_head(c.b, v);
            const float sx = c.b.src[0] - v.x, sy = c.b.src[1] - v.y, sz = c.b.src[2] - v.z;
            const float dsrc = sqrtf(sx * sx + sy * sy + sz * sz);
            const VoxTerms T = vox_terms(c.m, L, wet, dsrc);

            // r: within reach of a term; h: of the halo term only (beyond the primary radius)
            bool r0 = v0, r1 = v1, h0 = false, h1 = false;
            if (CULL) {
                const float rc = wave_max_nonneg(act ? fmaxf(c.cut * T.st, halo_reach(T, c.cut)) : 0.0f) + c.rbrick;
                const float rp = fmaf(c.cut, wave_max_nonneg(act ? T.st : 0.0f), c.rbrick);
                r0 = v0 && d2s0 < rc * rc;
                r1 = v1 && d2s1 < rc * rc;
                h0 = r0 && !(d2s0 < rp * rp);
                h1 = r1 && !(d2s1 < rp * rp);
            }
            const unsigned long long mp0 = __ballot(r0 && !h0), mp1 = __ballot(r1 && !h1);
            const unsigned long long mh0 = __ballot(h0), mh1 = __ballot(h1);
            const int np0 = __popcll(mp0), nP = np0 + __popcll(mp1);
            const int nh0 = __popcll(mh0), nrel = nP + nh0 + __popcll(mh1);
            if (nrel == 0) continue;
            if (COUNT) {
                n_exec += (unsigned long long)__popcll(am) * (unsigned)nrel;
                continue;
            }
            // list positions: spots with both terms (first set, second set), then
            // halo-only spots (first set, second set); passes past nP take the halo term alone
            const int p0 = h0 ? nP + __popcll(mh0 & lt_mask) : __popcll(mp0 & lt_mask);
            const int p1 = h1 ? nP + nh0 + __popcll(mh1 & lt_mask) : np0 + __popcll(mp1 & lt_mask);
            if (r0) S.list[p0] = lane;
            if (r1) S.list[p1] = lane + WAVE;
            const int na = __popcll(am);
            if (act) {
                LdsVox e;
                e.hx = h.x;
                e.hy = h.y;
                e.hz = h.z;
                e.a1 = T.a1;
                e.a2 = T.a2;
                e.GA = gv * T.Af;
                e.GB = gv * T.Bf;
                e.GAa = e.GA * (T.a1 * (1.0f / kLog2e));
                e.GBb = e.GB * (T.a2 * (1.0f / kLog2e));
                e.pad0 = e.pad1 = e.pad2 = 0.0f;
                S.vox[__popcll(am & lt_mask)] = e;
            }
            // zero voxels past na: every term of a zero voxel is an exact +0
            if (lane >= na) S.vox[lane] = LdsVox{};
            // passes: the plan with the fewest loop trips (1 <= nrel <= 128, 1 <= na <= 64)
            const int np_full = kPackPlan.spots[nrel - 1][na - 1];
            __syncthreads();
            for (int s0 = 0; s0 < nrel; s0 += np_full) {
                const int np = min(np_full, nrel - s0);
                const int G = kPackPlan.groups[np];
                const int group = small_div(lane, np), slot = lane - group * np;
                const bool on = group < G;
                const int js = on ? S.list[s0 + slot] : 0;
                const int src = js & (WAVE - 1);
                const bool hi = js >= WAVE;
                // the pass's spot from its owner lane (every lane takes part)
                const float ax = __shfl(x0, src), bx = __shfl(x1, src);
                const float ay = __shfl(y0, src), by = __shfl(y1, src);
                const float atx = __shfl(tx0, src), btx = __shfl(tx1, src);
                const float aty = __shfl(ty0, src), bty = __shfl(ty1, src);
                const float ait = __shfl(it0, src), bit = __shfl(it1, src);
                const float qx = hi ? bx : ax, qy = hi ? by : ay, qtx = hi ? btx : atx, qty = hi ? bty : aty;
                const float qit = hi ? bit : ait;
                const float ttx = qtx * qit, tty = qty * qit;
                const int nit = kPackPlan.trips[na][np];
                float K = 0.0f, X0 = 0.0f, X1 = 0.0f, Y0 = 0.0f, Y1 = 0.0f;
                if (on) {
                    if (s0 < nP) {
                        for (int it = 0; it < nit; ++it) {
                            const LdsVox e = S.vox[it * G + group];
                            // same terms as adjoint_item
                            const float a = e.hx - qx, b = e.hy - qy;
                            const float u = fmaf(e.hz, qtx, a), w = fmaf(e.hz, qty, b);
                            const float qr = fmaf(u, ttx, fmaf(w, tty, -e.hz));
                            const float px = fmaf(-qr, qtx, a), py = fmaf(-qr, qty, b);
                            const float d2 = fmaf(u, px, w * py);
                            const float e1 = __builtin_amdgcn_exp2f(e.a1 * d2);
                            const float e2 = __builtin_amdgcn_exp2f(e.a2 * d2);
                            K = fmaf(e.GA, e1, K);
                            K = fmaf(e.GB, e2, K);
                            const float F = fmaf(e.GAa, e1, e.GBb * e2);
                            const float Fq = F * qr;
                            X0 = fmaf(F, px, X0);
                            X1 = fmaf(Fq, px, X1);
                            Y0 = fmaf(F, py, Y0);
                            Y1 = fmaf(Fq, py, Y1);
                        }
                    } else {
                        // every spot of the pass is beyond the primary radius: the halo term alone
                        for (int it = 0; it < nit; ++it) {
                            const LdsVox e = S.vox[it * G + group];
                            const float a = e.hx - qx, b = e.hy - qy;
                            const float u = fmaf(e.hz, qtx, a), w = fmaf(e.hz, qty, b);
                            const float qr = fmaf(u, ttx, fmaf(w, tty, -e.hz));
                            const float px = fmaf(-qr, qtx, a), py = fmaf(-qr, qty, b);
                            const float d2 = fmaf(u, px, w * py);
                            const float e2 = __builtin_amdgcn_exp2f(e.a2 * d2);
                            K = fmaf(e.GB, e2, K);
                            const float F = e.GBb * e2;
                            const float Fq = F * qr;
                            X0 = fmaf(F, px, X0);
                            X1 = fmaf(Fq, px, X1);
                            Y0 = fmaf(F, py, Y0);
                            Y1 = fmaf(Fq, py, Y1);
                        }
                    }
                }
                S.part[lane] = make_float4(K, fmaf(X1, ivx, X0), fmaf(Y1, ivy, Y0), 0.0f);
                __syncthreads();
                // owners add their spot's G subset sums in group order
                const int o0 = p0 - s0, o1 = p1 - s0;
                if (r0 && o0 >= 0 && o0 < np) {
                    float4 t = S.part[o0];
                    for (int g = 1; g < G; ++g) {
                        const float4 y = S.part[g * np + o0];
                        t.x += y.x;
                        t.y += y.y;
                        t.z += y.z;
                    }
                    K0 += t.x;
                    FX0 += t.y;
                    FY0 += t.z;
                }
                if (r1 && o1 >= 0 && o1 < np) {
                    float4 t = S.part[o1];
                    for (int g = 1; g < G; ++g) {
                        const float4 y = S.part[g * np + o1];
                        t.x += y.x;
                        t.y += y.y;
                        t.z += y.z;
                    }
                    K1 += t.x;
                    FX1 += t.y;
                    FY1 += t.z;
                }
                __syncthreads();
            }
        }
    }
    if (COUNT) return;
    const long long row = (long long)slab * c.n_spots + L.start + c0;
    if (v0) gst(c.partial, row + lane, make_float4(K0, FX0, FY0, 0.0f));
    if (v1) gst(c.partial, row + lane + WAVE, make_float4(K1, FX1, FY1, 0.0f));
}

template <bool CULL, int CS, bool COUNT = false>
__global__ __launch_bounds__(64) void k_pb_adjoint(BeamCtx c, unsigned long long *__restrict__ counts = nullptr) {
    unsigned long long nd = 0, ne = 0;
    if constexpr (CS == PACK_CS) {
        __shared__ PackLds s_pack;
        adjoint_item_packed<CULL, COUNT>(c, blockIdx.x, blockIdx.y, threadIdx.x, s_pack, nd, ne);
    } else {
        __shared__ LdsVox s_vox[WAVE];
        adjoint_item<CULL, CS, COUNT>(c, blockIdx.x, blockIdx.y, threadIdx.x, s_vox, nd, ne);
    }
    if (COUNT && threadIdx.x == 0) {
        atomicAdd(&counts[2], nd);
        atomicAdd(&counts[3], ne);
    }
}

// Batched adjoint: beam b owns items [off[b], off[b + 1]) = chunk x slab.  One
// launch per chunk size: beams of the other size own no items of it.
template <bool CULL, int CS>
__global__ __launch_bounds__(64) void k_pb_adjoint_batch(const BeamCtx *__restrict__ ctx,
                                                        const long long *__restrict__ off, int n_beams,
                                                        long long base, long long total) {
    const long long item = base + blockIdx.x;
    if (item >= total) return;
    const int bm = find_owner(off, n_beams, item);
    const BeamCtx &c = load_ctx(ctx, bm);
    const long long local = item - off[bm];
    // items of one beam fit in 32 bits (chunks x slabs)
    const unsigned lu = (unsigned)local, ns = (unsigned)c.nslab;
    const int cid = (int)(lu / ns), slab = (int)(lu - (unsigned)(lu / ns) * ns);
    unsigned long long nd = 0, ne = 0;
    if constexpr (CS == PACK_CS) {
        __shared__ PackLds s_pack;
        adjoint_item_packed<CULL, false>(c, cid, slab, threadIdx.x, s_pack, nd, ne);
    } else {
        __shared__ LdsVox s_vox[WAVE];
        adjoint_item<CULL, CS, false>(c, cid, slab, threadIdx.x, s_vox, nd, ne);
    }
}

// 256 threads = 64 spots x 4 slab groups; each group sums every 4th slab,
// then the 4 group sums are added in a fixed order.  Every adjoint workgroup of
// a usable layer writes its chunk's partials for its slab, so the partials of
// a spot in a usable layer are all written and nothing needs clearing first;
// a spot outside every usable layer (layer records are disjoint runs) gets
// zero gradients without reading its partials.
__device__ __forceinline__ void reduce_spots(const BeamCtx &c, long long block, int tid) {
    __shared__ float4 s_part[4][64];
    const int sl = tid & 63, grp = tid >> 6;
    const long long s = block * 64 + sl;
    bool covered = false;
    for (int l = 0; l < c.n_layers; ++l) {
        const LayerC L = cld(c.lc, l);  // n = 0 for a layer the kernels skip
        covered = covered || (unsigned long long)(s - L.start) < (unsigned long long)L.n;
    }
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    if (s < c.n_spots && covered) {
        // eight slabs' loads in flight at a time; the sums keep slab order
        int d = grp;
        for (; d + 28 < c.nslab; d += 32) {
            float4 p[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) p[u] = gld(c.partial, (long long)(d + 4 * u) * c.n_spots + s);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                acc.x += p[u].x;
                acc.y += p[u].y;
                acc.z += p[u].z;
            }
        }
        for (; d < c.nslab; d += 4) {
            const float4 p = gld(c.partial, (long long)d * c.n_spots + s);
            acc.x += p.x;
            acc.y += p.y;
            acc.z += p.z;
        }
    }
    s_part[grp][sl] = acc;
    __syncthreads();
    if (grp == 0 && s < c.n_spots) {
        float K = 0.f, FX = 0.f, FY = 0.f;
        for (int q = 0; q < 4; ++q) {
            K += s_part[q][sl].x;
            FX += s_part[q][sl].y;
            FY += s_part[q][sl].z;
        }
        const float mu2 = -2.0f * gld(c.sp, s).mu;
        gst(c.grad_mu, s, K);
        gst(c.grad_xy, 2 * s, mu2 * FX);
        gst(c.grad_xy, 2 * s + 1, mu2 * FY);
    }
}

__global__ __launch_bounds__(256) void k_pb_reduce(BeamCtx c) { reduce_spots(c, blockIdx.x, threadIdx.x); }

__global__ __launch_bounds__(256) void k_pb_reduce_batch(const BeamCtx *__restrict__ ctx,
                                                        const long long *__restrict__ off, int n_beams) {
    const long long blk = blockIdx.x;
    const int bm = find_owner(off, n_beams, blk);
    const BeamCtx &c = load_ctx(ctx, bm);
    reduce_spots(c, blk - off[bm], threadIdx.x);
}

// ---------------------------------------------------------------------------
// Host side.
static constexpr long long kAlign = 256;
static long long align_up(long long x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Spots per adjoint workgroup when the schedule leaves it open.  With culling,
// the packed chunks, which fill the lanes per brick with the spots in reach
// (the autotuner chose them over 32- and 64-spot chunks on the head-phantom,
// the uneven-plan and the 1 mm body workloads).  Without culling every spot
// reaches every brick, the fixed chunks fill their lanes as well and cost
// less per brick: 64, or 32 (two voxel subsets per wave) when the average
// layer is small enough that 64-spot chunks would leave many lanes idle.
// Scores: lane use / (1 + per-brick overhead share), the overhead share
// growing as chunks shrink.
int adjoint_chunk(int n_spots, int n_layers, bool cull) {
    if (cull) return PACK_CS;
    if (n_layers <= 0) return 64;
    const double n = (double)n_spots / n_layers;
    const double u64 = n / (ceil(n / 64.0) * 64.0), u32 = n / (ceil(n / 32.0) * 32.0);
    return (u32 / 1.38 > u64 / 1.23) ? 32 : 64;
}

// Slabs of bricks per spot chunk: enough (chunk, slab) workgroups (wg_target)
// that the few heavy ones (chunks and slabs in the beam path) do not leave the
// chip idle at the end, with at least min_bricks bricks per slab and at most
// max_slabs slabs.
int adjoint_slabs(long long nbricks, int n_spots, int n_layers, int cs, const AdjTune &t) {
    const long long nch = adjoint_chunks(n_spots, n_layers, cs);
    long long d = (t.wg_target + nch - 1) / nch;
    if (d > t.max_slabs) d = t.max_slabs;
    if (d > nbricks / t.min_bricks) d = nbricks / t.min_bricks;
    if (d > nbricks) d = nbricks;
    if (d < 1) d = 1;
    return (int)d;
}

AdjTune default_tune() {
    AdjTune t;
    t.cs = 0;
    t.wg_target = 32768;
    t.min_bricks = 64;
    t.max_slabs = 1024;
    return t;
}

bool tune_ok(const AdjTune &t) {
    return (t.cs == 0 || t.cs == 32 || t.cs == 64 || t.cs == PACK_CS) && t.wg_target >= 1 && t.min_bricks >= 1 && t.max_slabs >= 1 &&
           t.max_slabs <= 65535;
}

// Workspace of one beam: [SpotExt x S][LayerC x L][adjoint partials float4 x
// nslab*S][brick summaries float2 x bricks], each part 256-byte aligned.
BeamWs carve_beam(void *base, long long nbricks, int S, int L, const AdjTune &t, bool cull) {
    BeamWs w;
    char *p = reinterpret_cast<char *>(base);
    w.cs = t.cs ? t.cs : adjoint_chunk(S, L, cull);
    w.nslab = adjoint_slabs(nbricks, S, L, w.cs, t);
    long long off = 0;
    w.ext = reinterpret_cast<SpotExt *>(p + off);
    off += align_up((long long)S * (long long)sizeof(SpotExt));
    w.layer = reinterpret_cast<LayerC *>(p + off);
    off += align_up((long long)(L > 0 ? L : 1) * (long long)sizeof(LayerC));
    w.partial = reinterpret_cast<float4 *>(p + off);
    off += align_up((long long)w.nslab * S * 16LL);
    w.bsum = reinterpret_cast<float2 *>(p + off);
    off += align_up(nbricks * 8LL);
    w.bytes = off;
    return w;
}

static long long grid_bricks(const dh_grid &g) {
    return (long long)((g.nx + 3) / 4) * ((g.ny + 3) / 4) * ((g.nz + 3) / 4);
}

bool fill_model(const dh_beam_model *bm, Model &m) {
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

bool pencil_grid_ok(const dh_grid &g) {
    if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0 || !(g.spacing > 0.0f)) {
        set_error("invalid grid %d x %d x %d, spacing %g", g.nx, g.ny, g.nz, (double)g.spacing);
        return false;
    }
    if ((double)g.nx * g.ny * g.nz >= 2147483647.0) {
        set_error("grid %d x %d x %d too large (below 2^31 voxels required)", g.nx, g.ny, g.nz);
        return false;
    }
    return true;
}

bool spot_set_ok(const dh_spot_set *ss) {
    // spots without layers are allowed (no spot is delivered: zero dose and
    // gradients); the layer table is only read when n_layers > 0
    if (!ss || ss->n_spots < 0 || ss->n_layers < 0 || (ss->n_spots > 0 && !ss->spots) ||
        (ss->n_spots > 0 && ss->n_layers > 0 && !ss->layers)) {
        set_error("invalid spot set (n_spots=%d, n_layers=%d)", ss ? ss->n_spots : -1, ss ? ss->n_layers : -1);
        return false;
    }
    return true;
}

float brick_radius(float spacing) { return 0.5f * (BRICK - 1) * 1.7320508075688772f * spacing * 1.0001f; }

// Fill the pencil-beam part of a beam context.
void fill_pencil_ctx(BeamCtx &c, const dh_grid &g, const dh_beam_geom &geom, const Model &m, const dh_spot_set &ss,
                     const BeamWs &w, float cutoff_sigma) {
    c.g = make_grid(g);
    c.b = make_geom(geom);
    c.bg = brick_grid(c.g);
    c.m = m;
    c.spots_in = ss.spots;
    c.layers_in = ss.layers;
    c.n_spots = ss.n_spots;
    c.n_layers = ss.n_spots > 0 ? ss.n_layers : 0;
    c.sp = w.ext;
    c.lc = w.layer;
    c.partial = w.partial;
    c.bsum = w.bsum;
    c.nslab = w.nslab;
    c.cs = w.cs;
    c.cut = cutoff_sigma > 0.0f ? cutoff_sigma : 0.0f;
    c.rbrick = brick_radius(g.spacing);
}

// Launch helpers.  Item spaces larger than kMaxBlocks workgroups go out in
// several launches (offset by `base`).
static constexpr long long kMaxBlocks = 1LL << 22;

static int prepare(hipStream_t st, const BeamCtx &c) {
    if (c.n_spots == 0) return DH_OK;
    hipLaunchKernelGGL(k_spot_prep, dim3((c.n_spots + 255) / 256), dim3(256), 0, st, c);
    int rc = check_launch("k_spot_prep");
    if (rc || c.n_layers == 0) return rc;
    hipLaunchKernelGGL(k_layer_prep, dim3(c.n_layers), dim3(64), 0, st, c);
    return check_launch("k_layer_prep");
}

template <bool COUNT>
static int launch_forward(hipStream_t st, const BeamCtx &c, int accumulate, unsigned long long *counts) {
    const unsigned nb = (unsigned)((c.bg.count() + 3) / 4);
    if (c.cut > 0.0f)
        hipLaunchKernelGGL((k_pb_forward<true, COUNT>), dim3(nb), dim3(256), 0, st, c, accumulate, counts);
    else
        hipLaunchKernelGGL((k_pb_forward<false, COUNT>), dim3(nb), dim3(256), 0, st, c, accumulate, counts);
    return check_launch(COUNT ? "k_pb_forward(count)" : "k_pb_forward");
}

template <bool COUNT>
static int launch_adjoint(hipStream_t st, const BeamCtx &c, unsigned long long *counts) {
    hipLaunchKernelGGL(k_brick_summary, dim3((unsigned)((c.bg.count() + 3) / 4)), dim3(256), 0, st, c);
    int rc = check_launch("k_brick_summary");
    if (rc) return rc;
    const dim3 grid((unsigned)adjoint_chunks(c.n_spots, c.n_layers, c.cs), c.nslab);
    const bool cull = c.cut > 0.0f;
    if (c.cs == PACK_CS) {
        if (cull) hipLaunchKernelGGL((k_pb_adjoint<true, PACK_CS, COUNT>), grid, dim3(64), 0, st, c, counts);
        else hipLaunchKernelGGL((k_pb_adjoint<false, PACK_CS, COUNT>), grid, dim3(64), 0, st, c, counts);
    } else if (c.cs == 32) {
        if (cull) hipLaunchKernelGGL((k_pb_adjoint<true, 32, COUNT>), grid, dim3(64), 0, st, c, counts);
        else hipLaunchKernelGGL((k_pb_adjoint<false, 32, COUNT>), grid, dim3(64), 0, st, c, counts);
    } else {
        if (cull) hipLaunchKernelGGL((k_pb_adjoint<true, 64, COUNT>), grid, dim3(64), 0, st, c, counts);
        else hipLaunchKernelGGL((k_pb_adjoint<false, 64, COUNT>), grid, dim3(64), 0, st, c, counts);
    }
    return check_launch(COUNT ? "k_pb_adjoint(count)" : "k_pb_adjoint");
}

// Batched launches: ctx is a device table of beam contexts, the offset tables
// are device arrays of ascending item starts (first entry 0).
int launch_prep_batch(const BeamCtx *ctx, const long long *spot_off, const long long *layer_off, int n_beams,
                      long long total_spots, long long total_layers, hipStream_t st) {
    if (total_spots > 0) {
        hipLaunchKernelGGL(k_spot_prep_batch, dim3((unsigned)((total_spots + 255) / 256)), dim3(256), 0, st, ctx,
                           spot_off, n_beams, total_spots);
        const int rc = check_launch("k_spot_prep_batch");
        if (rc) return rc;
    }
    if (total_layers > 0) {
        hipLaunchKernelGGL(k_layer_prep_batch, dim3((unsigned)total_layers), dim3(64), 0, st, ctx, layer_off,
                           n_beams);
        return check_launch("k_layer_prep_batch");
    }
    return DH_OK;
}

int launch_wet_need_batch(const BeamCtx *ctx, const long long *brick_off, int n_beams, long long total_bricks,
                          hipStream_t st, void *const *ev) {
    mark(ev, 2 * DH_TIME_NEED, st);
    // 256 bricks per workgroup (one per lane)
    constexpr long long kPer = 4 * WAVE;
    for (int pass = 0; pass < 2; ++pass) {
        for (long long base = 0; base < total_bricks; base += kPer * kMaxBlocks) {
            const long long n = total_bricks - base < kPer * kMaxBlocks ? total_bricks - base : kPer * kMaxBlocks;
            const dim3 grid((unsigned)((n + kPer - 1) / kPer));
            if (pass == 0)
                hipLaunchKernelGGL(k_wet_need_batch, grid, dim3(256), 0, st, ctx, brick_off, n_beams, base,
                                   total_bricks);
            else
                hipLaunchKernelGGL(k_wet_dilate_batch, grid, dim3(256), 0, st, ctx, brick_off, n_beams, base,
                                   total_bricks);
            const int rc = check_launch(pass == 0 ? "k_wet_need_batch" : "k_wet_dilate_batch");
            if (rc) return rc;
        }
    }
    mark(ev, 2 * DH_TIME_NEED + 1, st);
    return DH_OK;
}

int launch_forward_batch(const BeamCtx *ctx, const long long *brick_off, int n_beams, long long total_bricks,
                         const PlanCtx *plans, const long long *sum_off, int n_plans, long long total_sum_blocks,
                         bool cull, int persist_blocks, hipStream_t st, void *const *ev) {
    mark(ev, 2 * DH_TIME_FORWARD, st);
    if (persist_blocks > 0 && total_bricks > 0) {
        const long long nb = (total_bricks + 3) / 4 < persist_blocks ? (total_bricks + 3) / 4 : persist_blocks;
        const long long stride = 4 * nb;
        if (cull)
            hipLaunchKernelGGL((k_pb_forward_batch<true, true>), dim3((unsigned)nb), dim3(256), 0, st, ctx, brick_off,
                               n_beams, 0LL, total_bricks, stride);
        else
            hipLaunchKernelGGL((k_pb_forward_batch<false, true>), dim3((unsigned)nb), dim3(256), 0, st, ctx,
                               brick_off, n_beams, 0LL, total_bricks, stride);
        const int rc = check_launch("k_pb_forward_batch");
        if (rc) return rc;
    }
    for (long long base = 0; persist_blocks <= 0 && base < total_bricks; base += 4 * kMaxBlocks) {
        const long long n = total_bricks - base < 4 * kMaxBlocks ? total_bricks - base : 4 * kMaxBlocks;
        const dim3 grid((unsigned)((n + 3) / 4));
        if (cull)
            hipLaunchKernelGGL((k_pb_forward_batch<true, false>), grid, dim3(256), 0, st, ctx, brick_off, n_beams,
                               base, total_bricks, total_bricks);
        else
            hipLaunchKernelGGL((k_pb_forward_batch<false, false>), grid, dim3(256), 0, st, ctx, brick_off, n_beams,
                               base, total_bricks, total_bricks);
        const int rc = check_launch("k_pb_forward_batch");
        if (rc) return rc;
    }
    mark(ev, 2 * DH_TIME_FORWARD + 1, st);
    if (total_sum_blocks > 0) {
        hipLaunchKernelGGL(k_plan_sum, dim3((unsigned)total_sum_blocks), dim3(256), 0, st, plans, ctx, sum_off,
                           n_plans);
        return check_launch("k_plan_sum");
    }
    return DH_OK;
}

template <int CS>
static int launch_adjoint_items(const BeamCtx *ctx, int n_beams, const long long *item_off, long long total,
                                bool cull, hipStream_t st) {
    for (long long base = 0; base < total; base += kMaxBlocks) {
        const long long n = total - base < kMaxBlocks ? total - base : kMaxBlocks;
        if (cull)
            hipLaunchKernelGGL((k_pb_adjoint_batch<true, CS>), dim3((unsigned)n), dim3(64), 0, st, ctx, item_off,
                               n_beams, base, total);
        else
            hipLaunchKernelGGL((k_pb_adjoint_batch<false, CS>), dim3((unsigned)n), dim3(64), 0, st, ctx, item_off,
                               n_beams, base, total);
        const int rc = check_launch("k_pb_adjoint_batch");
        if (rc) return rc;
    }
    return DH_OK;
}

int launch_adjoint_batch(const BeamCtx *ctx, int n_beams, const long long *row_off, long long total_rows,
                         const long long *item_off32, long long total_items32, const long long *item_off64,
                         long long total_items64, const long long *item_off128, long long total_items128,
                         const long long *reduce_off, long long total_reduce, bool cull, hipStream_t st,
                         void *const *ev) {
    int rc;
    for (long long base = 0; base < total_rows; base += kMaxBlocks) {
        const long long n = total_rows - base < kMaxBlocks ? total_rows - base : kMaxBlocks;
        hipLaunchKernelGGL(k_brick_summary_batch, dim3((unsigned)n), dim3(256), 0, st, ctx, row_off, n_beams, base);
        if ((rc = check_launch("k_brick_summary_batch"))) return rc;
    }
    mark(ev, 2 * DH_TIME_ADJOINT, st);
    if ((rc = launch_adjoint_items<PACK_CS>(ctx, n_beams, item_off128, total_items128, cull, st))) return rc;
    if ((rc = launch_adjoint_items<64>(ctx, n_beams, item_off64, total_items64, cull, st))) return rc;
    if ((rc = launch_adjoint_items<32>(ctx, n_beams, item_off32, total_items32, cull, st))) return rc;
    mark(ev, 2 * DH_TIME_ADJOINT + 1, st);
    if (total_reduce > 0) {
        hipLaunchKernelGGL(k_pb_reduce_batch, dim3((unsigned)total_reduce), dim3(256), 0, st, ctx, reduce_off,
                           n_beams);
        return check_launch("k_pb_reduce_batch");
    }
    return DH_OK;
}

}  // namespace dh

using namespace dh;

static bool common_args(const dh_grid &g, const dh_beam_geom *geom, const dh_spot_set *ss, void *ws,
                        long long ws_bytes) {
    if (!pencil_grid_ok(g)) return false;
    if (!geom) {
        set_error("null beam geometry");
        return false;
    }
    if (!spot_set_ok(ss)) return false;
    const long long need = dh_pb_workspace_size(ss->n_spots, ss->n_layers, g);
    if (ss->n_spots > 0 && (!ws || ws_bytes < need || ((uintptr_t)ws % kAlign) != 0)) {
        set_error("workspace too small or misaligned (%lld bytes given, %lld needed, 256-byte alignment)",
                  (long long)ws_bytes, need);
        return false;
    }
    return true;
}

static BeamCtx single_ctx(const dh_grid &g, const dh_beam_geom *geom, const Model &m, const dh_spot_set *ss,
                          void *ws, float cutoff_sigma) {
    BeamCtx c = {};
    const BeamWs w = carve_beam(ws, grid_bricks(g), ss->n_spots, ss->n_layers, default_tune(), cutoff_sigma > 0.0f);
    fill_pencil_ctx(c, g, *geom, m, *ss, w, cutoff_sigma);
    return c;
}

extern "C" int dh_spot_params(dh_spot *spots, const int64_t *order, const float *mu, const float *xy, int32_t n,
                              void *stream) {
    if (n < 0 || (n > 0 && (!spots || !order))) {
        set_error("dh_spot_params: invalid arguments (n=%d)", n);
        return DH_ERR_ARG;
    }
    if (n == 0 || (!mu && !xy)) return DH_OK;
    hipLaunchKernelGGL(k_spot_params, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, spots,
                       reinterpret_cast<const long long *>(order), mu, xy, n);
    return check_launch("k_spot_params");
}

extern "C" int64_t dh_pb_workspace_size(int32_t n_spots, int32_t n_layers, dh_grid g) {
    if (n_spots <= 0) return 0;
    // the schedule depends on the cutoff, which this query does not know: the larger layout
    const long long a = carve_beam(nullptr, grid_bricks(g), n_spots, n_layers, default_tune(), true).bytes;
    const long long b = carve_beam(nullptr, grid_bricks(g), n_spots, n_layers, default_tune(), false).bytes;
    return a > b ? a : b;
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
    BeamCtx c = single_ctx(g, geom, m, spots, workspace, cutoff_sigma);
    c.wet = const_cast<float *>(wet);
    c.dose = dose;
    const int rc = prepare(st, c);
    if (rc) return rc;
    return launch_forward<false>(st, c, accumulate, nullptr);
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
    BeamCtx c = single_ctx(g, geom, m, spots, workspace, cutoff_sigma);
    c.wet = const_cast<float *>(wet);
    c.gdose = grad_dose;
    c.grad_mu = grad_mu;
    c.grad_xy = grad_xy;
    int rc = prepare(st, c);
    if (rc) return rc;
    if ((rc = launch_adjoint<false>(st, c, nullptr))) return rc;
    hipLaunchKernelGGL(k_pb_reduce, dim3((S + 63) / 64), dim3(256), 0, st, c);
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
    if (spots->n_spots == 0) return DH_OK;  // counts are accumulated into; nothing to add
    Model m;
    if (!fill_model(model, m)) return DH_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    BeamCtx c = single_ctx(g, geom, m, spots, workspace, cutoff_sigma);
    c.wet = const_cast<float *>(wet);
    c.gdose = grad_dose;
    int rc = prepare(st, c);
    if (rc) return rc;
    if ((rc = launch_forward<true>(st, c, 0, counts))) return rc;
    if (!grad_dose) return DH_OK;
    return launch_adjoint<true>(st, c, counts);
}

// ---------------------------------------------------------------------------
// dh_pb_adjoint without a workspace argument, for callers that cannot
// allocate device memory.  The beam's workspace (prepared spot and layer
// records, adjoint partial sums, brick summaries) lives in `scratch`, one
// array per device for the life of the process.  Nothing in it needs to start
// at zero: every part is written by an earlier kernel of the same call before
// any kernel reads it (prepare, k_brick_summary, every adjoint workgroup of a
// usable layer writes its partial rows, k_pb_reduce reads only rows of spots
// in usable layers).  Calls that overlap on one device share the array and
// race.
static __device__ __align__(256) unsigned char scratch[DH_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept; nullptr on failure (error set).
void *scratch_address() {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error("dh_pb_adjoint_noalloc: no usable current device (%s, device %d)", hipGetErrorString(e), dev);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("dh_pb_adjoint_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" int dh_pb_adjoint_noalloc(const float *wet, const float *grad_dose, dh_grid g, const dh_beam_geom *geom,
                                     const dh_beam_model *model, const dh_spot_set *spots, float cutoff_sigma,
                                     float *grad_mu, float *grad_xy, void *stream) {
    if (!pencil_grid_ok(g)) return DH_ERR_ARG;
    if (!geom) {
        set_error("null beam geometry");
        return DH_ERR_ARG;
    }
    if (!spot_set_ok(spots)) return DH_ERR_ARG;
    const int S = spots->n_spots;
    if (S == 0) return DH_OK;
    if (!wet || !grad_dose || !grad_mu || !grad_xy) {
        set_error("dh_pb_adjoint_noalloc: null pointer");
        return DH_ERR_ARG;
    }
    Model m;
    if (!fill_model(model, m)) return DH_ERR_ARG;
    // dh_pb_adjoint's schedule when its layout fits the array; otherwise the
    // most slabs that fit (the partial sums then add in another order)
    const long long nbricks = grid_bricks(g);
    const bool cull = cutoff_sigma > 0.0f;
    AdjTune t = default_tune();
    BeamWs w = carve_beam(nullptr, nbricks, S, spots->n_layers, t, cull);
    if (w.bytes > DH_NOALLOC_SCRATCH_BYTES) {
        const long long row = (long long)S * 16LL;
        const long long fixed = w.bytes - (long long)w.nslab * row;  // upper bound of the slab-free parts
        const long long fit = (DH_NOALLOC_SCRATCH_BYTES - fixed - kAlign) / row;
        if (fit >= 1) {
            t.max_slabs = (int)(fit < t.max_slabs ? fit : t.max_slabs);
            w = carve_beam(nullptr, nbricks, S, spots->n_layers, t, cull);
        }
        if (fit < 1 || w.bytes > DH_NOALLOC_SCRATCH_BYTES) {
            set_error("dh_pb_adjoint_noalloc: %d spots in %d layers on a %d x %d x %d grid need %lld bytes of "
                      "scratch with one slab, more than the %lld bytes the library holds per device; use "
                      "dh_pb_adjoint with a workspace",
                      S, spots->n_layers, g.nx, g.ny, g.nz, fixed + row, (long long)DH_NOALLOC_SCRATCH_BYTES);
            return DH_ERR_ARG;
        }
    }
    void *ws = scratch_address();
    if (!ws) return DH_ERR_HIP;
    w = carve_beam(ws, nbricks, S, spots->n_layers, t, cull);
    hipStream_t st = (hipStream_t)stream;
    BeamCtx c = {};
    fill_pencil_ctx(c, g, *geom, m, *spots, w, cutoff_sigma);
    c.wet = const_cast<float *>(wet);
    c.gdose = grad_dose;
    c.grad_mu = grad_mu;
    c.grad_xy = grad_xy;
    int rc = prepare(st, c);
    if (rc) return rc;
    if ((rc = launch_adjoint<false>(st, c, nullptr))) return rc;
    hipLaunchKernelGGL(k_pb_reduce, dim3((S + 63) / 64), dim3(256), 0, st, c);
    return check_launch("k_pb_reduce");
}
