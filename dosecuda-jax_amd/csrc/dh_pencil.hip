// Double-Gaussian pencil-beam spot x voxel kernels (gfx950, wave64).
//
//  k_spot_prep     spot records + per-spot constants of the beam-axis geometry
//  k_pb_forward    dose: lanes = voxels of a 4x4x4 brick, spots wave-uniform
//  k_pb_adjoint    dL/dmu, dL/dx, dL/dy partials: lanes = spots, voxel terms
//                  broadcast from LDS
//  k_pb_reduce     fixed-order sum of the adjoint partials
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

constexpr float kLog2e = 1.4426950408889634f;
constexpr float kTwoPi = 6.283185307179586f;

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
            float2 r = row[0];
            idd = r.x; sig = r.y;
            return;
        }
        if (q > (float)(L - 1)) {
            float2 r = row[L - 1];
            idd = r.x; sig = r.y;
            return;
        }
        lo = min((int)q, L - 2);
        f = q - (float)lo;
    } else {
        const float *dep = m.depths + (long long)eid * L;
        int l = 0, r = L;
        while (l < r) {
            int mid = (l + r) >> 1;
            if (wet <= dep[mid]) r = mid; else l = mid + 1;
        }
        if (l == 0 || l == L) {
            float2 v = row[l == 0 ? 0 : L - 1];
            idd = v.x; sig = v.y;
            return;
        }
        lo = l - 1;
        f = (wet - dep[lo]) / (dep[l] - dep[lo]);
    }
    const float2 a = row[lo], c = row[lo + 1];
    idd = fmaf(c.x - a.x, f, a.x);
    sig = fmaf(c.y - a.y, f, a.y);
}

// Per-layer constants, wave-uniform.
struct LayerK {
    int start, n, eid;
    float r80, lim, A, B, C;
};

__device__ __forceinline__ LayerK load_layer(const dh_layer *__restrict__ layers, const Model &m, int l) {
    dh_layer L = layers[l];
    LayerK k;
    k.start = L.spot_start;
    k.n = L.n_spots;
    k.eid = L.energy_id;
    k.r80 = L.r80;
    k.lim = 1.1f * L.r80;
    const int e = (unsigned)L.energy_id < (unsigned)m.n_energies ? L.energy_id : 0;
    const float *c = m.dvp + (long long)e * m.dvp_len;
    k.A = c[2];
    k.B = c[3];
    k.C = c[4];
    return k;
}

// A layer record the kernels can use safely: energy id inside the tables and
// spot range inside the spot array (the host builds such tables; records from
// elsewhere that break this are skipped rather than read out of bounds).
__device__ __forceinline__ bool layer_ok(const LayerK &L, const Model &m, int n_spots) {
    return (unsigned)L.eid < (unsigned)m.n_energies && L.start >= 0 && L.n > 0 && L.start <= n_spots - L.n;
}

// Per (voxel, layer) Gaussian parameters.
struct VoxTerms {
    float Af, Bf;    // primary / halo amplitude (idd * weight / (2 pi sigma^2 + 1e-8))
    float a1, a2;    // exponent scales in the exp2 domain: -0.5 log2(e) / sigma^2 (0 if sigma <= 0)
    float sh;        // halo sigma (culling bound)
};

__device__ __forceinline__ VoxTerms vox_terms(const Model &m, const LayerK &L, float wet, float dsrc) {
    float idd, sms;
    lut_lookup(m, L.eid, wet, idd, sms);
    const float r80 = L.r80;
    const float d = dsrc - wet + 0.7f * r80;
    const float st = fmaf(fmaf(L.A, d, L.B), d, L.C) + sms;

    // nuclear halo (nuclearHalo, IMPTClasses.cu:104-121)
    const float w = fminf(fmaxf(wet, 0.1f), r80 - 0.1f);
    const float e5 = w - r80 - 5.0f;
    const float r3 = r80 + 3.0f;
    float sn = 2.85f + 0.0014f * r80 * logf(w + 3.0f) + 0.06f * w - 7.4e-5f * w * w - 0.22f * r80 / (e5 * e5);
    float un = 0.052f * logf(1.13f + w / (11.2f - 0.023f * r80)) +
               0.35f * (0.0017f * r80 * r80 - r80) / (r3 * r3 - w * w) - 1.61e-9f * w * r3 * r3;
    sn = fmaxf(sn, 0.0f);
    un = fminf(fmaxf(un, 0.0f), 0.9f);

    const float st2 = st * st;
    const float sh2 = st2 + sn * sn;
    VoxTerms t;
    const bool pos1 = st > 0.0f, pos2 = sh2 > 0.0f;
    t.Af = pos1 ? (1.0f - un) * idd / (kTwoPi * st2 + 1e-8f) : 0.0f;
    t.Bf = pos2 ? un * idd / (kTwoPi * sh2 + 1e-8f) : 0.0f;
    t.a1 = pos1 ? (-0.5f * kLog2e) / st2 : 0.0f;
    t.a2 = pos2 ? (-0.5f * kLog2e) / sh2 : 0.0f;
    t.sh = sqrtf(sh2);
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
    long long t = brick / bg.bx;
    bj = (int)(t % bg.by);
    bk = (int)(t / bg.by);
}

// ---------------------------------------------------------------------------
// Forward.  One wave per 4x4x4 brick, one voxel per lane; a lane owns its
// voxel's sum (deterministic, no atomics).  Layers whose range no lane reaches
// are skipped by ballot.  With culling, each 64-spot chunk of a layer is tested
// lane-parallel against the brick (distance from the brick centre to the spot
// axis vs cutoff * max halo sigma + brick radius) and only the surviving spots
// are iterated, wave-uniformly (their records come in through scalar loads).
// COUNT: no dose; counts[0] += in-range voxels x layer spots (the reference's
// work), counts[1] += in-range voxels x spots surviving the cull (work done).
template <bool CULL, bool COUNT = false>
__global__ __launch_bounds__(256) void k_pb_forward(const float *__restrict__ wetv, float *__restrict__ dose, Grid g,
                                                   Geom b, BrickGrid bg, unsigned nblocks, Model m,
                                                   const dh_layer *__restrict__ layers, int n_layers,
                                                   const SpotExt *__restrict__ sp, int n_spots, float cut,
                                                   float rbrick, int accumulate,
                                                   unsigned long long *__restrict__ counts = nullptr) {
    const unsigned blk = xcd_remap(blockIdx.x, nblocks);
    const long long brick = (long long)blk * 4 + (threadIdx.x >> 6);
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

    float3 hc = make_float3(0.f, 0.f, 0.f);
    if (CULL) {
        const float half = 0.5f * (BRICK - 1);
        const float3 vc = make_float3(fmaf((float)bi * BRICK + half, g.s, -b.iso[0]),
                                      fmaf((float)bj * BRICK + half, g.s, -b.iso[1]),
                                      fmaf((float)bk * BRICK + half, g.s, -b.iso[2]));
        hc = image_to_head(b, vc);
    }

    float acc = 0.0f;
    unsigned long long n_dense = 0, n_exec = 0;
    for (int l = 0; l < n_layers; ++l) {
        const LayerK L = load_layer(layers, m, l);
        if (!layer_ok(L, m, n_spots)) continue;
        const bool act = wet <= L.lim;
        const unsigned long long amask = __ballot(act);
        if (!amask) continue;
        if (COUNT) n_dense += (unsigned long long)__popcll(amask) * (unsigned)L.n;
        const VoxTerms T = vox_terms(m, L, wet, dsrc);
        float rc2 = 0.0f;
        if (CULL) {
            const float shmax = wave_max(act ? T.sh : 0.0f);
            const float rc = fmaf(cut, shmax, rbrick);
            rc2 = rc * rc;
        }
        float P = 0.0f, H = 0.0f;
        const SpotExt *lsp = sp + L.start;
        for (int c0 = 0; c0 < L.n; c0 += WAVE) {
            const int nin = min(WAVE, L.n - c0);
            unsigned long long mask;
            if (CULL) {
                bool rel = false;
                if (lane < nin) {
                    const SpotExt q = lsp[c0 + lane];
                    rel = axis_d2(hc.x, hc.y, hc.z, q) < rc2;
                }
                mask = __ballot(rel);
            } else {
                mask = nin == WAVE ? ~0ull : ((1ull << nin) - 1ull);
            }
            if (COUNT) {
                n_exec += (unsigned long long)__popcll(amask) * (unsigned)__popcll(mask);
                continue;
            }
            while (mask) {
                const int s = __builtin_ctzll(mask);
                mask &= mask - 1ull;
                const SpotExt q = lsp[c0 + s];
                const float d2 = axis_d2(h.x, h.y, h.z, q);
                const float e1 = __builtin_amdgcn_exp2f(T.a1 * d2);
                const float e2 = __builtin_amdgcn_exp2f(T.a2 * d2);
                P = fmaf(q.mu, e1, P);
                H = fmaf(q.mu, e2, H);
            }
        }
        if (act) acc = fmaf(T.Af, P, fmaf(T.Bf, H, acc));
    }
    if (COUNT) {
        if (lane == 0) {
            atomicAdd(&counts[0], n_dense);
            atomicAdd(&counts[1], n_exec);
        }
        return;
    }
    if (valid) dose[idx] = accumulate ? dose[idx] + acc : acc;
}

// ---------------------------------------------------------------------------
// Adjoint.  Workgroup = one wave = one 64-spot chunk of one layer x one slab of
// bricks.  Lane = spot: its three sums stay in registers across the whole
// slab.  Per brick the wave first acts as lanes = voxels to compute the voxel
// terms, compacts the voxels that matter (in range, nonzero dL/dD) into LDS,
// then every spot lane walks them with broadcast LDS reads.
//   dL/dmu_s = sum_v g_v [A e1 + B e2]
//   dL/dx_s  = -2 mu_s sum_v g_v [A a1 e1 + B a2 e2] (1 + q/VSADX) p_x
//   dL/dy_s  = -2 mu_s sum_v g_v [A a1 e1 + B a2 e2] (1 + q/VSADY) p_y
// with r = (hx - x_s, hy - y_s, hz) the voxel relative to the spot,
// q = (t . r) / |t|^2, p = r - q t (its part perpendicular to the axis) and
// a1, a2 = -1/(2 sigma^2) (natural-log scale).
struct LdsVox {
    float hx, hy, hz, a1;
    float a2, GA, GB, GAa;
    float GBb, pad0, pad1, pad2;
};

// COUNT: no gradient; counts[2] += voxels (dL/dD != 0, in range) x chunk
// spots (the reference's work), counts[3] += those voxels x surviving spots.
template <bool CULL, bool COUNT = false>
__global__ __launch_bounds__(64) void k_pb_adjoint(const float *__restrict__ wetv, const float *__restrict__ gdose,
                                                  Grid g, Geom b, BrickGrid bg, Model m,
                                                  const dh_layer *__restrict__ layers, int n_layers,
                                                  const SpotExt *__restrict__ sp, int n_spots, float cut,
                                                  float rbrick, int nslab, const unsigned char *__restrict__ flags,
                                                  float4 *__restrict__ partial,
                                                  unsigned long long *__restrict__ counts = nullptr) {
    __shared__ LdsVox s_vox[WAVE];
    const int lane = threadIdx.x;
    // chunk -> (layer, first spot of the chunk within the layer)
    int cid = blockIdx.x, l = 0, c0 = -1;
    for (; l < n_layers; ++l) {
        const int nl = layers[l].n_spots;
        const int nch = (nl + WAVE - 1) / WAVE;
        if (cid < nch) {
            c0 = cid * WAVE;
            break;
        }
        cid -= nch;
    }
    if (c0 < 0) return;  // surplus workgroup of the upper-bound grid
    const LayerK L = load_layer(layers, m, l);
    if (!layer_ok(L, m, n_spots)) return;
    const int slab = blockIdx.y;
    const bool sval = c0 + lane < L.n;
    const int spot = L.start + c0 + lane;
    SpotExt q;
    if (sval) q = sp[spot];
    else q = SpotExt{0.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 1.f};
    const float itt = 1.0f / (q.tx * q.tx + q.ty * q.ty + 1.0f);
    const float ivx = 1.0f / b.vsadx, ivy = 1.0f / b.vsady;

    const long long nb = bg.count();
    const long long b0 = nb * slab / nslab, b1 = nb * (slab + 1) / nslab;
    float K = 0.0f, FX = 0.0f, FY = 0.0f;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    const unsigned long long chunk_n = (unsigned long long)min(WAVE, L.n - c0);
    unsigned long long n_dense = 0, n_exec = 0;

    // bricks flagged by k_brick_flags (any voxel with dL/dD != 0 in range),
    // 64 flags per load, visited in ascending order
    for (long long base = b0; base < b1; base += WAVE) {
      const bool fl = base + lane < b1 && flags[base + lane] != 0;
      unsigned long long fm = __ballot(fl);
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
            const float shmax = wave_max(act ? T.sh : 0.0f);
            const float rc = fmaf(cut, shmax, rbrick);
            const float half = 0.5f * (BRICK - 1);
            const float3 hc = image_to_head(b, make_float3(fmaf((float)bi * BRICK + half, g.s, -b.iso[0]),
                                                           fmaf((float)bj * BRICK + half, g.s, -b.iso[1]),
                                                           fmaf((float)bk * BRICK + half, g.s, -b.iso[2])));
            rel = sval && axis_d2(hc.x, hc.y, hc.z, q) < rc * rc;
            if (!__ballot(rel)) continue;
        }
        if (COUNT) {
            n_exec += (unsigned long long)__popcll(am) * (unsigned)__popcll(__ballot(rel));
            continue;
        }
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
            for (int t = 0; t < na; ++t) {
                const LdsVox e = s_vox[t];
                // in-plane offset e_xy, its part perpendicular to the axis p,
                // q_e = (e.t)/|t|^2; d2 = e_xy . p (no cancellation)
                const float u = fmaf(e.hz, q.tx, e.hx - q.x);
                const float v = fmaf(e.hz, q.ty, e.hy - q.y);
                const float qe = fmaf(u, q.tx, v * q.ty) * itt;
                const float px = fmaf(-qe, q.tx, u), py = fmaf(-qe, q.ty, v);
                const float d2 = fmaf(u, px, v * py);
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
    if (sval) partial[(long long)slab * n_spots + spot] = make_float4(K, FX, FY, 0.0f);
}

// One wave per brick: flag bricks holding a voxel with dL/dD != 0 whose WET
// is within reach of the deepest layer (wet <= 1.1 max R80).
__global__ __launch_bounds__(256) void k_brick_flags(const float *__restrict__ wetv, const float *__restrict__ gdose,
                                                    Grid g, BrickGrid bg, const dh_layer *__restrict__ layers,
                                                    int n_layers, unsigned char *__restrict__ flags) {
    const long long brick = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (brick >= bg.count()) return;
    const int lane = threadIdx.x & 63;
    float lim = -3.0e38f;
    for (int l = 0; l < n_layers; ++l) lim = fmaxf(lim, 1.1f * layers[l].r80);
    int bi, bj, bk;
    brick_of(bg, brick, bi, bj, bk);
    const int i = bi * BRICK + (lane & 3), j = bj * BRICK + ((lane >> 2) & 3), k = bk * BRICK + (lane >> 4);
    bool act = false;
    if (i < g.nx && j < g.ny && k < g.nz) {
        const long long idx = i + (long long)g.nx * (j + (long long)g.ny * k);
        act = gdose[idx] != 0.0f && wetv[idx] * 10.0f <= lim;
    }
    const unsigned long long am = __ballot(act);
    if (lane == 0) flags[brick] = am ? 1 : 0;
}

// Spots that no chunk covers (never happens for a consistent layer table, but
// keeps the outputs defined) are zeroed by the reduce reading zeroed partials.
__global__ __launch_bounds__(256) void k_zero4(float4 *__restrict__ p, long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(256) void k_pb_reduce(const float4 *__restrict__ partial, const SpotExt *__restrict__ sp,
                                                  int n_spots, int nslab, float *__restrict__ grad_mu,
                                                  float *__restrict__ grad_xy) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_spots) return;
    float K = 0.f, FX = 0.f, FY = 0.f;
    for (int d = 0; d < nslab; ++d) {
        const float4 p = partial[(long long)d * n_spots + s];
        K += p.x;
        FX += p.y;
        FY += p.z;
    }
    const float mu2 = -2.0f * sp[s].mu;
    grad_mu[s] = K;
    grad_xy[2 * s] = mu2 * FX;
    grad_xy[2 * s + 1] = mu2 * FY;
}

// ---------------------------------------------------------------------------
static constexpr long long kAlign = 256;
static long long align_up(long long x) { return (x + kAlign - 1) / kAlign * kAlign; }

static int adjoint_slabs(const dh_grid &g, int n_spots, int n_layers) {
    const long long nb = (long long)((g.nx + 3) / 4) * ((g.ny + 3) / 4) * ((g.nz + 3) / 4);
    const long long nch = ((long long)n_spots + 63LL * n_layers) / 64 + 1;
    long long d = (8192 + nch - 1) / nch;
    if (d > 256) d = 256;
    if (d > nb) d = nb;
    if (d < 1) d = 1;
    return (int)d;
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

}  // namespace dh

using namespace dh;

extern "C" int64_t dh_pb_workspace_size(int32_t n_spots, int32_t n_layers, dh_grid g) {
    if (n_spots <= 0) return 0;
    const int d = adjoint_slabs(g, n_spots, n_layers);
    const long long nb = (long long)((g.nx + 3) / 4) * ((g.ny + 3) / 4) * ((g.nz + 3) / 4);
    return align_up((long long)n_spots * (long long)sizeof(SpotExt)) + align_up((long long)d * n_spots * 16LL) +
           align_up(nb);
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
    const Grid gg = make_grid(g);
    const Geom gb = make_geom(*geom);
    SpotExt *ext = reinterpret_cast<SpotExt *>(workspace);
    const int S = spots->n_spots;
    if (S > 0) {
        hipLaunchKernelGGL(k_spot_prep, dim3((S + 255) / 256), dim3(256), 0, st, spots->spots, ext, S,
                           1.0f / geom->vsadx, 1.0f / geom->vsady);
        int rc = check_launch("k_spot_prep");
        if (rc) return rc;
    }
    const BrickGrid bg = brick_grid(gg);
    const unsigned nb = (unsigned)((bg.count() + 3) / 4);
    const float rbrick = 0.5f * (BRICK - 1) * 1.7320508075688772f * g.spacing * 1.0001f;
    const int nl = S > 0 ? spots->n_layers : 0;
    if (cutoff_sigma > 0.0f)
        hipLaunchKernelGGL(k_pb_forward<true>, dim3(nb), dim3(256), 0, st, wet, dose, gg, gb, bg, nb, m,
                           spots->layers, nl, ext, S, cutoff_sigma, rbrick, accumulate);
    else
        hipLaunchKernelGGL(k_pb_forward<false>, dim3(nb), dim3(256), 0, st, wet, dose, gg, gb, bg, nb, m,
                           spots->layers, nl, ext, S, 0.0f, rbrick, accumulate);
    return check_launch("k_pb_forward");
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
    const Grid gg = make_grid(g);
    const Geom gb = make_geom(*geom);
    SpotExt *ext = reinterpret_cast<SpotExt *>(workspace);
    float4 *partial = reinterpret_cast<float4 *>(reinterpret_cast<char *>(workspace) +
                                                 align_up((long long)S * (long long)sizeof(SpotExt)));
    const int nslab = adjoint_slabs(g, S, spots->n_layers);
    hipLaunchKernelGGL(k_spot_prep, dim3((S + 255) / 256), dim3(256), 0, st, spots->spots, ext, S,
                       1.0f / geom->vsadx, 1.0f / geom->vsady);
    int rc = check_launch("k_spot_prep");
    if (rc) return rc;
    const long long np = (long long)nslab * S;
    hipLaunchKernelGGL(k_zero4, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, partial, np);
    rc = check_launch("k_zero4");
    if (rc) return rc;
    const BrickGrid bg = brick_grid(gg);
    unsigned char *flags = reinterpret_cast<unsigned char *>(partial) + align_up(np * 16LL);
    hipLaunchKernelGGL(k_brick_flags, dim3((unsigned)((bg.count() + 3) / 4)), dim3(256), 0, st, wet, grad_dose, gg,
                       bg, spots->layers, spots->n_layers, flags);
    rc = check_launch("k_brick_flags");
    if (rc) return rc;
    const unsigned nch = (unsigned)(((long long)S + 63LL * spots->n_layers) / 64 + 1);
    const float rbrick = 0.5f * (BRICK - 1) * 1.7320508075688772f * g.spacing * 1.0001f;
    if (cutoff_sigma > 0.0f)
        hipLaunchKernelGGL(k_pb_adjoint<true>, dim3(nch, nslab), dim3(64), 0, st, wet, grad_dose, gg, gb, bg, m,
                           spots->layers, spots->n_layers, ext, S, cutoff_sigma, rbrick, nslab, flags, partial);
    else
        hipLaunchKernelGGL(k_pb_adjoint<false>, dim3(nch, nslab), dim3(64), 0, st, wet, grad_dose, gg, gb, bg, m,
                           spots->layers, spots->n_layers, ext, S, 0.0f, rbrick, nslab, flags, partial);
    rc = check_launch("k_pb_adjoint");
    if (rc) return rc;
    hipLaunchKernelGGL(k_pb_reduce, dim3((S + 255) / 256), dim3(256), 0, st, partial, ext, S, nslab, grad_mu,
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
    const Grid gg = make_grid(g);
    const Geom gb = make_geom(*geom);
    SpotExt *ext = reinterpret_cast<SpotExt *>(workspace);
    hipLaunchKernelGGL(k_spot_prep, dim3((S + 255) / 256), dim3(256), 0, st, spots->spots, ext, S,
                       1.0f / geom->vsadx, 1.0f / geom->vsady);
    int rc = check_launch("k_spot_prep");
    if (rc) return rc;
    const BrickGrid bg = brick_grid(gg);
    const unsigned nb = (unsigned)((bg.count() + 3) / 4);
    const float rbrick = 0.5f * (BRICK - 1) * 1.7320508075688772f * g.spacing * 1.0001f;
    const bool cull = cutoff_sigma > 0.0f;
    if (cull)
        hipLaunchKernelGGL((k_pb_forward<true, true>), dim3(nb), dim3(256), 0, st, wet, nullptr, gg, gb, bg, nb, m,
                           spots->layers, spots->n_layers, ext, S, cutoff_sigma, rbrick, 0, counts);
    else
        hipLaunchKernelGGL((k_pb_forward<false, true>), dim3(nb), dim3(256), 0, st, wet, nullptr, gg, gb, bg, nb, m,
                           spots->layers, spots->n_layers, ext, S, 0.0f, rbrick, 0, counts);
    rc = check_launch("k_pb_forward(count)");
    if (rc || !grad_dose) return rc;
    const int nslab = adjoint_slabs(g, S, spots->n_layers);
    float4 *partial = reinterpret_cast<float4 *>(reinterpret_cast<char *>(workspace) +
                                                 align_up((long long)S * (long long)sizeof(SpotExt)));
    const long long np = (long long)nslab * S;
    unsigned char *flags = reinterpret_cast<unsigned char *>(partial) + align_up(np * 16LL);
    hipLaunchKernelGGL(k_brick_flags, dim3((unsigned)((bg.count() + 3) / 4)), dim3(256), 0, st, wet, grad_dose, gg,
                       bg, spots->layers, spots->n_layers, flags);
    rc = check_launch("k_brick_flags");
    if (rc) return rc;
    const unsigned nch = (unsigned)(((long long)S + 63LL * spots->n_layers) / 64 + 1);
    if (cull)
        hipLaunchKernelGGL((k_pb_adjoint<true, true>), dim3(nch, nslab), dim3(64), 0, st, wet, grad_dose, gg, gb, bg,
                           m, spots->layers, spots->n_layers, ext, S, cutoff_sigma, rbrick, nslab, flags, partial,
                           counts);
    else
        hipLaunchKernelGGL((k_pb_adjoint<false, true>), dim3(nch, nslab), dim3(64), 0, st, wet, grad_dose, gg, gb,
                           bg, m, spots->layers, spots->n_layers, ext, S, 0.0f, rbrick, nslab, flags, partial,
                           counts);
    return check_launch("k_pb_adjoint(count)");
}
