// Shared device/host helpers of libdosehip (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dosehip.h"

namespace dh {

// ---------------------------------------------------------------------------
// errors (dh_api.cpp)
void set_error(const char *fmt, ...);
int check_launch(const char *what);

// Run radii of a block table (dh_wet.hip): float2 (value or NaN, radius) per
// 4x4x4 block of cells, nbx x nby x nbz blocks, x fastest; s0, s1 two int
// arrays of scratch of one int per block.  Radius r (0 .. 15): every block
// within Chebyshev distance r of a uniform block that lies in the table holds
// the same value.
int block_radius_passes(float2 *table, int *s0, int *s1, int nbx, int nby, int nbz, hipStream_t st);

// ---------------------------------------------------------------------------
// kernel-argument copies of the ABI structs (passed by value)
struct Grid {
    int nx, ny, nz;
    float s, inv_s;
};

struct Geom {
    float iso[3];
    float src[3];
    float sg, cg, sc, cc;
    float vsadx, vsady;
    float ivsadx, ivsady;   // 1 / vsad: correctly rounded on the host, as a device division rounds
};

inline Grid make_grid(const dh_grid &g) {
    Grid r;
    r.nx = g.nx; r.ny = g.ny; r.nz = g.nz;
    r.s = g.spacing;
    r.inv_s = 1.0f / g.spacing;
    return r;
}

inline Geom make_geom(const dh_beam_geom &b) {
    Geom r;
    for (int a = 0; a < 3; ++a) { r.iso[a] = b.iso[a]; r.src[a] = b.src[a]; }
    r.sg = b.sin_g; r.cg = b.cos_g; r.sc = b.sin_c; r.cc = b.cos_c;
    r.vsadx = b.vsadx; r.vsady = b.vsady;
    r.ivsadx = 1.0f / b.vsadx; r.ivsady = 1.0f / b.vsady;
    return r;
}

// Voxel bricks: 4x4x4 voxels, one wave64 per brick, lane = bx + 4*by + 16*bz.
constexpr int BRICK = 4;
constexpr int WAVE = 64;

struct BrickGrid {
    int bx, by, bz;      // bricks per axis
    __host__ __device__ long long count() const { return (long long)bx * by * bz; }
};

// Brick index -> brick coordinates.  Bricks of one beam fit in 32 bits (at
// most nx*ny*nz < 2^28), and 32-bit unsigned division costs the VALU a few
// instructions where 64-bit division is a long software sequence.
__device__ __forceinline__ void brick_coords(const BrickGrid &bg, long long brick, int &bi, int &bj, int &bk) {
    const unsigned b = (unsigned)brick, bx = (unsigned)bg.bx, by = (unsigned)bg.by;
    const unsigned t = b / bx;
    bi = (int)(b - t * bx);
    bk = (int)(t / by);
    bj = (int)(t - (unsigned)bk * by);
}

inline BrickGrid brick_grid(const Grid &g) {
    BrickGrid b;
    b.bx = (g.nx + BRICK - 1) / BRICK;
    b.by = (g.ny + BRICK - 1) / BRICK;
    b.bz = (g.nz + BRICK - 1) / BRICK;
    return b;
}

// gfx950 has 8 XCDs and the dispatcher hands consecutive workgroups to XCDs
// round-robin.  Remap so each XCD walks one contiguous eighth of the work,
// which keeps neighbouring bricks (and the volume rows their rays read) in one
// XCD's L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned bid, unsigned nblocks) {
    constexpr unsigned NXCD = 8;
    if (nblocks < NXCD) return bid;
    unsigned per = nblocks / NXCD, rem = nblocks % NXCD;
    unsigned xcd = bid % NXCD, slot = bid / NXCD;
    // XCDs [0, rem) own per+1 blocks, the rest own per.
    unsigned base = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
    return base + slot;
}

// Chunked form: XCD x takes chunks of C consecutive workgroups, chunk k of XCD
// x being workgroups ((k * 8 + x) * C .. + C).  All XCDs move through the work
// together (the working set of the whole chip stays one region of the item
// space, which the shared Infinity Cache holds) while each XCD's L2 sees runs
// of C neighbouring workgroups.  Bijective on [0, nblocks): the last partial
// round of 8 chunks keeps the identity mapping.
template <unsigned C>
__device__ __forceinline__ unsigned xcd_chunk(unsigned bid, unsigned nblocks) {
    constexpr unsigned NXCD = 8;
    const unsigned full = nblocks / (NXCD * C) * (NXCD * C);
    if (bid >= full) return bid;
    const unsigned xcd = bid % NXCD, slot = bid / NXCD;
    return ((slot / C) * NXCD + xcd) * C + slot % C;
}

// Voxel centre in image coordinates relative to the isocentre (pointIJKtoXYZ,
// CudaClasses.cuh:199-203): p = ijk * s - iso.
__device__ __forceinline__ float3 voxel_xyz(const Grid &g, const Geom &b, int i, int j, int k) {
    return make_float3(fmaf((float)i, g.s, -b.iso[0]),
                       fmaf((float)j, g.s, -b.iso[1]),
                       fmaf((float)k, g.s, -b.iso[2]));
}

// Image -> beam's-eye-view ("head") frame: couch rotation about y, gantry
// rotation about z, then (x, y, z)_head = (-x_g, z_g, y_g)
// (pointXYZImageToHead, CudaClasses.cu:65-95).
__device__ __forceinline__ float3 image_to_head(const Geom &b, float3 p) {
    float xt = p.x * b.cc - p.z * b.sc;
    float yt = p.y;
    float zt = p.x * b.sc + p.z * b.cc;
    float xg = xt * b.cg + yt * b.sg;
    float yg = yt * b.cg - xt * b.sg;
    return make_float3(-xg, zt, yg);
}

// Inverse of image_to_head (pointXYZHeadToImage, CudaClasses.cu:97-126).
__device__ __forceinline__ float3 head_to_image(const Geom &b, float3 h) {
    float xz = -h.x, yz = h.z, zz = h.y;
    float xg = xz * b.cg - yz * b.sg;
    float yg = xz * b.sg + yz * b.cg;
    float xt = xg * b.cc + zz * b.sc;
    float zt = zz * b.cc - xg * b.sc;
    return make_float3(xt, yg, zt);
}

// ---------------------------------------------------------------------------
// Shared records (device memory)

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

// Beam-model tables on the device.
struct Model {
    const float2 *idd_sigma;
    const float *depths;
    const float *dvp;
    int n_energies, lut_len, dvp_len, uniform;
    float depth0, inv_ddepth;
};

// Layer record (96 bytes): the layer's spot range and energy, the constants of
// its sigma_air polynomial, and the reach of its spots: bounding box of the
// positions at the isocentre plane, the smallest eigenvalue 1/(1 + |t_xy|^2)
// of their distance forms and max |sigma_ms| of the energy's lookup row; then
// the factors of the halo fit and the range that depend on R80 alone, formed
// once per layer exactly as the per-voxel expressions would form them.
struct LayerC {
    int start, n, eid, ok;
    float r80, A, B, C;
    float xmin, xmax, ymin, ymax;
    float lam_min, sms_max, lim, vx;  // vx: vertex -B / (2A) of sigma_air's quadratic (0 when A = 0)
    float sn_log, sn_pole, un_lin, un_pole;  // 0.0014 R80, 0.22 R80, 1 / (11.2 - 0.023 R80), 0.35 (0.0017 R80^2 - R80)
    float r3, d07, w_top, pad;               // R80 + 3, 0.7 R80, R80 - 0.1
};

// Smoothing neighbour stencil of a beam: per (direction, dr) entry, direction
// major, dr = 1..9 mm: integer base offsets, tie bits (1 = x, 2 = y, 4 = z).
constexpr int N_STENCIL = 54;
struct Stencil {
    signed char dx[N_STENCIL], dy[N_STENCIL], dz[N_STENCIL];
    unsigned char tie[N_STENCIL], dr[N_STENCIL];
    unsigned char pad[2];  // 272 bytes: the batched smoothing reads the table in whole dwords
};
static_assert(sizeof(Stencil) % 4 == 0, "the batched smoothing reads Stencil in dwords");

// Everything the kernels of one beam read or write.  Passed by value to the
// single-beam kernels and read from a device table by the batched ones.
struct BeamCtx {
    const float4 *rlsp_q;   // quad RLSP of the beam's plan (ray-march input)
    float *wet_raw;         // ray-march output
    float *wet;             // smoothed WET (cm): smoothing output, dose input
    float *dose;            // dose output of the beam's plan
    const float *gdose;     // dL/dD of the beam's plan (adjoint input)
    const dh_spot *spots_in;
    const dh_layer *layers_in;
    SpotExt *sp;            // prepared spot records
    LayerC *lc;             // prepared layer records
    float4 *partial;        // adjoint partials [nslab][n_spots]
    float2 *bsum;           // per-brick WET range of voxels with dL/dD != 0
    float *grad_mu, *grad_xy;
    const Stencil *stencil; // smoothing stencil (batched smoothing)
    unsigned char *need;    // per brick: bit 0 smoothed WET needed, bit 1 raw WET needed
    const int *touch;       // brick offsets smoothing reads raw WET from (make_touch)
    Model m;
    Grid g;
    Geom b;
    BrickGrid bg;
    int n_spots, n_layers, nslab, cs;
    float cut, rbrick;
    int n_touch;
    int sbox[3];            // tiled smoothing: LDS box (tile plus stencil halo) per axis
    int slo[3];             // tiled smoothing: halo below the tile per axis
};

// Tiled smoothing (batched): a workgroup smooths a 32 x 4 x 8 voxel tile
// (8 x 1 x 2 bricks) from an LDS box holding the tile plus the stencil's
// halo; beams whose box exceeds SMOOTH_BOX_MAX floats use the gather kernel.
constexpr int STX = 32, STY = 4, STZ = 8;
constexpr int SMOOTH_BOX_MAX = 16384;  // 64 KB of LDS

// Smoothing halo of a stencil (voxels below / above the tile per axis) and
// the resulting box; returns the box volume in floats.
int smooth_box(const Stencil &st, int slo[3], int sbox[3]);

// Voxels per block of the plan-sum pass (256 threads x 4).
constexpr long long PLAN_SUM_VOXELS = 1024;

// Most brick offsets a smoothing stencil reaches (54 entries x 2 x 2 x 2).
constexpr int MAX_TOUCH = 432;
static_assert(MAX_TOUCH % 8 == 0, "the dilation pass reads the touch table eight entries at a time");

// WET (cm) written where the reach-limited WET pass skips a brick: beyond any
// layer's range (1.1 R80), so the dose kernels skip those voxels.
constexpr float kWetOutside = 1.0e30f;

// A plan of a batch: its dose volume and its beams [beam0, beam0 + n_beams) of
// the beam-context table.
struct PlanCtx {
    float *dose;
    long long nvox;
    int beam0, n_beams;
};

// Adjoint schedule: spots per workgroup (cs: 32, 64, or 0 = chosen from the
// average layer size) and how bricks are cut into slabs (about wg_target
// workgroups per beam, at least min_bricks bricks per slab, at most max_slabs
// slabs).
struct AdjTune {
    int cs, wg_target, min_bricks, max_slabs;
};

// Schedules of the batched brick kernels: 0 = one wave per brick (hardware
// dispatch), n > 0 = a persistent grid of n 256-thread blocks.
struct BrickTune {
    int fwd_blocks, ray_blocks;
};

// Number of cs-spot adjoint chunks provided for a beam: an upper bound of the
// sum over layers of ceil(n_l / cs).
__host__ __device__ inline long long adjoint_chunks(int n_spots, int n_layers, int cs) {
    return ((long long)n_spots + (cs - 1LL) * n_layers) / cs + 1;
}

// Pencil-beam workspace of one beam (device memory).
struct BeamWs {
    SpotExt *ext;
    LayerC *layer;
    float4 *partial;
    float2 *bsum;
    int nslab, cs;
    long long bytes;
};

// Record timing event `slot` (dh_batch_timing) if timing is on.
inline void mark(void *const *ev, int slot, hipStream_t st) {
    if (ev && ev[slot]) (void)hipEventRecord((hipEvent_t)ev[slot], st);
}

// ---------------------------------------------------------------------------
// host helpers shared by the single-beam and batched entry points
// dh_wet.hip
void make_stencil(const dh_beam_geom &b, double s, Stencil &st);
int make_touch(const Stencil &st, int *out, int cap);
bool grid_ok(const dh_grid &g);
// block_off / total_blocks: smoothing items per beam -- 32 x 4 x 8 tiles when
// smooth_box_floats > 0 (the largest box of the batch), else 256-voxel blocks.
int launch_wet_batch(const BeamCtx *ctx, const long long *brick_off, const long long *block_off, int n_beams,
                     long long total_bricks, long long total_blocks, bool use_need, int persist_blocks,
                     int ray_shape, int smooth_box_floats, hipStream_t st, void *const *ev = nullptr);
// dh_pencil.hip
int adjoint_chunk(int n_spots, int n_layers, bool cull);
int adjoint_slabs(long long nbricks, int n_spots, int n_layers, int cs, const AdjTune &t);
AdjTune default_tune();
bool tune_ok(const AdjTune &t);
BeamWs carve_beam(void *base, long long nbricks, int S, int L, const AdjTune &t, bool cull);
bool fill_model(const dh_beam_model *bm, Model &m);
bool pencil_grid_ok(const dh_grid &g);
bool spot_set_ok(const dh_spot_set *ss);
float brick_radius(float spacing);
void fill_pencil_ctx(BeamCtx &c, const dh_grid &g, const dh_beam_geom &geom, const Model &m, const dh_spot_set &ss,
                     const BeamWs &w, float cutoff_sigma);
int launch_wet_need_batch(const BeamCtx *ctx, const long long *brick_off, int n_beams, long long total_bricks,
                          hipStream_t st, void *const *ev = nullptr);  // two launches: reach test, dilation
int launch_prep_batch(const BeamCtx *ctx, const long long *spot_off, const long long *layer_off, int n_beams,
                      long long total_spots, long long total_layers, hipStream_t st);
int launch_forward_batch(const BeamCtx *ctx, const long long *brick_off, int n_beams, long long total_bricks,
                         const PlanCtx *plans, const long long *sum_off, int n_plans, long long total_sum_blocks,
                         bool cull, int persist_blocks, hipStream_t st, void *const *ev = nullptr);
int launch_adjoint_batch(const BeamCtx *ctx, int n_beams, const long long *row_off, long long total_rows,
                         const long long *item_off32, long long total_items32, const long long *item_off64,
                         long long total_items64, const long long *item_off128, long long total_items128,
                         const long long *reduce_off, long long total_reduce, bool cull, hipStream_t st,
                         void *const *ev);

// ---------------------------------------------------------------------------
// wave64 helpers
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// Index of this thread's wave in its workgroup, as a wave-uniform value (the
// compiler cannot tell that threadIdx.x / 64 is uniform; readfirstlane lets
// everything derived from it live in scalar registers and load through the
// scalar cache).
__device__ __forceinline__ int wave_in_block() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// Pointers read from device memory (the batched kernels' context table) are
// generic pointers to the compiler, which then uses flat loads with 64-bit
// per-lane addresses and no scalar loads.  Every pointer in a context points to
// global memory: forming the address in the global address space
// (address_space(1)) and converting only the final element address back lets
// address-space inference turn the access into a global (or, when uniform,
// scalar) load.  A bare round trip p -> global -> generic would be folded away,
// hence base + offset inside the global space.
#define DH_GLOBAL __attribute__((address_space(1)))

template <class T>
__device__ __forceinline__ T *gptr(T *p, long long off) {
    return (T *)((DH_GLOBAL T *)p + off);
}

template <class T>
__device__ __forceinline__ T gld(const T *p, long long i) {
    return *gptr(p, i);
}

template <class T>
__device__ __forceinline__ void gst(T *p, long long i, const T &v) {
    *gptr(p, i) = v;
}

// Tables that an earlier kernel wrote and the current one only reads (layer
// records, touch sets, stencils), read through the constant address space: a
// wave-uniform index then becomes a scalar load even where the kernel also
// stores through other pointers.
#define DH_CONST __attribute__((address_space(4)))

template <class T>
__device__ __forceinline__ T cld(const T *p, long long i) {
    static_assert(sizeof(T) % 4 == 0, "cld reads whole dwords");
    const DH_CONST int *q = (const DH_CONST int *)(p + i);
    union {
        T t;
        int w[sizeof(T) / 4];
    } u;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) u.w[k] = q[k];
    return u.t;
}

__device__ __forceinline__ const BeamCtx &load_ctx(const BeamCtx *__restrict__ ctx, int b) { return ctx[b]; }

// Owner of an item in a table of ascending start offsets off[0..n) with
// off[0] = 0: the largest b with off[b] <= item.
__device__ __forceinline__ int find_owner(const long long *__restrict__ off, int n, long long item) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= item) lo = mid; else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Piecewise-linear calibration curve (HU -> RLSP, plan_impt.py:158-168; HU ->
// density, plan_imrt.py:114-124), evaluated as np.interp in float64.
constexpr int MAX_KNOTS = 32;
struct Knots {
    double hu[MAX_KNOTS];
    double val[MAX_KNOTS];   // RLSP or density at the knots
    int n;
};

__device__ __forceinline__ double interp_knots(const Knots &kn, double x) {
    // np.interp: clamp outside [hu[0], hu[n-1]]; otherwise linear on the
    // bracketing interval, slope * (x - hu[j]) + val[j].
    if (!(x == x)) return x;  // NaN stays NaN
    if (x <= kn.hu[0]) return kn.val[0];
    if (x >= kn.hu[kn.n - 1]) return kn.val[kn.n - 1];
    int j = 0;
    for (int t = 1; t < kn.n - 1; ++t) j = (x >= kn.hu[t]) ? t : j;
    double slope = __ddiv_rn(kn.val[j + 1] - kn.val[j], kn.hu[j + 1] - kn.hu[j]);
    return __dadd_rn(__dmul_rn(slope, x - kn.hu[j]), kn.val[j]);
}

}  // namespace dh
