// HU -> RLSP and water-equivalent-thickness kernels (gfx950).
//
//  k_rlsp_from_hu  fused np.interp HU->RLSP into the quad volume layout
//  k_rlsp_pad      plain RLSP -> quad layout (drop-in path: RLSP comes from the host)
//  k_wet_raymarch  per-voxel 1 mm march toward the source, manual trilinear
//  k_wet_smooth    6-direction lateral mean of the raw WET
//
// Quad layout: element (i, j, k), i < nx, j < ny, k <= nz, holds the bilinear
// form of the cell face at slice k, with a = rho(i,j,k), b = rho(i+1,j,k),
// c = rho(i,j+1,k), d = rho(i+1,j+1,k) and rho = 0 past the high end of every
// axis (i = nx, j = ny, k = nz):
//     (a, b - a, c - a, (d - c) - (b - a))
// so that rho(i + x, j + y, k) = a + x (b - a) + y ((c - a) + x (...)) is three
// fmas.  A trilinear sample with continuous index c in [0, n) needs two
// 16-byte loads (slices k0 and k0 + 1) and no per-corner bounds tests.  The
// ray-march is bound by the address and data rate of its gathers (one address
// per lane per load) together with its VALU work, so two 16-byte loads per
// sample instead of four 8-byte ones halve the former and the bilinear form
// trims the latter (8 instead of 14 operations per sample).
//
// Block table (after the quads): one float2 per 4x4x4 block of trilinear cells
// (cells with floor(c) in [4b, 4b + 3] per axis).  x: the common value when
// every corner those cells touch (voxels 4b .. 4b + 4, zero past the high end)
// holds the same rho, NaN otherwise.  y: for a uniform block, its run radius r
// (0 .. RUN_RMAX, as a float): every block within Chebyshev distance r of it
// that lies in the grid is uniform with the same value, so all cells of that
// cube of (2r + 1)^3 blocks have that value.  Every sample whose cell lies in
// the cube is exactly the value, so the march adds the run of samples that
// stays in the cube at once (empty-space skipping; air around the patient is
// one value because the HU table clamps at -1000 HU).  The table is followed by
// one int per block of scratch for the radius passes.
#include <algorithm>
#include "dh_common.h"

namespace dh {

// Blocks of 4x4x4 cells per axis and the table's position after the quads.
__host__ __device__ inline long long quad_count(int nx, int ny, int nz) { return (long long)nx * ny * (nz + 1); }
__host__ __device__ inline long long block_count(int nx, int ny, int nz) {
    return (long long)((nx + 3) / 4) * ((ny + 3) / 4) * ((nz + 3) / 4);
}

// Largest run radius (in blocks) the table records.
constexpr int RUN_RMAX = 15;

// One thread per block of cells: compare every corner the block's cells touch
// (voxels 4b .. min(4b + 4, n) per axis; rho(i, j, k) is the first component
// of quad (i, j, k), zero at i = nx or j = ny, and quad (i, j, nz) is zero)
// with the first one.  The radius is filled in by k_block_radius.
__global__ __launch_bounds__(256) void k_rlsp_blocks(const float4 *__restrict__ q, float2 *__restrict__ table, Grid g,
                                                    long long nblk) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nblk) return;
    const int nbx = (g.nx + 3) / 4, nby = (g.ny + 3) / 4;
    const int bx = (int)(t % nbx);
    const long long r = t / nbx;
    const int by = (int)(r % nby), bz = (int)(r / nby);
    const int i0 = 4 * bx, i1 = min(4 * bx + 4, g.nx);
    const int j0 = 4 * by, j1 = min(4 * by + 4, g.ny);
    const int k0 = 4 * bz, k1 = min(4 * bz + 4, g.nz);
    const float v = q[i0 + (long long)g.nx * (j0 + (long long)g.ny * k0)].x;
    bool same = v == v;
    for (int k = k0; k <= k1 && same; ++k)
        for (int j = j0; j <= j1 && same; ++j)
            for (int i = i0; i <= i1 && same; ++i) {
                const float e = (i < g.nx && j < g.ny) ? q[i + (long long)g.nx * (j + (long long)g.ny * k)].x : 0.0f;
                same = e == v;
            }
    table[t] = make_float2(same ? v : __int_as_float(0x7fc00000), 0.0f);
}

// Run radii by three separable passes of the Chebyshev distance to the nearest
// "obstacle" of a block B -- a block that is not uniform with B's value (a
// non-uniform block is an obstacle to every block, itself included; blocks
// outside the grid are none).  Pass a (x, y, z in turn) gives
//     f_a(B) = min over |d| <= RUN_RMAX + 1 of max(|d|, same(B + d e_a) ? f_{a-1}(B + d e_a) : 0),
// with f_{-1} = RUN_RMAX + 1 everywhere: f_x is the distance along x, f_y the
// distance in the (x, y) plane, f_z the 3-D one (capped at RUN_RMAX + 1).  A
// block whose value equals B's has the same obstacles as B, and one whose value
// differs is itself an obstacle, which is why the term is 0 there.  Radius =
// f_z - 1.  Blocks are indexed bx + nbx * (by + nby * bz).
__global__ __launch_bounds__(256) void k_block_radius(const float2 *__restrict__ table, const int *__restrict__ fin,
                                                     int *__restrict__ fout, int nbx, int nby, int nbz, int axis) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= nbx * nby * nbz) return;
    const int b[3] = {t % nbx, (t / nbx) % nby, t / (nbx * nby)};
    const int n = axis == 0 ? nbx : (axis == 1 ? nby : nbz);
    const int stride = axis == 0 ? 1 : (axis == 1 ? nbx : nbx * nby);
    const float v = table[t].x;
    int best = RUN_RMAX + 1;
    for (int d = 0; d <= RUN_RMAX + 1 && d < best; ++d) {
        for (int sgn = -1; sgn <= 1; sgn += 2) {
            if (d == 0 && sgn > 0) continue;
            const int p = b[axis] + sgn * d;
            if (p < 0 || p >= n) continue;
            const int u = t + sgn * d * stride;
            const int f = table[u].x == v ? (fin ? fin[u] : RUN_RMAX + 1) : 0;  // NaN never equals
            best = min(best, max(d, f));
        }
    }
    fout[t] = best;
}

// Radius = f_z - 1 into the table (0 for non-uniform blocks).
__global__ __launch_bounds__(256) void k_block_radius_set(float2 *__restrict__ table, const int *__restrict__ f,
                                                         long long nblk) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nblk) return;
    table[t].y = (float)max(f[t] - 1, 0);
}

int block_radius_passes(float2 *table, int *s0, int *s1, int nbx, int nby, int nbz, hipStream_t st) {
    const long long nblk = (long long)nbx * nby * nbz;
    const dim3 grid((unsigned)((nblk + 255) / 256));
    int *const out[3] = {s0, s1, s0};
    const int *const in[3] = {nullptr, s0, s1};
    for (int axis = 0; axis < 3; ++axis) {
        hipLaunchKernelGGL(k_block_radius, grid, dim3(256), 0, st, (const float2 *)table, in[axis], out[axis], nbx,
                           nby, nbz, axis);
        const int rc = check_launch("k_block_radius");
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_block_radius_set, grid, dim3(256), 0, st, table, (const int *)s0, nblk);
    return check_launch("k_block_radius_set");
}

// ---------------------------------------------------------------------------
// HU -> RLSP (IMPTDoseGrid.RLSPFromHU, plan_impt.py:158-168)
// (Knots, interp_knots: dh_common.h)

__device__ __forceinline__ void quad_coords(const Grid &g, long long t, int &i, int &j, int &k) {
    i = (int)(t % g.nx);
    const long long r = t / g.nx;
    j = (int)(r % g.ny);
    k = (int)(r / g.ny);
}

template <typename F>
__device__ __forceinline__ float4 make_quad(const Grid &g, int i, int j, int k, F rho) {
    // rho(i, j, k) for in-range indices, 0 past the high end
    const bool xi = i + 1 < g.nx, yj = j + 1 < g.ny, zk = k < g.nz;
    const float a = zk ? rho(i, j, k) : 0.0f;
    const float b = (zk && xi) ? rho(i + 1, j, k) : 0.0f;
    const float c = (zk && yj) ? rho(i, j + 1, k) : 0.0f;
    const float d = (zk && xi && yj) ? rho(i + 1, j + 1, k) : 0.0f;
    return make_float4(a, b - a, c - a, (d - c) - (b - a));
}

__global__ __launch_bounds__(256) void k_rlsp_from_hu(const float *__restrict__ hu, float4 *__restrict__ out,
                                                     Grid g, Knots kn, long long nquad) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nquad) return;
    int i, j, k;
    quad_coords(g, t, i, j, k);
    out[t] = make_quad(g, i, j, k, [&](int a, int b, int c) {
        return (float)interp_knots(kn, (double)hu[a + (long long)g.nx * (b + (long long)g.ny * c)]);
    });
}

__global__ __launch_bounds__(256) void k_rlsp_pad(const float *__restrict__ in, float4 *__restrict__ out, Grid g,
                                                 long long nquad) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nquad) return;
    int i, j, k;
    quad_coords(g, t, i, j, k);
    out[t] = make_quad(g, i, j, k,
                       [&](int a, int b, int c) { return in[a + (long long)g.nx * (b + (long long)g.ny * c)]; });
}

// ---------------------------------------------------------------------------
// Ray-march (rayTraceKernel, CudaClasses.cu:180-228), with the sampling and
// bounds of impt_jax.py:254-281: sample n sits at p = v + n*u (1 mm steps,
// u = unit vector to the source), continuous index c = (p + iso) / s; a sample
// counts only if every c_a is in [0, n_a); its RLSP is the zero-padded
// trilinear interpolant.  WET(cm) = -0.05 + 0.1 * sum max(rlsp, 0).
//
// The samples inside the grid form one run n = 0 .. n_exit-1 (the grid is
// convex and sample 0 is the voxel centre).  The run length is found
// analytically; all but the last couple of samples are provably inside by a
// margin far above float32 rounding and are read without tests, the tail is
// tested sample by sample exactly as the per-step mask does.
__device__ __forceinline__ float blend_quads(float4 q0, float4 q1, float ax, float ay, float az) {
    const float y0 = fmaf(ay, fmaf(ax, q0.w, q0.z), fmaf(ax, q0.y, q0.x));
    const float y1 = fmaf(ay, fmaf(ax, q1.w, q1.z), fmaf(ax, q1.y, q1.x));
    return fmaf(az, y1 - y0, y0);
}


// Tail samples (plain loads; c inside the grid).
__device__ __forceinline__ float trilinear_quad(const float4 *__restrict__ qv, int nx, int nxy, float cx, float cy,
                                                float cz) {
    const float fx = floorf(cx), fy = floorf(cy), fz = floorf(cz);
    const long long base = (long long)(int)fx + (long long)nx * (int)fy + (long long)nxy * (int)fz;
    return blend_quads(gld(qv, base), gld(qv, base + nxy), cx - fx, cy - fy, cz - fz);
}

// Steps until the ray leaves [0, n) along one axis (real-valued).
__device__ __forceinline__ float axis_exit(float c0, float du, int n) {
    if (du > 0.0f) return ((float)n - c0) / du;
    if (du < 0.0f) return c0 / (-du);
    return 3.0e38f;
}

// Bulk samples: position by one fma per axis from the voxel-centre index, the
// two quads fetched with buffer loads (32-bit offsets, slice stride as a scalar
// offset; out-of-range reads give 0).
__device__ __forceinline__ float trilinear_buf(__amdgpu_buffer_rsrc_t rs, float cx, float cy, float cz, int nx,
                                               int nxy, int so_z) {
    // c >= 0 here, so truncation is floor
    const int ix = (int)cx, iy = (int)cy, iz = (int)cz;
    const float ax = __builtin_amdgcn_fractf(cx), ay = __builtin_amdgcn_fractf(cy), az = __builtin_amdgcn_fractf(cz);
    const unsigned off = ((unsigned)ix + __umul24((unsigned)iy, (unsigned)nx) + __umul24((unsigned)iz, (unsigned)nxy))
                         << 4;
    const auto r0 = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, 0, 0);
    const auto r1 = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, so_z, 0);
    const float4 q0 = make_float4(__uint_as_float(r0[0]), __uint_as_float(r0[1]), __uint_as_float(r0[2]),
                                  __uint_as_float(r0[3]));
    const float4 q1 = make_float4(__uint_as_float(r1[0]), __uint_as_float(r1[1]), __uint_as_float(r1[2]),
                                  __uint_as_float(r1[3]));
    return blend_quads(q0, q1, ax, ay, az);
}

// The same sample with the cell's two quads kept in q0, q1 between calls: a
// sample in the cell of the lane's previous one (byte offset off_prev) reads
// no memory.  Same values, same arithmetic as trilinear_buf.
__device__ __forceinline__ float trilinear_buf_cached(__amdgpu_buffer_rsrc_t rs, float cx, float cy, float cz, int nx,
                                                      int nxy, int so_z, unsigned &off_prev, float4 &q0,
                                                      float4 &q1) {
    const int ix = (int)cx, iy = (int)cy, iz = (int)cz;  // c >= 0: truncation is floor
    const float ax = __builtin_amdgcn_fractf(cx), ay = __builtin_amdgcn_fractf(cy), az = __builtin_amdgcn_fractf(cz);
    const unsigned off = ((unsigned)ix + __umul24((unsigned)iy, (unsigned)nx) + __umul24((unsigned)iz, (unsigned)nxy))
                         << 4;
    if (off != off_prev) {
        const auto r0 = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, 0, 0);
        const auto r1 = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, so_z, 0);
        q0 = make_float4(__uint_as_float(r0[0]), __uint_as_float(r0[1]), __uint_as_float(r0[2]),
                         __uint_as_float(r0[3]));
        q1 = make_float4(__uint_as_float(r1[0]), __uint_as_float(r1[1]), __uint_as_float(r1[2]),
                         __uint_as_float(r1[3]));
        off_prev = off;
    }
    return blend_quads(q0, q1, ax, ay, az);
}

// Samples from sample n on that stay in the 4x4x4 cell block of position c,
// minus one step of margin against the rounding of the positions (negative or
// zero: no run).  Per axis the ray meets the block's far face (the one it moves
// toward) at (face - c) / du steps; lim[a] = face offset 4 or 0, rds = 1 / du
// (0 when du = 0), add = 0 (3e38 when du = 0).
struct BlockExit {
    int offx, offy, offz;
    float sgx, sgy, sgz;
    float rdx, rdy, rdz, addx, addy, addz;
};

__device__ __forceinline__ BlockExit block_exit_setup(float dux, float duy, float duz) {
    BlockExit e;
    e.offx = dux > 0.0f ? 4 : 0;
    e.offy = duy > 0.0f ? 4 : 0;
    e.offz = duz > 0.0f ? 4 : 0;
    e.sgx = dux > 0.0f ? 1.0f : -1.0f;
    e.sgy = duy > 0.0f ? 1.0f : -1.0f;
    e.sgz = duz > 0.0f ? 1.0f : -1.0f;
    e.rdx = dux != 0.0f ? 1.0f / dux : 0.0f;
    e.rdy = duy != 0.0f ? 1.0f / duy : 0.0f;
    e.rdz = duz != 0.0f ? 1.0f / duz : 0.0f;
    e.addx = dux != 0.0f ? 0.0f : 3.0e38f;
    e.addy = duy != 0.0f ? 0.0f : 3.0e38f;
    e.addz = duz != 0.0f ? 0.0f : 3.0e38f;
    return e;
}

// Samples from n on that stay in the cube of blocks within Chebyshev distance
// r of block (bx, by, bz) (cells [4 (b - r), 4 (b + r + 1)) per axis), minus the
// step of margin: the far face along an axis is 4 b + off + 4 r sg with
// sg = +1 moving up, -1 moving down.
__device__ __forceinline__ int cube_run(const BlockExit &e, float cx, float cy, float cz, int bx, int by, int bz,
                                        float r4) {
    const float tx = fmaf(fmaf(r4, e.sgx, (float)((bx << 2) + e.offx)) - cx, e.rdx, e.addx);
    const float ty = fmaf(fmaf(r4, e.sgy, (float)((by << 2) + e.offy)) - cy, e.rdy, e.addy);
    const float tz = fmaf(fmaf(r4, e.sgz, (float)((bz << 2) + e.offz)) - cz, e.rdz, e.addz);
    return (int)fminf(fminf(fminf(tx, ty), tz), 1.0e9f) - 1;
}

// Raw WET of voxel (i, j, k) of a beam; the caller guarantees (i, j, k) inside.
__device__ __forceinline__ float raymarch_voxel(const float4 *__restrict__ qv, const Grid &g, const Geom &b, int i,
                                                int j, int k) {
    const float3 v = voxel_xyz(g, b, i, j, k);
    const float dx = b.src[0] - v.x, dy = b.src[1] - v.y, dz = b.src[2] - v.z;
    // correctly rounded sqrt and divisions: an axis-aligned ray gets an exactly
    // axis-aligned u, so its samples land exactly on grid planes as in float64
    const float len = sqrtf(dx * dx + dy * dy + dz * dz);
    const float ux = dx / len, uy = dy / len, uz = dz / len;

    const float c0x = (v.x + b.iso[0]) * g.inv_s, c0y = (v.y + b.iso[1]) * g.inv_s, c0z = (v.z + b.iso[2]) * g.inv_s;
    const float dux = ux * g.inv_s, duy = uy * g.inv_s, duz = uz * g.inv_s;
    float t_exit = fminf(axis_exit(c0x, dux, g.nx), fminf(axis_exit(c0y, duy, g.ny), axis_exit(c0z, duz, g.nz)));
    t_exit = fminf(t_exit, 1.0e7f);
    const int n_safe = max(0, (int)floorf(t_exit) - 1);
    const int n_end = (int)floorf(t_exit) + 3;

    const int nxy = g.nx * g.ny;
    // one buffer over the quads and the block table after them (byte offsets
    // below 2^32: grid_ok)
    const unsigned tab_off = (unsigned)nxy * (unsigned)(g.nz + 1) * 16u;
    const unsigned nbx = (unsigned)(g.nx + 3) >> 2, nby = (unsigned)(g.ny + 3) >> 2;
    const unsigned nbz = (unsigned)(g.nz + 3) >> 2;
    const unsigned nbxy = nbx * nby;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)qv, (short)0, (int)(tab_off + nbx * nby * nbz * 8u), 0x00020000);
    const int so_z = nxy * 16;
    const BlockExit ex = block_exit_setup(dux, duy, duz);
    float acc0 = -0.05f, acc1 = 0.0f;
    int n = 0;
    // quads of the cell of the lane's last sample (none yet: an offset no cell has)
    unsigned off_prev = 0xffffffffu;
    float4 qc0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), qc1 = qc0;
    // Per step: the table entry of sample n's block; in a uniform block the run
    // of samples n .. n + T that stay in its cube of same-valued blocks adds
    // T + 1 times its value at once, otherwise samples n and n + 1 are read (no
    // quad loads for skipped runs; a sample in its lane's previous cell reuses
    // the quads read for it: about half the samples at 1 mm steps in 2 mm cells).
    while (n < n_safe) {
        const float f0 = (float)n;
        const float cx = fmaf(dux, f0, c0x), cy = fmaf(duy, f0, c0y), cz = fmaf(duz, f0, c0z);
        const int bx = (int)cx >> 2, by = (int)cy >> 2, bz = (int)cz >> 2;  // c >= 0: truncation is floor
        // two 24-bit products (the nested form compiled to a 64-bit multiply-add)
        const unsigned bidx =
            (unsigned)bx + (unsigned)__umul24(nbx, (unsigned)by) + (unsigned)__umul24(nbxy, (unsigned)bz);
        const auto te = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)(tab_off + (bidx << 3)), 0, 0);
        const float u = __uint_as_float(te[0]);
        if (u == u) {
            const float r4 = 4.0f * __uint_as_float(te[1]);
            const int T = min(cube_run(ex, cx, cy, cz, bx, by, bz, r4), n_safe - 1 - n);
            if (T >= 1) {
                acc1 = fmaf(fmaxf(u, 0.0f) * (float)(T + 1), 0.1f, acc1);
                n += T + 1;
                continue;
            }
        }
        const float v0 = trilinear_buf_cached(rs, cx, cy, cz, g.nx, nxy, so_z, off_prev, qc0, qc1);
        acc0 = fmaf(fmaxf(v0, 0.0f), 0.1f, acc0);
        // an empty statement that orders the first sample's sum (and memory) before the second sample: the
        // compiler otherwise sinks that sum below the second sample's conditional loads and keeps two copies
        // of the cached quads (16 register moves per pair); it emits no instruction
        asm volatile("" : "+v"(acc0)::"memory");
        if (n + 1 < n_safe) {
            const float f1 = (float)(n + 1);
            const float v1 = trilinear_buf_cached(rs, fmaf(dux, f1, c0x), fmaf(duy, f1, c0y), fmaf(duz, f1, c0z),
                                                  g.nx, nxy, so_z, off_prev, qc0, qc1);
            acc1 = fmaf(fmaxf(v1, 0.0f), 0.1f, acc1);
            n += 2;
        } else {
            n += 1;
        }
    }
    float acc = acc0 + acc1;
    // tail: the reference's per-sample position arithmetic and bounds test
    for (; n < n_end; ++n) {
        const float fn = (float)n;
        const float cx = (fmaf(ux, fn, v.x) + b.iso[0]) * g.inv_s;
        const float cy = (fmaf(uy, fn, v.y) + b.iso[1]) * g.inv_s;
        const float cz = (fmaf(uz, fn, v.z) + b.iso[2]) * g.inv_s;
        const bool inside = cx >= 0.0f && cx < (float)g.nx && cy >= 0.0f && cy < (float)g.ny && cz >= 0.0f &&
                            cz < (float)g.nz;
        if (inside) acc = fmaf(fmaxf(trilinear_quad(qv, g.nx, nxy, cx, cy, cz), 0.0f), 0.1f, acc);
    }
    return acc;
}

// Work mapping.  A workgroup takes four consecutive bricks; when they lie in
// one row of the beam's brick grid (bx0 .. bx0 + 3), its four waves tile their
// 16 x 4 x 4 voxels with patches of one shape (kRayShape,
// dh_batch_tune.ray_shape), e.g. 16 x 4 x 1: wave w marches layer z = 4 bk + w
// of all four bricks, lane (x, y) = (lane & 15, lane >> 4).  The shape decides
// how many cache lines each gather of the march touches (wide x-runs touch
// fewer) and how often a wave's lanes reuse their last cell together (the
// rays of a compact patch enter new cells at nearly the same steps).  On the
// head batch with the cell cache the 4 x 4 x 4 brick is the fastest (the
// default) and 16 x 4 x 1 the slowest, 6% apart.  Other quadruples (a row end,
// a beam end) go one wave per brick.  Every voxel's value is the same
// whatever the shape.
__device__ __forceinline__ void raymarch_voxel_at(const BeamCtx &c, int i, int j, int k) {
    if (i >= c.g.nx || j >= c.g.ny || k >= c.g.nz) return;
    gst(c.wet_raw, i + (long long)c.g.nx * (j + (long long)c.g.ny * k), raymarch_voxel(c.rlsp_q, c.g, c.b, i, j, k));
}

__device__ __forceinline__ void raymarch_brick(const BeamCtx &c, long long brick, int lane) {
    int bi, bj, bk;
    brick_coords(c.bg, brick, bi, bj, bk);
    raymarch_voxel_at(c, bi * BRICK + (lane & 3), bj * BRICK + ((lane >> 2) & 3), bk * BRICK + (lane >> 4));
}

// Ray-march patch shapes (log2 of the x, y, z extents of a wave's 64 voxels):
// the four waves of a workgroup tile the 16 x 4 x 4 voxels of bricks b0 .. b0 + 3
// of one row with patches of this shape.
constexpr int RAY_SHAPES = 6;
__constant__ const unsigned char kRayShape[RAY_SHAPES][3] = {
    {2, 2, 2},  // 4 x 4 x 4 (one brick; the default)
    {3, 2, 1},  // 8 x 4 x 2
    {4, 2, 0},  // 16 x 4 x 1
    {4, 1, 1},  // 16 x 2 x 2
    {4, 0, 2},  // 16 x 1 x 4
    {3, 1, 2},  // 8 x 2 x 4
};

// Bricks b0 .. b0 + 3 of one row: wave w, lane -> voxel of its patch of the
// given shape.  Lanes of bricks the reach pass does not need stay idle.
__device__ __forceinline__ void raymarch_row(const BeamCtx &c, long long b0, int w, int lane, int use_need,
                                             int shape) {
    int bi, bj, bk;
    brick_coords(c.bg, b0, bi, bj, bk);
    const int lx = kRayShape[shape][0], ly = kRayShape[shape][1], lz = kRayShape[shape][2];
    // waves per axis: 16 >> lx along x, 4 >> ly along y (the rest along z)
    const int wlx = 4 - lx, wly = 2 - ly;
    const int wx = w & ((1 << wlx) - 1), wy = (w >> wlx) & ((1 << wly) - 1), wz = w >> (wlx + wly);
    const int x = (wx << lx) + (lane & ((1 << lx) - 1));
    const int y = (wy << ly) + ((lane >> lx) & ((1 << ly) - 1));
    const int z = (wz << lz) + (lane >> (lx + ly));
    if (use_need && !(gld(c.need, b0 + (x >> 2)) & 2u)) return;
    raymarch_voxel_at(c, bi * BRICK + x, bj * BRICK + y, bk * BRICK + z);
}

__device__ __forceinline__ bool in_one_row(const BrickGrid &bg, long long b0, long long end) {
    return b0 + 3 < end && (int)((unsigned)b0 % (unsigned)bg.bx) + 3 < bg.bx;
}

// Single beam: workgroup q takes bricks 4q .. 4q + 3.
__global__ __launch_bounds__(256) void k_wet_raymarch(BeamCtx c, unsigned nblocks) {
    const unsigned blk = xcd_remap(blockIdx.x, nblocks);
    const long long b0 = (long long)blk * 4, nbr = c.bg.count();
    const int w = wave_in_block(), lane = threadIdx.x & 63;
    if (in_one_row(c.bg, b0, nbr)) {
        raymarch_row(c, b0, w, lane, 0, 0);  // the default shape
    } else if (b0 + w < nbr) {
        raymarch_brick(c, b0 + w, lane);
    }
}

// Batched: bricks of every beam in one launch; beam b owns bricks
// [off[b], off[b + 1]) of the item space.  A workgroup takes items
// q .. q + 3 (q a multiple of 4); the row mapping needs all four in one beam
// and one row.  Workgroups step over items by `stride` (4 * gridDim.x for the
// persistent grid).
__device__ __forceinline__ void raymarch_item(const BeamCtx *__restrict__ ctx, const long long *__restrict__ off,
                                              int n_beams, long long item, int use_need) {
    const int bm = find_owner(off, n_beams, item);
    const BeamCtx &c = load_ctx(ctx, bm);
    const long long brick = item - off[bm];
    // reach-limited pass: bricks no smoothed voxel reads stay unwritten
    if (use_need && !(gld(c.need, brick) & 2u)) return;
    raymarch_brick(c, brick, threadIdx.x & 63);
}

__device__ __forceinline__ void raymarch_quad(const BeamCtx *__restrict__ ctx, const long long *__restrict__ off,
                                              int n_beams, long long q, long long total, int use_need,
                                              int shape) {
    const int w = wave_in_block();
    const int bm = find_owner(off, n_beams, q);
    // off holds the first item of each beam; the last beam ends at total
    const long long end = bm + 1 < n_beams ? off[bm + 1] : total;
    const BeamCtx &c = load_ctx(ctx, bm);
    if (in_one_row(c.bg, q - off[bm], end - off[bm])) {
        raymarch_row(c, q - off[bm], w, threadIdx.x & 63, use_need, shape);
    } else if (q + w < total) {
        raymarch_item(ctx, off, n_beams, q + w, use_need);
    }
}

// The persistent form (PERSIST) is its own instantiation so that the one-shot
// form keeps the registers of the plain kernel.
template <bool PERSIST>
__global__ __launch_bounds__(256) void k_wet_raymarch_batch(const BeamCtx *__restrict__ ctx,
                                                           const long long *__restrict__ off, int n_beams,
                                                           long long base, long long total, int use_need,
                                                           long long stride, int shape) {
    // one-shot grid: chunks of 256 workgroups per XCD (measured best of 4 .. 4096
    // on the batched head workload, 5-8% over the dispatcher's round-robin)
    long long q = base + (long long)(PERSIST ? blockIdx.x : xcd_chunk<256>(blockIdx.x, gridDim.x)) * 4;
    if (PERSIST) {
        for (; q < total; q += stride) raymarch_quad(ctx, off, n_beams, q, total, use_need, shape);
    } else if (q < total) {
        raymarch_quad(ctx, off, n_beams, q, total, use_need, shape);
    }
}

// ---------------------------------------------------------------------------
// Lateral WET smoothing (smoothRayKernel, IMPTClasses.cu:139-202;
// _smooth_wet_kernel, impt_jax.py:317-418).
//
// A neighbour of voxel (i, j, k) is the voxel nearest to the voxel centre
// moved by dr mm along beam's-eye-view direction a*60 deg.  The move is the
// same for every voxel: in voxel units it is D = R^T (dr cos, dr sin, 0) / s
// (R = image -> head rotation, orthogonal), so the neighbour is
// round_half_even((i, j, k) + D) per axis.  The host tabulates D for the 54
// (direction, dr) pairs in float64 as an integer base offset plus a tie flag
// for components whose fractional part is a half (within 1e-6): those round to
// the even index, which depends on the voxel's parity.  This is the
// real-number value of the reference's rotate / round sequence; float32 (or
// float64) evaluation of the rotations would break exact half-voxel ties at
// random, and such ties are common (gantry angles that are multiples of 30 deg
// put the 60-deg directions on rational offsets).
__device__ __forceinline__ int stencil_axis(int i, int base, bool tie) {
    const int n = i + base;
    return tie ? n + (n & 1) : n;
}

// Stencil entry e, from a kernel-argument copy or from a table in global memory.
struct StencilEntry {
    int dx, dy, dz;
    unsigned tie;
    float dr;
};

__device__ __forceinline__ StencilEntry stencil_entry(const Stencil &st, int e) {
    return StencilEntry{st.dx[e], st.dy[e], st.dz[e], st.tie[e], (float)st.dr[e]};
}

// The device table is read-only while kernels run: read it through the
// constant address space in whole dwords, so the (wave-uniform) entries come
// in through scalar loads and scalar bit-field extracts rather than one
// vector byte load per field.

__device__ __forceinline__ int stencil_byte(const Stencil *st, int off, bool is_signed) {
    const int w = ((const DH_CONST int *)st)[off >> 2];
    const int sh = (off & 3) * 8;
    return is_signed ? (w << (24 - sh)) >> 24 : (w >> sh) & 0xff;
}

__device__ __forceinline__ StencilEntry stencil_entry(const Stencil *st, int e) {
    return StencilEntry{stencil_byte(st, e, true), stencil_byte(st, N_STENCIL + e, true),
                        stencil_byte(st, 2 * N_STENCIL + e, true), (unsigned)stencil_byte(st, 3 * N_STENCIL + e, false),
                        (float)stencil_byte(st, 4 * N_STENCIL + e, false)};
}

// Indices fit in 32 bits: the grid check keeps nx * ny * (nz + 1) below 2^28.
// Every entry issues its load (a skipped one reads the centre voxel again and
// is not added), so the 54 loads of a voxel need no branches and can be in
// flight together; the sum runs over the kept entries in table order.  (A
// separate interior path with wave-uniform linear offsets was measured slower:
// 109 VGPRs.)
template <class ST>
__device__ __forceinline__ float smooth_voxel(const float *__restrict__ raw, const Grid &g, const ST &st, long long t) {
    const int ti = (int)t;
    const int i = ti % g.nx;
    const int r = ti / g.nx;
    const int j = r % g.ny, k = r / g.ny;
    const float wc = gld(raw, ti);
    const float max_dr = fminf(wc * 10.0f, 10.0f);
    float sum = wc;
    int cnt = 1;
    // entries are ordered direction-major, dr = 1..9 within a direction: an
    // entry is kept while dr < max_dr, so a direction stops at its first miss
#pragma unroll
    for (int e = 0; e < N_STENCIL; ++e) {
        const StencilEntry se = stencil_entry(st, e);
        const int ni = stencil_axis(i, se.dx, se.tie & 1u);
        const int nj = stencil_axis(j, se.dy, se.tie & 2u);
        const int nk = stencil_axis(k, se.dz, se.tie & 4u);
        const bool ok = se.dr < max_dr && (unsigned)ni < (unsigned)g.nx && (unsigned)nj < (unsigned)g.ny &&
                        (unsigned)nk < (unsigned)g.nz;
        const float v = gld(raw, ok ? ni + g.nx * (nj + g.ny * nk) : ti);
        sum = ok ? sum + v : sum;
        cnt += ok ? 1 : 0;
    }
    return sum / (float)cnt;
}

__global__ __launch_bounds__(256) void k_wet_smooth(const float *__restrict__ raw, float *__restrict__ out, Grid g,
                                                   Stencil st, long long nvox) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nvox) return;
    out[t] = smooth_voxel(raw, g, st, t);
}

// Batched: beam b owns blocks [boff[b], boff[b + 1]) of 256 voxels each.
__global__ __launch_bounds__(256) void k_wet_smooth_batch(const BeamCtx *__restrict__ ctx,
                                                         const long long *__restrict__ boff, int n_beams,
                                                         long long base, long long total_blocks, int use_need) {
    const long long blk = base + xcd_chunk<256>(blockIdx.x, gridDim.x);
    if (blk >= total_blocks) return;
    const int bm = find_owner(boff, n_beams, blk);
    const BeamCtx &c = load_ctx(ctx, bm);
    const long long t = (blk - boff[bm]) * 256 + threadIdx.x;
    const long long nvox = (long long)c.g.nx * c.g.ny * c.g.nz;
    if (t >= nvox) return;
    if (use_need) {
        // voxel indices fit in 32 bits (the grid check keeps nx*ny*(nz+1) below 2^28)
        const unsigned tu = (unsigned)t, nx = (unsigned)c.g.nx, ny = (unsigned)c.g.ny;
        const unsigned r = tu / nx;
        const int i = (int)(tu - r * nx), k = (int)(r / ny), j = (int)(r - (unsigned)k * ny);
        const long long brick = (i >> 2) + (long long)c.bg.bx * ((j >> 2) + (long long)c.bg.by * (k >> 2));
        if (!(gld(c.need, brick) & 1u)) {
            gst(c.wet, t, kWetOutside);
            return;
        }
    }
    gst(c.wet, t, smooth_voxel(c.wet_raw, c.g, c.stencil, t));
}

// Box index of stencil entry e for the voxel (i, j, k) at box index cidx:
// the entry's linear offset (lane e of fast_e, with its tie bits) plus, for a
// half-voxel tie, the neighbour's parity (stencil_axis).
__device__ __forceinline__ int tile_box_index(int fast_e, int gen_e, int e, int cidx, int i, int j, int k, int BX,
                                              int sxy) {
    const int f = __builtin_amdgcn_readlane(fast_e, e);
    int idx = cidx + (f >> 3);
    if (f & 7) {  // wave-uniform
        const int g = __builtin_amdgcn_readlane(gen_e, e);
        if (f & 1) idx += (i + (g & 127) - 64) & 1;
        if (f & 2) idx += BX * ((j + ((g >> 7) & 127) - 64) & 1);
        if (f & 4) idx += sxy * ((k + ((g >> 14) & 127) - 64) & 1);
    }
    return idx;
}

// Tiled batched smoothing: one workgroup per 32 x 4 x 8 tile of a beam.  The
// workgroup first copies the tile and the stencil's halo (a box of the raw
// WET, coalesced rows) into LDS, then each thread smooths four voxels with the
// 54 neighbour reads coming from LDS.  Entries, the kept-entry test and the
// summation order are those of smooth_voxel, so the results are the same.
// Tiles whose bricks all lack bit 0 of the need mask only write kWetOutside.
// A wave whose voxels have all their neighbours inside the grid skips the
// bounds tests, and one whose voxels also keep every entry (centre WET above
// 0.9 cm) skips the radius test too.
__global__ __launch_bounds__(256) void k_wet_smooth_tile_batch(const BeamCtx *__restrict__ ctx,
                                                              const long long *__restrict__ toff, int n_beams,
                                                              long long base, long long total, int use_need) {
    extern __shared__ float box[];
    const long long tile = base + xcd_chunk<256>(blockIdx.x, gridDim.x);
    if (tile >= total) return;  // whole workgroup leaves together
    const int bm = find_owner(toff, n_beams, tile);
    const BeamCtx &c = load_ctx(ctx, bm);
    const int nx = c.g.nx, ny = c.g.ny, nz = c.g.nz;
    // tile coordinates (tile counts fit in 32 bits)
    const unsigned t = (unsigned)(tile - toff[bm]);
    const unsigned ntx = (unsigned)(nx + STX - 1) / STX, nty = (unsigned)(ny + STY - 1) / STY;
    const unsigned tr = t / ntx;
    const int x0 = (int)(t - tr * ntx) * STX, y0 = (int)(tr % nty) * STY, z0 = (int)(tr / nty) * STZ;
    const int tid = threadIdx.x;

    if (use_need) {
        bool nd = false;
        if (tid < 16) {  // the tile's 8 x 1 x 2 bricks
            const int bi = (x0 >> 2) + (tid & 7), bj = y0 >> 2, bk = (z0 >> 2) + (tid >> 3);
            if (bi < c.bg.bx && bk < c.bg.bz)
                nd = (gld(c.need, bi + (long long)c.bg.bx * (bj + (long long)c.bg.by * bk)) & 1u) != 0;
        }
        if (!__syncthreads_or(nd)) {
            for (int v = tid; v < STX * STY * STZ; v += 256) {
                const int i = x0 + (v & (STX - 1)), j = y0 + ((v / STX) & (STY - 1)), k = z0 + v / (STX * STY);
                if (i < nx && j < ny && k < nz)
                    gst(c.wet, (unsigned)i + (unsigned)nx * ((unsigned)j + (unsigned)ny * (unsigned)k), kWetOutside);
            }
            return;
        }
    }

    // box of the raw WET, by rows along x: each wave takes every fourth row and
    // issues 16 rows' loads before storing them, so they are in flight together
    const int BX = c.sbox[0], BY = c.sbox[1], BZ = c.sbox[2];
    const int gx0 = x0 - c.slo[0], gy0 = y0 - c.slo[1], gz0 = z0 - c.slo[2];
    {
        const int wave = tid >> 6, lane = tid & 63, rows = BY * BZ;
        for (int cx = 0; cx < BX; cx += 64) {  // one pass unless BX > 64 (spacing below 1 mm)
            const int bx = cx + lane, gx = gx0 + bx;
            const bool col = bx < BX && (unsigned)gx < (unsigned)nx;
            int by = wave, bz = 0;  // row = by + BY * bz, advancing by 4 rows
            while (by >= BY) {
                by -= BY;
                ++bz;
            }
            for (int r0 = wave; r0 < rows; r0 += 64) {
                float vals[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int gy = gy0 + by, gz = gz0 + bz;
                    const bool in = col && r0 + 4 * u < rows && (unsigned)gy < (unsigned)ny && (unsigned)gz < (unsigned)nz;
                    // voxel indices fit in 32 bits (the grid check keeps nx*ny*(nz+1) below 2^28); a
                    // position outside the grid loads voxel 0 and stores 0 (such entries are never kept)
                    const unsigned at = in ? (unsigned)gx + (unsigned)nx * ((unsigned)gy + (unsigned)ny * (unsigned)gz) : 0u;
                    const float v = gld(c.wet_raw, at);
                    vals[u] = in ? v : 0.0f;
                    by += 4;
                    while (by >= BY) {
                        by -= BY;
                        ++bz;
                    }
                }
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (bx < BX && r0 + 4 * u < rows) box[bx + BX * (r0 + 4 * u)] = vals[u];
            }
        }
    }
    __syncthreads();

    // the 54 entries live in lane e of two registers (read back with
    // readlane, so the loop below touches no memory but LDS): the entry's
    // linear box offset with its tie bits, and its packed fields
    const int sxy = BX * BY;
    int fast_e = 0, gen_e = 0;
    {
        const int e = tid & 63;
        if (e < N_STENCIL) {
            const signed char *sb = reinterpret_cast<const signed char *>(c.stencil);
            const int dx = gld(sb, e), dy = gld(sb, N_STENCIL + e), dz = gld(sb, 2 * N_STENCIL + e);
            const int tie = (unsigned char)gld(sb, 3 * N_STENCIL + e), dr = (unsigned char)gld(sb, 4 * N_STENCIL + e);
            fast_e = ((dx + BX * dy + sxy * dz) << 3) | tie;
            gen_e = (dx + 64) | (dy + 64) << 7 | (dz + 64) << 14 | tie << 21 | dr << 24;
        }
    }
    // halo above the tile per axis: a voxel at least slo below / shi above
    // the grid's edges has every stencil neighbour inside the grid
    const int shx = BX - STX - c.slo[0], shy = BY - STY - c.slo[1], shz = BZ - STZ - c.slo[2];
#pragma unroll 1
    for (int q = 0; q < STX * STY * STZ / 256; ++q) {
        const int v = tid + 256 * q;
        const int lx = v & (STX - 1), ly = (v / STX) & (STY - 1), lz = v / (STX * STY);
        const int i = x0 + lx, j = y0 + ly, k = z0 + lz;
        bool live = i < nx && j < ny && k < nz;
        const unsigned vidx = live ? (unsigned)i + (unsigned)nx * ((unsigned)j + (unsigned)ny * (unsigned)k) : 0u;
        if (use_need && live &&
            !(gld(c.need, (i >> 2) + (long long)c.bg.bx * ((j >> 2) + (long long)c.bg.by * (k >> 2))) & 1u)) {
            gst(c.wet, vidx, kWetOutside);
            live = false;
        }
        const bool safe = i >= c.slo[0] && i + shx < nx && j >= c.slo[1] && j + shy < ny && k >= c.slo[2] &&
                          k + shz < nz;
        const int cidx = (lx + c.slo[0]) + BX * ((ly + c.slo[1]) + BY * (lz + c.slo[2]));
        const float wc = box[cidx];
        const float max_dr = fminf(wc * 10.0f, 10.0f);
        float sum = wc;
        int cnt = 1;
        // entry e has dr = e % 9 + 1 (direction major) and is kept where dr <
        // max_dr and its neighbour is inside the grid
        if (__all(!live || (safe && 9.0f < max_dr))) {
            // every entry kept: sum in table order, 55 terms
#pragma unroll
            for (int e = 0; e < N_STENCIL; ++e) sum += box[tile_box_index(fast_e, gen_e, e, cidx, i, j, k, BX, sxy)];
            cnt = 1 + N_STENCIL;
        } else {
            // no live lane of the wave keeps an entry with dr > dmax (wave-uniform)
            int kk = live ? (int)ceilf(max_dr) - 1 : 0;
            kk = kk < 0 ? 0 : (kk > 9 ? 9 : kk);
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) kk = max(kk, __shfl_xor(kk, off));
            const int dmax = __builtin_amdgcn_readfirstlane(kk);
            if (__all(!live || safe)) {
                // every neighbour inside the grid: only the radius test
#pragma unroll
                for (int e = 0; e < N_STENCIL; ++e) {
                    if (e % 9 + 1 > dmax) continue;
                    const float val = box[tile_box_index(fast_e, gen_e, e, cidx, i, j, k, BX, sxy)];
                    const bool ok = (float)(e % 9 + 1) < max_dr;
                    sum = ok ? sum + val : sum;
                    cnt += ok ? 1 : 0;
                }
            } else {
#pragma unroll
                for (int e = 0; e < N_STENCIL; ++e) {
                    if (e % 9 + 1 > dmax) continue;
                    const int g = __builtin_amdgcn_readlane(gen_e, e);
                    const int tie = (g >> 21) & 7;
                    const int ni = stencil_axis(i, (g & 127) - 64, tie & 1);
                    const int nj = stencil_axis(j, ((g >> 7) & 127) - 64, tie & 2);
                    const int nk = stencil_axis(k, ((g >> 14) & 127) - 64, tie & 4);
                    const bool ok = (float)(e % 9 + 1) < max_dr && (unsigned)ni < (unsigned)nx &&
                                    (unsigned)nj < (unsigned)ny && (unsigned)nk < (unsigned)nz;
                    const float val = box[ok ? (ni - gx0) + BX * ((nj - gy0) + BY * (nk - gz0)) : cidx];
                    sum = ok ? sum + val : sum;
                    cnt += ok ? 1 : 0;
                }
            }
        }
        if (live) gst(c.wet, vidx, sum / (float)cnt);
    }
}

// Host: tabulate the smoothing stencil of a beam.
void make_stencil(const dh_beam_geom &b, double s, Stencil &st) {
    static const double kc[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
    static const double ks[6] = {0.0, 0.86602540378443864676, 0.86602540378443864676,
                                 0.0, -0.86602540378443864676, -0.86602540378443864676};
    const double sg = b.sin_g, cg = b.cos_g, sc = b.sin_c, cc = b.cos_c;
    for (int a = 0; a < 6; ++a) {
        for (int d = 1; d <= 9; ++d) {
            const int e = a * 9 + (d - 1);
            // head offset (hx, hy, 0) -> image offset, as head_to_image without translation
            const double hx = d * kc[a], hy = d * ks[a];
            const double xz = -hx, yz = 0.0, zz = hy;
            const double xg = xz * cg - yz * sg, yg = xz * sg + yz * cg;
            const double off[3] = {xg * cc + zz * sc, yg, -xg * sc + zz * cc};
            int base[3];
            unsigned tie = 0;
            for (int ax = 0; ax < 3; ++ax) {
                const double q = off[ax] / s;
                const double fl = floor(q), fr = q - fl;
                if (fabs(fr - 0.5) < 1e-6) {
                    base[ax] = (int)fl;
                    tie |= 1u << ax;
                } else {
                    base[ax] = (int)fl + (fr > 0.5 ? 1 : 0);
                }
            }
            st.dx[e] = (signed char)base[0];
            st.dy[e] = (signed char)base[1];
            st.dz[e] = (signed char)base[2];
            st.tie[e] = (unsigned char)tie;
            st.dr[e] = (unsigned char)d;
        }
    }
    st.pad[0] = st.pad[1] = 0;
}

// Brick offsets o such that smoothing a voxel of brick b reads raw WET in
// brick b + o: per stencil entry and axis the neighbour index is l + base
// (+1 on a tie) for l = 0..3 inside the brick, so the brick offset is
// floor((l + base [+ 1]) / 4).  Packed as (ox + 128) | (oy + 128) << 8 |
// (oz + 128) << 16; returns the count (at most 54 * 8).
static int floor4(int n) { return n >= 0 ? n / 4 : -((-n + 3) / 4); }

int smooth_box(const Stencil &st, int slo[3], int sbox[3]) {
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int e = 0; e < N_STENCIL; ++e) {
        const int base[3] = {st.dx[e], st.dy[e], st.dz[e]};
        for (int ax = 0; ax < 3; ++ax) {
            const int t = (st.tie[e] >> ax) & 1;  // a tie may add one voxel
            lo[ax] = std::max(lo[ax], -base[ax]);
            hi[ax] = std::max(hi[ax], base[ax] + t);
        }
    }
    const int tile[3] = {STX, STY, STZ};
    for (int ax = 0; ax < 3; ++ax) {
        slo[ax] = lo[ax];
        sbox[ax] = tile[ax] + lo[ax] + hi[ax];
    }
    return sbox[0] * sbox[1] * sbox[2];
}

int make_touch(const Stencil &st, int *out, int cap) {
    int n = 0;
    for (int e = 0; e < N_STENCIL; ++e) {
        const int base[3] = {st.dx[e], st.dy[e], st.dz[e]};
        int lo[3], hi[3];
        for (int ax = 0; ax < 3; ++ax) {
            const int t = (st.tie[e] >> ax) & 1;
            lo[ax] = floor4(base[ax]);
            hi[ax] = floor4(base[ax] + 3 + t);
        }
        for (int oz = lo[2]; oz <= hi[2]; ++oz)
            for (int oy = lo[1]; oy <= hi[1]; ++oy)
                for (int ox = lo[0]; ox <= hi[0]; ++ox) {
                    const int key = (ox + 128) | (oy + 128) << 8 | (oz + 128) << 16;
                    bool seen = false;
                    for (int q = 0; q < n && !seen; ++q) seen = out[q] == key;
                    if (!seen && n < cap) out[n++] = key;
                }
    }
    // nearest offsets first: the dilation pass stops at the first hit
    auto len2 = [](int key) {
        const int ox = (key & 255) - 128, oy = ((key >> 8) & 255) - 128, oz = ((key >> 16) & 255) - 128;
        return ox * ox + oy * oy + oz * oz;
    };
    std::stable_sort(out, out + n, [&](int a, int b) { return len2(a) < len2(b); });
    return n;
}

bool grid_ok(const dh_grid &g) {
    if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0 || !(g.spacing > 0.0f)) {
        set_error("invalid grid %d x %d x %d, spacing %g", g.nx, g.ny, g.nz, (double)g.spacing);
        return false;
    }
    // 32-bit byte offsets of the ray-march buffer loads: the quads and the block
    // table after them, 16 * nx * ny * (nz + 1) + 8 * blocks < 2^32
    if (16.0 * g.nx * g.ny * (g.nz + 1) + 8.0 * (double)block_count(g.nx, g.ny, g.nz) >= 4294967296.0) {
        set_error("grid %d x %d x %d too large (16*nx*ny*(nz+1) + 8*blocks must stay below 2^32)", g.nx, g.ny, g.nz);
        return false;
    }
    return true;
}

}  // namespace dh

using namespace dh;

// Block table of a quad volume (written right after its quads): values, then
// run radii by three separable passes through the two scratch int arrays that
// follow the table (x -> s0, y -> s1, z -> s0, radius = s0 - 1).
static int launch_blocks(float *rlsp_padded, const dh_grid &g, hipStream_t st) {
    const long long nquad = quad_count(g.nx, g.ny, g.nz), nblk = block_count(g.nx, g.ny, g.nz);
    float2 *table = reinterpret_cast<float2 *>(rlsp_padded + 4 * nquad);
    int *s0 = reinterpret_cast<int *>(rlsp_padded + 4 * nquad + 2 * nblk);
    int *s1 = s0 + nblk;
    const dim3 grid((unsigned)((nblk + 255) / 256));
    hipLaunchKernelGGL(k_rlsp_blocks, grid, dim3(256), 0, st, reinterpret_cast<const float4 *>(rlsp_padded), table,
                       make_grid(g), nblk);
    const int rc = check_launch("k_rlsp_blocks");
    if (rc) return rc;
    return dh::block_radius_passes(table, s0, s1, (g.nx + 3) / 4, (g.ny + 3) / 4, (g.nz + 3) / 4, st);
}

extern "C" int dh_rlsp_from_hu(const float *hu, float *rlsp_padded, dh_grid g, const double *hu_knots,
                               const double *rlsp_knots, int32_t n_knots, void *stream) {
    if (!grid_ok(g)) return DH_ERR_ARG;
    if (!hu || !rlsp_padded || !hu_knots || !rlsp_knots || n_knots < 2 || n_knots > MAX_KNOTS) {
        set_error("dh_rlsp_from_hu: bad arguments (n_knots=%d, max %d)", n_knots, MAX_KNOTS);
        return DH_ERR_ARG;
    }
    Knots kn;
    kn.n = n_knots;
    for (int t = 0; t < n_knots; ++t) {
        kn.hu[t] = hu_knots[t];
        kn.val[t] = rlsp_knots[t];
        if (t && !(hu_knots[t] > hu_knots[t - 1])) {
            set_error("dh_rlsp_from_hu: HU knots must be strictly ascending");
            return DH_ERR_ARG;
        }
    }
    for (int t = n_knots; t < MAX_KNOTS; ++t) kn.hu[t] = kn.val[t] = 0.0;
    const long long nquad = quad_count(g.nx, g.ny, g.nz);
    const unsigned nb = (unsigned)((nquad + 255) / 256);
    hipLaunchKernelGGL(k_rlsp_from_hu, dim3(nb), dim3(256), 0, (hipStream_t)stream, hu,
                       reinterpret_cast<float4 *>(rlsp_padded), make_grid(g), kn, nquad);
    const int rc = check_launch("k_rlsp_from_hu");
    return rc ? rc : launch_blocks(rlsp_padded, g, (hipStream_t)stream);
}

extern "C" int dh_rlsp_pad(const float *rlsp, float *rlsp_padded, dh_grid g, void *stream) {
    if (!grid_ok(g)) return DH_ERR_ARG;
    if (!rlsp || !rlsp_padded) {
        set_error("dh_rlsp_pad: null pointer");
        return DH_ERR_ARG;
    }
    const long long nquad = quad_count(g.nx, g.ny, g.nz);
    const unsigned nb = (unsigned)((nquad + 255) / 256);
    hipLaunchKernelGGL(k_rlsp_pad, dim3(nb), dim3(256), 0, (hipStream_t)stream, rlsp,
                       reinterpret_cast<float4 *>(rlsp_padded), make_grid(g), nquad);
    const int rc = check_launch("k_rlsp_pad");
    return rc ? rc : launch_blocks(rlsp_padded, g, (hipStream_t)stream);
}

extern "C" int dh_wet_raymarch(const float *rlsp_padded, float *wet_raw, dh_grid g, const dh_beam_geom *geom,
                               void *stream) {
    if (!grid_ok(g)) return DH_ERR_ARG;
    if (!rlsp_padded || !wet_raw || !geom) {
        set_error("dh_wet_raymarch: null pointer");
        return DH_ERR_ARG;
    }
    if (((uintptr_t)rlsp_padded & 15u) != 0) {
        set_error("dh_wet_raymarch: the quad RLSP volume must be 16-byte aligned");
        return DH_ERR_ARG;
    }
    BeamCtx c = {};
    c.rlsp_q = reinterpret_cast<const float4 *>(rlsp_padded);
    c.wet_raw = wet_raw;
    c.g = make_grid(g);
    c.b = make_geom(*geom);
    c.bg = brick_grid(c.g);
    const unsigned nb = (unsigned)((c.bg.count() + 3) / 4);
    hipLaunchKernelGGL(k_wet_raymarch, dim3(nb), dim3(256), 0, (hipStream_t)stream, c, nb);
    return check_launch("k_wet_raymarch");
}

// Batched launches: ctx is a device table of n_beams entries.
namespace dh {
int launch_wet_batch(const BeamCtx *ctx, const long long *brick_off, const long long *block_off, int n_beams,
                     long long total_bricks, long long total_blocks, bool use_need, int persist_blocks,
                     int ray_shape, int smooth_box_floats, hipStream_t st, void *const *ev) {
    // item spaces above 2^24 bricks go out in several launches
    constexpr long long kMaxItems = 1LL << 24;
    mark(ev, 2 * DH_TIME_RAYMARCH, st);
    if (persist_blocks > 0 && total_bricks > 0) {
        const long long nb = (total_bricks + 3) / 4 < persist_blocks ? (total_bricks + 3) / 4 : persist_blocks;
        hipLaunchKernelGGL(k_wet_raymarch_batch<true>, dim3((unsigned)nb), dim3(256), 0, st, ctx, brick_off,
                           n_beams, 0LL, total_bricks, use_need ? 1 : 0, 4 * nb, ray_shape);
        const int rc = check_launch("k_wet_raymarch_batch");
        if (rc) return rc;
    }
    for (long long base = 0; persist_blocks <= 0 && base < total_bricks; base += kMaxItems) {
        const long long n = total_bricks - base < kMaxItems ? total_bricks - base : kMaxItems;
        hipLaunchKernelGGL(k_wet_raymarch_batch<false>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, ctx,
                           brick_off, n_beams, base, total_bricks, use_need ? 1 : 0, total_bricks, ray_shape);
        const int rc = check_launch("k_wet_raymarch_batch");
        if (rc) return rc;
    }
    mark(ev, 2 * DH_TIME_RAYMARCH + 1, st);
    mark(ev, 2 * DH_TIME_SMOOTH, st);
    for (long long base = 0; base < total_blocks; base += kMaxItems / 4) {
        const long long n = total_blocks - base < kMaxItems / 4 ? total_blocks - base : kMaxItems / 4;
        if (smooth_box_floats > 0)
            hipLaunchKernelGGL(k_wet_smooth_tile_batch, dim3((unsigned)n), dim3(256),
                               (unsigned)smooth_box_floats * sizeof(float), st, ctx, block_off, n_beams, base,
                               total_blocks, use_need ? 1 : 0);
        else
            hipLaunchKernelGGL(k_wet_smooth_batch, dim3((unsigned)n), dim3(256), 0, st, ctx, block_off, n_beams,
                               base, total_blocks, use_need ? 1 : 0);
        const int rc = check_launch(smooth_box_floats > 0 ? "k_wet_smooth_tile_batch" : "k_wet_smooth_batch");
        if (rc) return rc;
    }
    mark(ev, 2 * DH_TIME_SMOOTH + 1, st);
    return DH_OK;
}
}  // namespace dh

extern "C" int dh_wet_smooth(const float *wet_raw, float *wet, dh_grid g, const dh_beam_geom *geom, void *stream) {
    if (!grid_ok(g)) return DH_ERR_ARG;
    if (!wet_raw || !wet || !geom || wet_raw == wet) {
        set_error("dh_wet_smooth: null or aliased pointers");
        return DH_ERR_ARG;
    }
    if (g.spacing < 0.2f) {
        // |offset| <= 9 mm must fit the signed-char stencil (45 voxels)
        set_error("dh_wet_smooth: spacing %g mm below 0.2 mm", (double)g.spacing);
        return DH_ERR_ARG;
    }
    Stencil st;
    make_stencil(*geom, (double)g.spacing, st);
    const long long nvox = (long long)g.nx * g.ny * g.nz;
    const unsigned nb = (unsigned)((nvox + 255) / 256);
    hipLaunchKernelGGL(k_wet_smooth, dim3(nb), dim3(256), 0, (hipStream_t)stream, wet_raw, wet, make_grid(g), st,
                       nvox);
    return check_launch("k_wet_smooth");
}
