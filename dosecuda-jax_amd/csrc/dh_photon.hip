// Photon collapsed-cone convolution (CCC) dose of one control point on gfx950:
// the photon half of the reference extension (photon_dose_cuda,
// IMRTClasses.cu:353-405; kernels CudaClasses.cu:180-228 and
// IMRTClasses.cu:209-350).
//
// In place of the reference's two border-addressed float textures (density,
// TERMA) the kernels read a record volume over the grid padded by one voxel on
// every side (include/dosehip.h): per padded cell 32 bytes, the bilinear face
// forms of TERMA and of density side by side (the cone sum reads both from the
// same cache lines), then a second copy of the density forms alone, 16 bytes
// per cell (the ray-trace reads only density: half the lines).  A trilinear
// sample is two cells' forms blended, 16-byte buffer loads, with the texture's
// border values (TERMA 0, density AIR_DENSITY) built into the padding and
// returned outright for samples whose every corner lies outside the grid.  A
// table of uniform 4x4x4 blocks of cells follows, for the ray-trace's runs.
// Work mapping as in the proton kernels: one wave per 4x4x4 brick of voxels,
// lane = voxel, four bricks per workgroup.
#include "dh_common.h"

namespace dh {

constexpr float kAirDensity = 0.0012f;  // AIR_DENSITY, IMRTClasses.cuh:6

struct PGrid {
    int nx, ny, nz;      // voxels
    int px, pxy;         // padded row and slice lengths (records)
    float s, inv_s;
    float hx, hy, hz;    // nx + 1, ny + 1, nz + 1: a sample at padded index c reads the grid iff c < h (and c >= 0)
    unsigned rho_base;   // byte offset of the density-only quads (32 bytes x padded cells)
    unsigned tab_off;    // byte offset of the block table (48 bytes x padded cells)
    int nbx, nby, nbz;   // blocks of 4x4x4 sample cells per axis (cells 0 .. n: (n + 4) / 4)
};

static PGrid make_pgrid(const dh_grid &g) {
    PGrid p;
    p.nx = g.nx; p.ny = g.ny; p.nz = g.nz;
    p.px = g.nx + 2;
    p.pxy = (g.nx + 2) * (g.ny + 2);
    p.s = g.spacing;
    p.inv_s = 1.0f / g.spacing;
    p.hx = (float)(g.nx + 1); p.hy = (float)(g.ny + 1); p.hz = (float)(g.nz + 1);
    p.rho_base = (unsigned)(32LL * (g.nx + 2) * (g.ny + 2) * (g.nz + 2));
    p.tab_off = (unsigned)(48LL * (g.nx + 2) * (g.ny + 2) * (g.nz + 2));
    p.nbx = (g.nx + 4) / 4; p.nby = (g.ny + 4) / 4; p.nbz = (g.nz + 4) / 4;
    return p;
}

static long long record_count(const dh_grid &g) {
    return (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
}

// Blocks of the density table: a sample's cell (padded index floor(c)) lies in
// 0 .. n per axis, so (n + 1) cells, (n + 4) / 4 blocks.
static long long photon_block_count(const dh_grid &g) {
    return (long long)((g.nx + 4) / 4) * ((g.ny + 4) / 4) * ((g.nz + 4) / 4);
}

static bool photon_grid_ok(const dh_grid &g, const char *what) {
    if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0 || !(g.spacing > 0.0f)) {
        set_error("%s: invalid grid %d x %d x %d, spacing %g", what, g.nx, g.ny, g.nz, (double)g.spacing);
        return false;
    }
    // 32-bit byte offsets of the record and table loads (vector offset + scalar
    // offset of the density copy and the slice, the table after them, below 2^32)
    if (48.0 * (double)record_count(g) + 8.0 * (double)photon_block_count(g) >= 4294967296.0) {
        set_error("%s: grid %d x %d x %d too large (48 (nx+2)(ny+2)(nz+2) + 8 blocks must stay below 2^32)", what,
                  g.nx, g.ny, g.nz);
        return false;
    }
    return true;
}

// Image -> photon head frame (IMRTBeam::pointXYZImageToHead, IMRTClasses.cu:41-79):
// couch about y, gantry about z, collimator about y, then (-x_c, z_c, y_c).
__device__ __forceinline__ float3 ph_image_to_head(const Geom &b, float sca, float cca, float3 p) {
    const float xt = p.x * b.cc - p.z * b.sc;
    const float zt = p.x * b.sc + p.z * b.cc;
    const float xg = xt * b.cg + p.y * b.sg;
    const float yg = p.y * b.cg - xt * b.sg;
    const float xc = xg * cca - zt * sca;
    const float zc = xg * sca + zt * cca;
    return make_float3(-xc, zc, yg);
}

// ---------------------------------------------------------------------------
// Record volume: padded cell t -> (i', j', k'); value f(i', j', k') of the
// voxel (i'-1, j'-1, k'-1) inside the grid, the border value outside.
template <typename F>
__device__ __forceinline__ float4 padded_quad(const PGrid &g, int i, int j, int k, float border, F val) {
    auto f = [&](int a, int b, int c) -> float {
        const bool in = a >= 1 && a <= g.nx && b >= 1 && b <= g.ny && c >= 1 && c <= g.nz;
        return in ? val(a - 1, b - 1, c - 1) : border;
    };
    const float a = f(i, j, k), b = f(i + 1, j, k), c = f(i, j + 1, k), d = f(i + 1, j + 1, k);
    return make_float4(a, b - a, c - a, (d - c) - (b - a));
}

__device__ __forceinline__ void padded_coords(const PGrid &g, long long t, int &i, int &j, int &k) {
    const unsigned u = (unsigned)t;
    const unsigned r = u / (unsigned)g.px;
    i = (int)(u - r * (unsigned)g.px);
    const unsigned py = (unsigned)(g.pxy / g.px);
    k = (int)(r / py);
    j = (int)(r - (unsigned)k * py);
}

__device__ __forceinline__ long long vox_index(const PGrid &g, int i, int j, int k) {
    return i + (long long)g.nx * (j + (long long)g.ny * k);
}

// field 0: TERMA (float4 2t), 1: density (float4 2t + 1 and its copy 2n + t)
__global__ __launch_bounds__(256) void k_ph_records_from(const float *__restrict__ src, float4 *__restrict__ rec,
                                                         PGrid g, long long n, int field, float border) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    padded_coords(g, t, i, j, k);
    const float4 q = padded_quad(g, i, j, k, border, [&](int a, int b, int c) { return gld(src, vox_index(g, a, b, c)); });
    gst(rec, 2 * t + field, q);
    if (field) gst(rec, 2 * n + t, q);
}

__global__ __launch_bounds__(256) void k_ph_density_from_hu(const float *__restrict__ hu, float4 *__restrict__ rec,
                                                            PGrid g, Knots kn, long long n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    padded_coords(g, t, i, j, k);
    const float4 q = padded_quad(g, i, j, k, kAirDensity, [&](int a, int b, int c) {
        return (float)interp_knots(kn, (double)gld(hu, vox_index(g, a, b, c)));
    });
    gst(rec, 2 * t + 1, q);
    gst(rec, 2 * n + t, q);
}

// Block table after the records: one float2 per block of 4x4x4 sample cells
// (padded cells 4b .. 4b + 3 per axis): the density every corner those cells
// touch holds (padded indices 4b .. 4b + 4; AIR_DENSITY past the padded grid,
// as the padding itself) or NaN where they differ, and the run radius, set by
// block_radius_passes.  A sample in a cell of a uniform block is exactly that
// density: every form it blends is (rho, 0, 0, 0).
__global__ __launch_bounds__(256) void k_ph_density_blocks(const float4 *__restrict__ rho_q,
                                                           float2 *__restrict__ table, PGrid g, long long nblk) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nblk) return;
    const int bx = (int)(t % g.nbx);
    const long long r = t / g.nbx;
    const int by = (int)(r % g.nby), bz = (int)(r / g.nby);
    const int py = g.pxy / g.px, pz = g.nz + 2;
    auto val = [&](int i, int j, int k) -> float {
        return (i < g.px && j < py && k < pz) ? gld(rho_q, (long long)i + (long long)g.px * (j + (long long)py * k)).x
                                              : kAirDensity;
    };
    const float v = val(4 * bx, 4 * by, 4 * bz);
    bool same = v == v;
    for (int k = 4 * bz; k <= 4 * bz + 4 && same; ++k)
        for (int j = 4 * by; j <= 4 * by + 4 && same; ++j)
            for (int i = 4 * bx; i <= 4 * bx + 4 && same; ++i) same = val(i, j, k) == v;
    table[t] = make_float2(same ? v : __int_as_float(0x7fc00000), 0.0f);
}

// ---------------------------------------------------------------------------
// Trilinear samples from the records (buffer loads; byte offsets < 2^32).
__device__ __forceinline__ float4 load4(__amdgpu_buffer_rsrc_t rs, unsigned off, int soff) {
    const auto r = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, soff, 0);
    return make_float4(__uint_as_float(r[0]), __uint_as_float(r[1]), __uint_as_float(r[2]), __uint_as_float(r[3]));
}

__device__ __forceinline__ float bilerp(float4 q, float ax, float ay) {
    return fmaf(ay, fmaf(ax, q.w, q.z), fmaf(ax, q.y, q.x));
}

// (the padded slice length and index stay below 2^24: 24-bit products, low 32 bits)
__device__ __forceinline__ unsigned cell_index(const PGrid &g, int ix, int iy, int iz) {
    return (unsigned)ix + __umul24((unsigned)iy, (unsigned)g.px) + __umul24((unsigned)iz, (unsigned)g.pxy);
}

// TERMA and density at padded index c, border values where the texture would
// read only border texels (some axis outside [0, n + 1)).  Lanes that are not
// live read nothing (their te, rho are not used).
__device__ __forceinline__ void sample_tr(__amdgpu_buffer_rsrc_t rs, const PGrid &g, float cx, float cy, float cz,
                                          bool live, float &te, float &rho) {
    const bool in = cx >= 0.0f && cx < g.hx && cy >= 0.0f && cy < g.hy && cz >= 0.0f && cz < g.hz;
    const float fx = floorf(cx), fy = floorf(cy), fz = floorf(cz);
    const unsigned off = in ? cell_index(g, (int)fx, (int)fy, (int)fz) << 5 : 0u;
    const int so_z = g.pxy << 5;
    const float ax = cx - fx, ay = cy - fy, az = cz - fz;
    float4 t0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), d0 = t0, t1 = t0, d1 = t0;
    if (live) {
        t0 = load4(rs, off, 0);
        d0 = load4(rs, off + 16u, 0);
        t1 = load4(rs, off, so_z);
        d1 = load4(rs, off + 16u, so_z);
    }
    const float tt0 = bilerp(t0, ax, ay), tt1 = bilerp(t1, ax, ay);
    const float dd0 = bilerp(d0, ax, ay), dd1 = bilerp(d1, ax, ay);
    te = in ? fmaf(az, tt1 - tt0, tt0) : 0.0f;
    rho = in ? fmaf(az, dd1 - dd0, dd0) : kAirDensity;
}

// A byte offset past every record volume's end (48 bytes per padded cell plus
// the block table stay below 2^32 - 32 for any grid that passes
// photon_grid_ok): a buffer load there, or 16 bytes on, returns zeros without
// a memory access.
constexpr unsigned kNoRecord = 0xffffffe0u;

__device__ __forceinline__ float4 bpermute4(int addr, float4 x) {
    return make_float4(__int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(x.x))),
                       __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(x.y))),
                       __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(x.z))),
                       __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(x.w))));
}

// The same for a padded index known to lie in [0, n + 1) on every axis, with
// the upper face shared: lane l + 16 is the voxel one z-layer up in the brick
// and its rays run parallel to lane l's to within a few hundredths of a cell,
// so at the same step it mostly samples the cell one slice up, whose lower
// face (the forms at byte offset off + slice) is lane l's upper face.  A lane
// takes those forms from lane l + 16 through the lane crossbar when that
// lane's cell offset is exactly off + slice, and reads them itself otherwise
// (the top layer of the brick, cells that differ near a cell boundary, a
// partner that is not live): the same bytes either way, so the same results,
// with fewer cache-line lookups for the upper faces.  up = byte address of
// lane l + 16 for ds_bpermute (the top layer wraps to the bottom one, whose
// offsets never match); a lane that is not live has offset kNoRecord.
__device__ __forceinline__ void sample_tr_inside(__amdgpu_buffer_rsrc_t rs, const PGrid &g, float cx, float cy,
                                                 float cz, bool live, int up, float &te, float &rho) {
    // c >= 0: truncation is floor.  Loads at offset kNoRecord (past the end of
    // the buffer) return zeros without a memory access: lanes that are not
    // live, and upper faces taken from the partner, read nothing.
    const unsigned so_z = (unsigned)g.pxy << 5;
    const unsigned off = live ? cell_index(g, (int)cx, (int)cy, (int)cz) << 5 : kNoRecord;
    const unsigned poff = (unsigned)__builtin_amdgcn_ds_bpermute(up, (int)off);
    const bool share = live && poff == off + so_z;
    const unsigned off1 = (share || !live) ? kNoRecord : off + so_z;
    const float ax = __builtin_amdgcn_fractf(cx), ay = __builtin_amdgcn_fractf(cy), az = __builtin_amdgcn_fractf(cz);
    const float4 t0 = load4(rs, off, 0), d0 = load4(rs, off + 16u, 0);
    float4 t1 = load4(rs, off1, 0), d1 = load4(rs, off1 + 16u, 0);
    const float4 tu = bpermute4(up, t0), du = bpermute4(up, d0);
    if (share) {
        t1 = tu;
        d1 = du;
    }
    const float tt0 = bilerp(t0, ax, ay), tt1 = bilerp(t1, ax, ay);
    const float dd0 = bilerp(d0, ax, ay), dd1 = bilerp(d1, ax, ay);
    te = fmaf(az, tt1 - tt0, tt0);
    rho = fmaf(az, dd1 - dd0, dd0);
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t record_rsrc(const float4 *rec, const PGrid &g) {
    const unsigned bytes = g.rho_base + g.rho_base / 2u;
    return __builtin_amdgcn_make_buffer_rsrc((void *)rec, (short)0, (int)bytes, 0x00020000);
}

// Brick mapping.  Bricks are numbered tile by tile, 4x4x4 bricks (16^3 voxels)
// per tile, x fastest inside a tile and across tiles; workgroup q takes virtual
// bricks 4q .. 4q + 3 (wave w the brick 4q + w, lane = x + 4 y + 16 z of the
// brick; virtual bricks past the grid's edge leave at once).  Workgroups go to
// the XCDs in chunks of 64, four tiles, in turn: the bricks a control point
// irradiates form one compact region of the grid, which contiguous eighths of
// the brick range would hand to one or two XCDs, while tile-sized chunks keep
// the rays of an XCD's concurrent waves in few cache lines of its L2.  Cone sum
// of the head plan per control point: 5.98 ms with contiguous eighths, 2.70 with
// chunks of 16 workgroups in raster brick order, 2.28 with tiles and chunks of 64.
__device__ __forceinline__ bool brick_voxel(const BrickGrid &bg, const PGrid &g, unsigned nblocks, int &i, int &j,
                                            int &k) {
    const unsigned q = xcd_chunk<64>(blockIdx.x, nblocks);
    const unsigned vb = 4u * q + (threadIdx.x >> 6);
    const unsigned tile = vb >> 6, w = vb & 63u;
    const unsigned tx = (unsigned)(bg.bx + 3) >> 2, ty = (unsigned)(bg.by + 3) >> 2;
    const unsigned t1 = tile / tx;
    const int bi = (int)((tile - t1 * tx) * 4u + (w & 3u));
    const int bj = (int)((t1 % ty) * 4u + ((w >> 2) & 3u));
    const int bk = (int)((t1 / ty) * 4u + (w >> 4));
    if (bi >= bg.bx || bj >= bg.by || bk >= bg.bz) return false;
    const int lane = threadIdx.x & 63;
    i = bi * BRICK + (lane & 3);
    j = bj * BRICK + ((lane >> 2) & 3);
    k = bk * BRICK + (lane >> 4);
    return i < g.nx && j < g.ny && k < g.nz;
}

__device__ __forceinline__ float3 ph_voxel_xyz(const PGrid &g, const Geom &b, int i, int j, int k) {
    return make_float3(fmaf((float)i, g.s, -b.iso[0]), fmaf((float)j, g.s, -b.iso[1]),
                       fmaf((float)k, g.s, -b.iso[2]));
}

// ---------------------------------------------------------------------------
// Radiological depth (rayTraceKernel, CudaClasses.cu:180-228) with the
// reference's sample positions p_n = fma(u, n, v), texture coordinate
// fma(1/s, p + iso, 1/2) and bounds test [0, n) per axis.
//
// Uniform runs: when sample n's block of cells is uniform with radius r, every
// cell of the cube of blocks within r of it has that density, so the samples
// that stay in the cube add it at once.  Per axis the run ends one step before
// the ray reaches the cube's far face, and before padded index n (moving up) or
// 1 (moving down): the run's samples lie inside the grid, and the margin of one
// step covers the rounding of the positions.  In air and other uniform regions
// one table read replaces the samples of up to 4 (2r + 1) cells.
__global__ __launch_bounds__(256) void k_ph_raytrace(const float4 *__restrict__ rec, float *__restrict__ depth,
                                                     PGrid g, Geom b, BrickGrid bg, unsigned nblocks) {
    int i, j, k;
    if (!brick_voxel(bg, g, nblocks, i, j, k)) return;
    const float3 v = ph_voxel_xyz(g, b, i, j, k);
    const float dx = b.src[0] - v.x, dy = b.src[1] - v.y, dz = b.src[2] - v.z;
    const float len = sqrtf(dx * dx + dy * dy + dz * dz);
    const float ux = dx / len, uy = dy / len, uz = dz / len;
    const unsigned nblk = (unsigned)g.nbx * (unsigned)g.nby * (unsigned)g.nbz;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)rec, (short)0, (int)(g.tab_off + 8u * nblk), 0x00020000);
    const float nxf = (float)g.nx, nyf = (float)g.ny, nzf = (float)g.nz;
    // per axis: cells per step, and the run's face: 4 b + off + 4 r sg, clamped at lim
    const float dux = ux * g.inv_s, duy = uy * g.inv_s, duz = uz * g.inv_s;
    const float rdx = dux != 0.0f ? 1.0f / dux : 0.0f, rdy = duy != 0.0f ? 1.0f / duy : 0.0f;
    const float rdz = duz != 0.0f ? 1.0f / duz : 0.0f;
    const float adx = dux != 0.0f ? 0.0f : 3.0e38f, ady = duy != 0.0f ? 0.0f : 3.0e38f;
    const float adz = duz != 0.0f ? 0.0f : 3.0e38f;
    const float sgx = dux > 0.0f ? 1.0f : -1.0f, sgy = duy > 0.0f ? 1.0f : -1.0f, sgz = duz > 0.0f ? 1.0f : -1.0f;
    const int ofx = dux > 0.0f ? 4 : 0, ofy = duy > 0.0f ? 4 : 0, ofz = duz > 0.0f ? 4 : 0;
    const unsigned nbxy = (unsigned)g.nbx * (unsigned)g.nby;
    float acc = -0.05f;
    // the two density forms of the cell of the lane's last sample: at 1 mm steps
    // a sample often lies in the cell of the one before and reads no memory
    unsigned off_prev = 0xffffffffu;
    float4 q0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), q1 = q0;
    const int so_z = g.pxy << 4;
    for (int n = 0;;) {
        const float fn = (float)n;
        const float tx = fmaf(g.inv_s, fmaf(ux, fn, v.x) + b.iso[0], 0.5f);
        const float ty = fmaf(g.inv_s, fmaf(uy, fn, v.y) + b.iso[1], 0.5f);
        const float tz = fmaf(g.inv_s, fmaf(uz, fn, v.z) + b.iso[2], 0.5f);
        if (!(tx >= 0.0f && tx < nxf && ty >= 0.0f && ty < nyf && tz >= 0.0f && tz < nzf)) break;
        // padded index tex + 1/2 >= 1/2: truncation is floor
        const float cx = tx + 0.5f, cy = ty + 0.5f, cz = tz + 0.5f;
        const int ix = (int)cx, iy = (int)cy, iz = (int)cz;
        const int bx = ix >> 2, by = iy >> 2, bz = iz >> 2;
        const unsigned bidx = (unsigned)bx + __umul24((unsigned)g.nbx, (unsigned)by) + __umul24(nbxy, (unsigned)bz);
        const auto te = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)(g.tab_off + (bidx << 3)), 0, 0);
        const float u = __uint_as_float(te[0]);
        if (u == u) {
            const float r4 = 4.0f * __uint_as_float(te[1]);
            const float fx = dux > 0.0f ? fminf(fmaf(r4, sgx, (float)((bx << 2) + ofx)), nxf)
                                        : fmaxf(fmaf(r4, sgx, (float)((bx << 2) + ofx)), 1.0f);
            const float fy = duy > 0.0f ? fminf(fmaf(r4, sgy, (float)((by << 2) + ofy)), nyf)
                                        : fmaxf(fmaf(r4, sgy, (float)((by << 2) + ofy)), 1.0f);
            const float fz = duz > 0.0f ? fminf(fmaf(r4, sgz, (float)((bz << 2) + ofz)), nzf)
                                        : fmaxf(fmaf(r4, sgz, (float)((bz << 2) + ofz)), 1.0f);
            const float sx = fmaf(fx - cx, rdx, adx), sy = fmaf(fy - cy, rdy, ady), sz = fmaf(fz - cz, rdz, adz);
            const int T = (int)fminf(fminf(fminf(sx, sy), sz), 1.0e9f) - 1;
            if (T >= 1) {
                acc = fmaf(fmaxf(u, 0.0f) * (float)(T + 1), 0.1f, acc);
                n += T + 1;
                continue;
            }
        }
        auto sample = [&](float sx, float sy, float sz) {
            const unsigned off = cell_index(g, (int)sx, (int)sy, (int)sz) << 4;
            if (off != off_prev) {
                q0 = load4(rs, off, (int)g.rho_base);
                q1 = load4(rs, off, (int)(g.rho_base + (unsigned)so_z));
                off_prev = off;
            }
            const float ax = __builtin_amdgcn_fractf(sx), ay = __builtin_amdgcn_fractf(sy);
            const float r0 = bilerp(q0, ax, ay), r1 = bilerp(q1, ax, ay);
            const float rho = fmaf(__builtin_amdgcn_fractf(sz), r1 - r0, r0);
            acc = fmaf(fmaxf(rho, 0.0f), 0.1f, acc);
        };
        sample(cx, cy, cz);
        // the next sample too, without a table read (outside a uniform block it
        // is most often in the same block or another non-uniform one)
        const float f1 = (float)(n + 1);
        const float ux1 = fmaf(g.inv_s, fmaf(ux, f1, v.x) + b.iso[0], 0.5f);
        const float uy1 = fmaf(g.inv_s, fmaf(uy, f1, v.y) + b.iso[1], 0.5f);
        const float uz1 = fmaf(g.inv_s, fmaf(uz, f1, v.z) + b.iso[2], 0.5f);
        if (!(ux1 >= 0.0f && ux1 < nxf && uy1 >= 0.0f && uy1 < nyf && uz1 >= 0.0f && uz1 < nzf)) break;
        sample(ux1 + 0.5f, uy1 + 0.5f, uz1 + 0.5f);
        n += 2;
    }
    gst(depth, vox_index(g, i, j, k), acc);
}

// ---------------------------------------------------------------------------
// TERMA and electron contamination (termaKernel, IMRTClasses.cu:209-250).
struct PhotonTermaArgs {
    const float *profile;   // [3][n_profile]
    const float *spectrum;  // [3][n_spectrum]
    const dh_mlc_pair *mlc;
    int n_profile, n_spectrum, n_pairs;
    float sad_p, sad_s, inv_sig_p, inv_sig_s;  // 1 / (sigma sqrt 2)
    float mlc_distance, mlc_scale;             // (SAD - mlc_distance) / SAD
    float ssw, ea, esw, el_floor, el_inv;      // exp(-ea dmax), 1 / (1 - exp(-ea dmax))
    float sca, cca;
};

// offAxisFactors (IMRTClasses.cu:152-171): lowerBound of r in the radius table,
// clamped at both ends, linear in between.
__device__ __forceinline__ void off_axis(const PhotonTermaArgs &a, float r, float &oaf, float &soft) {
    const float *rad = a.profile, *inten = a.profile + a.n_profile, *sf = a.profile + 2 * a.n_profile;
    int lo = 0;  // first index with r <= rad[lo]
    for (int t = 0; t < a.n_profile; ++t) lo += (r > rad[t]) ? 1 : 0;
    if (lo == 0 || lo == a.n_profile) {
        const int e = lo == 0 ? 0 : a.n_profile - 1;
        oaf = inten[e];
        soft = sf[e];
        return;
    }
    const float m = (r - rad[lo - 1]) / (rad[lo] - rad[lo - 1]);
    oaf = fmaf(inten[lo] - inten[lo - 1], m, inten[lo - 1]);
    soft = fmaf(sf[lo] - sf[lo - 1], m, sf[lo - 1]);
}

// One source's aperture term of a leaf pair, 0 without evaluating the error
// functions when the pair is closed or every edge lies 5 sigma sqrt 2 beyond the
// point on one side (erff is then +-1 exactly in float32, so the term is 0).
__device__ __forceinline__ float pair_term(float xl, float xr, float yb, float yt, float3 h, float div, float inv) {
    const float tip1 = fmaf(xl - h.x, div, h.x) * inv, tip2 = fmaf(xr - h.x, div, h.x) * inv;
    const float e1 = fmaf(yb - h.y, div, h.y) * inv, e2 = fmaf(yt - h.y, div, h.y) * inv;
    const bool far = (tip1 >= 5.0f && tip2 >= 5.0f) || (tip1 <= -5.0f && tip2 <= -5.0f) ||
                     (e1 >= 5.0f && e2 >= 5.0f) || (e1 <= -5.0f && e2 <= -5.0f);
    if (far) return 0.0f;
    const float ex = 0.5f * (erff(tip2) - erff(tip1));
    const float ey = 0.5f * (erff(e2) - erff(e1));
    return ex * ey;
}

__global__ __launch_bounds__(256) void k_ph_terma(const float *__restrict__ depth, float *__restrict__ terma,
                                                  float *__restrict__ electron, PGrid g, Geom b, PhotonTermaArgs a,
                                                  BrickGrid bg, unsigned nblocks) {
    int i, j, k;
    if (!brick_voxel(bg, g, nblocks, i, j, k)) return;
    const float3 h = ph_image_to_head(b, a.sca, a.cca, ph_voxel_xyz(g, b, i, j, k));
    const float d_p = a.sad_p - h.z, d_s = a.sad_s - h.z;
    float oaf, soft;
    off_axis(a, hypotf(h.x, h.y) * d_p / a.sad_p, oaf, soft);
    // headTransmission (IMRTClasses.cu:120-150) for the primary and the scatter source
    const float div_p = (a.sad_p - h.z) / (a.mlc_distance - h.z);
    const float div_s = (a.sad_s - h.z) / (a.mlc_distance - h.z);
    float tp = 0.0f, ts = 0.0f;
    for (int p = 0; p < a.n_pairs; ++p) {
        const dh_mlc_pair m = a.mlc[p];
        if (m.x1 == m.x2) continue;  // closed pair: both tips equal, the term is exactly 0
        const float yb = (m.y_offset - 0.5f * m.y_width) * a.mlc_scale;
        const float yt = (m.y_offset + 0.5f * m.y_width) * a.mlc_scale;
        const float xl = m.x1 * a.mlc_scale, xr = m.x2 * a.mlc_scale;
        tp += pair_term(xl, xr, yb, yt, h, div_p, a.inv_sig_p);
        ts += pair_term(xl, xr, yb, yt, h, div_s, a.inv_sig_s);
    }
    const long long v = vox_index(g, i, j, k);
    const float wet = gld(depth, v);
    const float tr = fminf(1.0f, ts / tp);  // fminf: a NaN quotient gives 1
    const float inv_sq = (a.sad_p / d_p) * (a.sad_p / d_p);
    const float inv_s = a.sad_s / d_s;
    float te = 0.0f;
    for (int e = 0; e < a.n_spectrum; ++e) {
        const float att = __expf(-a.spectrum[e] * wet * soft);
        te += (1.0f - tr) * (a.spectrum[a.n_spectrum + e] * att * inv_sq) +
              tr * a.spectrum[2 * a.n_spectrum + e] * att * inv_s;
    }
    const float el = fmaxf(0.0f, (__expf(-a.ea * wet) - a.el_floor) * a.el_inv);
    gst(terma, v, oaf * ((1.0f - a.ssw) * tp * te + a.ssw * ts * te));
    gst(electron, v, a.esw * (0.4f + 0.3f * tr) * el * tp);
}

// ---------------------------------------------------------------------------
// Collapsed-cone dose (cccKernel, IMRTClasses.cu:252-350).
constexpr int N_CONES = 6;
constexpr int N_AZ = 12;
struct PhotonCccArgs {
    float t[N_CONES * N_AZ][3];  // image-space cone directions before the tilt (head -> image of the head-frame set)
    float k1[N_CONES];           // sin(theta) Am / am^2
    float k2[N_CONES];           // sin(theta) Bm / bm
    float am_l2[N_CONES];        // am log2(e)
    float am[N_CONES], bm[N_CONES];
    float len0[N_CONES];         // initial ray length (cm, float32 counter)
    int nsteps[N_CONES];         // values of the counter len0, len0 - sp, ... that are >= 0
    float ax[3];                 // unit vector from the isocentre to the source
    float sp;                    // spacing / 10 (cm)
    float mu_cal, mu, scale;
    int accumulate;
};

__global__ __launch_bounds__(256) void k_ph_ccc(const float4 *__restrict__ rec, const float *__restrict__ terma,
                                                const float *__restrict__ electron, float *__restrict__ dose, PGrid g,
                                                Geom b, PhotonCccArgs a, BrickGrid bg, unsigned nblocks) {
    // Every lane of the wave stays to the end (the shared upper faces read
    // other lanes' offsets); a lane is live when its voxel is in the grid and
    // has TERMA > 0.01, and only live lanes read memory or write a dose.
    int i = 0, j = 0, k = 0;
    const bool in = brick_voxel(bg, g, nblocks, i, j, k);
    if (!in) i = j = k = 0;   // a lane past the grid's edge computes a voxel it does not keep
    const long long v = vox_index(g, i, j, k);
    const bool live = in && gld(terma, v) > 0.01f;
    if (in && !live && !a.accumulate) gst(dose, v, 0.0f);
    if (!__any(live)) return;   // wave-uniform
    const int up = (int)(((threadIdx.x & 63u) + 16u) << 2);
    const float3 p = ph_voxel_xyz(g, b, i, j, k);
    // kernelTilt (IMRTClasses.cu:173-203): rotation about a x u taking the
    // isocentre's direction a to the voxel's direction u to the source, as the
    // matrix R = c I + [k]x + k k^T / (1 + c), k = a x u, c = a . u
    const float dx = b.src[0] - p.x, dy = b.src[1] - p.y, dz = b.src[2] - p.z;
    const float rn = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    const float ux = dx * rn, uy = dy * rn, uz = dz * rn;
    const float c = a.ax[0] * ux + a.ax[1] * uy + a.ax[2] * uz;
    const float kx = a.ax[1] * uz - a.ax[2] * uy, ky = a.ax[2] * ux - a.ax[0] * uz, kz = a.ax[0] * uy - a.ax[1] * ux;
    const float f = 1.0f / (1.0f + c);
    const float r00 = fmaf(kx * kx, f, c), r01 = fmaf(kx * ky, f, -kz), r02 = fmaf(kx * kz, f, ky);
    const float r10 = fmaf(ky * kx, f, kz), r11 = fmaf(ky * ky, f, c), r12 = fmaf(ky * kz, f, -kx);
    const float r20 = fmaf(kz * kx, f, -ky), r21 = fmaf(kz * ky, f, kx), r22 = fmaf(kz * kz, f, c);
    // padded continuous index of the voxel (texture coordinate + 1/2)
    const float c0x = fmaf(p.x + b.iso[0], g.inv_s, 1.0f);
    const float c0y = fmaf(p.y + b.iso[1], g.inv_s, 1.0f);
    const float c0z = fmaf(p.z + b.iso[2], g.inv_s, 1.0f);
    const __amdgpu_buffer_rsrc_t rs = record_rsrc(rec, g);
    const float sp = a.sp;
    float sum = 0.0f;
    for (int cone = 0; cone < N_CONES; ++cone) {
        const float k1 = a.k1[cone], bm = a.bm[cone];
        const float am_l2sp = -a.am_l2[cone] * a.sp, k2_sp = a.k2[cone] * a.sp, bm_sp = -bm * a.sp;
        float cone_sum_p = 0.0f, cone_sum_s = 0.0f;
        for (int az = 0; az < N_AZ; ++az) {
            const float *t0 = a.t[cone * N_AZ + az];
            // tilted direction, in voxels per cm of ray length (x 10 mm / s)
            const float sc10 = 10.0f * g.inv_s;
            const float ddx = (r00 * t0[0] + r01 * t0[1] + r02 * t0[2]) * sc10;
            const float ddy = (r10 * t0[0] + r11 * t0[1] + r12 * t0[2]) * sc10;
            const float ddz = (r20 * t0[0] + r21 * t0[1] + r22 * t0[2]) * sc10;
            float rp = 0.0f, rsum = 0.0f;
            const float len0 = a.len0[cone];
            const int nsteps = a.nsteps[cone];
            // samples at ray lengths len0, len0 - sp, ... (the float32 counter of the
            // reference, nsteps values >= 0); the positions are monotone in the length,
            // so when the farthest sample is inside the padded grid every sample is
            // with D = max(rho, AIR) sp the reference's updates
            //   R_p <- R_p e + T k1 (1 - e),  e = exp(-am D)
            //   R_s <- R_s (1 - bm D) + T D k2
            // are evaluated as R_p <- T k1 + e (R_p - T k1) and
            // R_s <- R_s + rho' (T k2 sp - bm sp R_s), rho' = max(rho, AIR)
            auto step = [&](float te, float rho) {
                const float rc = fmaxf(rho, kAirDensity);
                const float e = __builtin_amdgcn_exp2f(rc * am_l2sp);
                const float tk = te * k1;
                rp = fmaf(e, rp - tk, tk);
                rsum = fmaf(rc, fmaf(bm_sp, rsum, te * k2_sp), rsum);
            };
            const float fx = fmaf(ddx, len0, c0x), fy = fmaf(ddy, len0, c0y), fz = fmaf(ddz, len0, c0z);
            const bool far_in =
                !live || (fx >= 0.0f && fx < g.hx && fy >= 0.0f && fy < g.hy && fz >= 0.0f && fz < g.hz);
            float rl = len0;
            if (__all(far_in)) {
                // two samples per trip (their loads in flight together)
                int t = 0;
                for (; t + 1 < nsteps; t += 2) {
                    const float rl1 = rl - sp;
                    float te0, rho0, te1, rho1;
                    sample_tr_inside(rs, g, fmaf(ddx, rl, c0x), fmaf(ddy, rl, c0y), fmaf(ddz, rl, c0z), live, up, te0,
                                     rho0);
                    sample_tr_inside(rs, g, fmaf(ddx, rl1, c0x), fmaf(ddy, rl1, c0y), fmaf(ddz, rl1, c0z), live, up,
                                     te1, rho1);
                    step(te0, rho0);
                    step(te1, rho1);
                    rl = rl1 - sp;
                }
                if (t < nsteps) {
                    float te, rho;
                    sample_tr_inside(rs, g, fmaf(ddx, rl, c0x), fmaf(ddy, rl, c0y), fmaf(ddz, rl, c0z), live, up, te,
                                     rho);
                    step(te, rho);
                }
            } else {
                for (int t = 0; t < nsteps; ++t, rl = rl - sp) {
                    float te, rho;
                    sample_tr(rs, g, fmaf(ddx, rl, c0x), fmaf(ddy, rl, c0y), fmaf(ddz, rl, c0z), live, te, rho);
                    step(te, rho);
                }
            }
            cone_sum_p += rp;
            cone_sum_s += rsum;
        }
        sum = fmaf(a.am[cone], cone_sum_p, fmaf(bm, cone_sum_s, sum));
    }
    if (!live) return;
    sum += gld(electron, v);
    const float d = (sum == sum && sum >= 0.0f) ? a.mu_cal * sum * a.mu : 0.0f;
    gst(dose, v, a.accumulate ? fmaf(a.scale, d, gld(dose, v)) : a.scale * d);
}

// Workgroups of a brick launch: 16 per tile of 4x4x4 bricks.
static unsigned brick_blocks(const BrickGrid &bg) {
    const long long tiles = (long long)((bg.bx + 3) / 4) * ((bg.by + 3) / 4) * ((bg.bz + 3) / 4);
    return (unsigned)(tiles * 16);
}

}  // namespace dh

using namespace dh;

extern "C" int64_t dh_photon_record_len(dh_grid g) {
    if (!photon_grid_ok(g, "dh_photon_record_len")) return -1;
    return 12 * record_count(g) + 4 * photon_block_count(g);
}

// The density's block table and run radii, after its records are written.
static int density_blocks(float *records, const dh_grid &g, hipStream_t st) {
    const long long P = record_count(g), nblk = photon_block_count(g);
    const PGrid pg = make_pgrid(g);
    float2 *table = reinterpret_cast<float2 *>(records + 12 * P);
    int *s0 = reinterpret_cast<int *>(records + 12 * P + 2 * nblk);
    hipLaunchKernelGGL(k_ph_density_blocks, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<const float4 *>(records) + 2 * P, table, pg, nblk);
    const int rc = check_launch("k_ph_density_blocks");
    if (rc) return rc;
    return block_radius_passes(table, s0, s0 + nblk, pg.nbx, pg.nby, pg.nbz, st);
}

static int records_ok(const void *rec, const char *what) {
    if (!rec || ((uintptr_t)rec & 31u)) {
        set_error("%s: the record volume must be a non-null, 32-byte aligned device pointer", what);
        return 0;
    }
    return 1;
}

extern "C" int dh_photon_density(const float *density, float *records, dh_grid g, void *stream) {
    if (!photon_grid_ok(g, "dh_photon_density")) return DH_ERR_ARG;
    if (!density || !records_ok(records, "dh_photon_density")) {
        if (!density) set_error("dh_photon_density: null density");
        return DH_ERR_ARG;
    }
    const long long n = record_count(g);
    hipLaunchKernelGGL(k_ph_records_from, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       density, reinterpret_cast<float4 *>(records), make_pgrid(g), n, 1, kAirDensity);
    const int rc = check_launch("k_ph_records_from");
    return rc ? rc : density_blocks(records, g, (hipStream_t)stream);
}

extern "C" int dh_photon_density_from_hu(const float *hu, float *records, dh_grid g, const double *hu_knots,
                                         const double *density_knots, int32_t n_knots, void *stream) {
    if (!photon_grid_ok(g, "dh_photon_density_from_hu")) return DH_ERR_ARG;
    if (!records_ok(records, "dh_photon_density_from_hu")) return DH_ERR_ARG;
    if (!hu || !hu_knots || !density_knots || n_knots < 2 || n_knots > MAX_KNOTS) {
        set_error("dh_photon_density_from_hu: bad arguments (n_knots=%d, max %d)", n_knots, MAX_KNOTS);
        return DH_ERR_ARG;
    }
    Knots kn;
    kn.n = n_knots;
    for (int t = 0; t < MAX_KNOTS; ++t) kn.hu[t] = kn.val[t] = 0.0;
    for (int t = 0; t < n_knots; ++t) {
        kn.hu[t] = hu_knots[t];
        kn.val[t] = density_knots[t];
        if (t && !(hu_knots[t] > hu_knots[t - 1])) {
            set_error("dh_photon_density_from_hu: HU knots must be strictly ascending");
            return DH_ERR_ARG;
        }
    }
    const long long n = record_count(g);
    hipLaunchKernelGGL(k_ph_density_from_hu, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, hu,
                       reinterpret_cast<float4 *>(records), make_pgrid(g), kn, n);
    const int rc = check_launch("k_ph_density_from_hu");
    return rc ? rc : density_blocks(records, g, (hipStream_t)stream);
}

extern "C" int dh_photon_terma_records(const float *terma, float *records, dh_grid g, void *stream) {
    if (!photon_grid_ok(g, "dh_photon_terma_records")) return DH_ERR_ARG;
    if (!terma || !records_ok(records, "dh_photon_terma_records")) {
        if (!terma) set_error("dh_photon_terma_records: null TERMA");
        return DH_ERR_ARG;
    }
    const long long n = record_count(g);
    hipLaunchKernelGGL(k_ph_records_from, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, terma,
                       reinterpret_cast<float4 *>(records), make_pgrid(g), n, 0, 0.0f);
    return check_launch("k_ph_records_from");
}

static bool beam_ok(const dh_photon_beam *beam, const char *what) {
    if (!beam) {
        set_error("%s: null beam", what);
        return false;
    }
    return true;
}

extern "C" int dh_photon_raytrace(const float *records, float *depth, dh_grid g, const dh_photon_beam *beam,
                                  void *stream) {
    if (!photon_grid_ok(g, "dh_photon_raytrace")) return DH_ERR_ARG;
    if (!records_ok(records, "dh_photon_raytrace") || !beam_ok(beam, "dh_photon_raytrace")) return DH_ERR_ARG;
    if (!depth) {
        set_error("dh_photon_raytrace: null output");
        return DH_ERR_ARG;
    }
    const PGrid pg = make_pgrid(g);
    const BrickGrid bg = brick_grid(make_grid(g));
    const unsigned nb = brick_blocks(bg);
    hipLaunchKernelGGL(k_ph_raytrace, dim3(nb), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<const float4 *>(records), depth, pg, make_geom(beam->geom), bg, nb);
    return check_launch("k_ph_raytrace");
}

static bool model_ok(const dh_photon_model *m, const char *what) {
    if (!m || !m->profile || !m->spectrum || m->n_profile < 1 || m->n_spectrum < 0) {
        set_error("%s: bad model (null tables or empty profile)", what);
        return false;
    }
    if (!(m->primary_src_dist > 0.0f) || !(m->primary_src_size > 0.0f) || !(m->scatter_src_size > 0.0f)) {
        set_error("%s: source distance and sizes must be positive", what);
        return false;
    }
    return true;
}

extern "C" int dh_photon_terma(const float *depth, float *terma, float *electron, dh_grid g,
                               const dh_photon_model *model, const dh_photon_beam *beam, void *stream) {
    if (!photon_grid_ok(g, "dh_photon_terma")) return DH_ERR_ARG;
    if (!model_ok(model, "dh_photon_terma") || !beam_ok(beam, "dh_photon_terma")) return DH_ERR_ARG;
    if (!depth || !terma || !electron || (beam->n_pairs > 0 && !beam->mlc) || beam->n_pairs < 0) {
        set_error("dh_photon_terma: null volume or leaf-pair array");
        return DH_ERR_ARG;
    }
    PhotonTermaArgs a;
    a.profile = model->profile;
    a.spectrum = model->spectrum;
    a.mlc = beam->mlc;
    a.n_profile = model->n_profile;
    a.n_spectrum = model->n_spectrum;
    a.n_pairs = beam->n_pairs;
    a.sad_p = model->primary_src_dist;
    a.sad_s = model->scatter_src_dist;
    a.inv_sig_p = 1.0f / (model->primary_src_size * sqrtf(2.0f));
    a.inv_sig_s = 1.0f / (model->scatter_src_size * sqrtf(2.0f));
    a.mlc_distance = model->mlc_distance;
    a.mlc_scale = (model->primary_src_dist - model->mlc_distance) / model->primary_src_dist;
    a.ssw = model->scatter_src_weight;
    a.ea = model->electron_attenuation;
    a.esw = model->electron_src_weight;
    a.el_floor = expf(-model->electron_attenuation * model->electron_fitted_dmax);
    a.el_inv = 1.0f / (1.0f - a.el_floor);
    a.sca = beam->sin_col;
    a.cca = beam->cos_col;
    const PGrid pg = make_pgrid(g);
    const BrickGrid bg = brick_grid(make_grid(g));
    const unsigned nb = brick_blocks(bg);
    hipLaunchKernelGGL(k_ph_terma, dim3(nb), dim3(256), 0, (hipStream_t)stream, depth, terma, electron, pg,
                       make_geom(beam->geom), a, bg, nb);
    return check_launch("k_ph_terma");
}

// Image-space cone directions of a control point: the head-frame set
// (sin th cos ph, sin th sin ph, cos th), ph = j pi / 6, taken to the image
// frame by the inverse of the photon head transform (IMRTClasses.cu:81-118,
// 286-309), in float32 as the reference evaluates them.
static void cone_directions(const dh_photon_model &m, const dh_photon_beam &b, float sp, PhotonCccArgs &a) {
    const float pi = 3.14159265358979f;
    const dh_beam_geom &gm = b.geom;
    for (int c = 0; c < N_CONES; ++c) {
        const float th = m.kernel[c][0] * pi / 180.0f;
        const float st = sinf(th), ct = cosf(th);
        const float Am = m.kernel[c][1], am = m.kernel[c][2], Bm = m.kernel[c][3], bm = m.kernel[c][4];
        a.k1[c] = st * (Am / (am * am));
        a.k2[c] = st * (Bm / bm);
        a.am[c] = am;
        a.bm[c] = bm;
        a.am_l2[c] = am * 1.44269504088896341f;
        a.len0[c] = m.kernel[c][5];
        int n = 0;
        for (float rl = m.kernel[c][5]; rl >= 0.0f; rl = rl - sp) ++n;  // float32 as the reference's counter
        a.nsteps[c] = n;
        for (int j = 0; j < N_AZ; ++j) {
            const float phi = (float)j * pi / 6.0f;
            const float hx = st * cosf(phi), hy = st * sinf(phi), hz = ct;
            // head -> image: undo the collimator, gantry and couch rotations
            const float xz = -hx, yz = hz, zz = hy;
            const float xc = xz * b.cos_col + zz * b.sin_col;
            const float yc = yz;
            const float zc = -xz * b.sin_col + zz * b.cos_col;
            const float xg = xc * gm.cos_g - yc * gm.sin_g;
            const float yg = xc * gm.sin_g + yc * gm.cos_g;
            const float zg = zc;
            float *t = a.t[c * N_AZ + j];
            t[0] = xg * gm.cos_c + zg * gm.sin_c;
            t[1] = yg;
            t[2] = -xg * gm.sin_c + zg * gm.cos_c;
        }
    }
    const float n = sqrtf(gm.src[0] * gm.src[0] + gm.src[1] * gm.src[1] + gm.src[2] * gm.src[2]);
    for (int d = 0; d < 3; ++d) a.ax[d] = gm.src[d] / n;
}

extern "C" int dh_photon_ccc(const float *records, const float *terma, const float *electron, float *dose, dh_grid g,
                             const dh_photon_model *model, const dh_photon_beam *beam, float scale, int32_t accumulate,
                             void *stream) {
    if (!photon_grid_ok(g, "dh_photon_ccc")) return DH_ERR_ARG;
    if (!records_ok(records, "dh_photon_ccc") || !beam_ok(beam, "dh_photon_ccc")) return DH_ERR_ARG;
    if (!terma || !electron || !dose) {
        set_error("dh_photon_ccc: null volume");
        return DH_ERR_ARG;
    }
    if (!model_ok(model, "dh_photon_ccc")) return DH_ERR_ARG;
    const float sp = g.spacing / 10.0f;
    for (int c = 0; c < N_CONES; ++c) {
        // the ray-length counter must reach below zero in a bounded number of steps
        // a negative length takes no sample, as the reference's loop; NaN or more than 4096 steps is refused
        const float L = model->kernel[c][5];
        if (L != L || L / sp >= 4096.0f) {
            set_error("dh_photon_ccc: cone %d ray length %g cm is not a number or beyond 4096 steps of %g cm", c,
                      (double)L, (double)sp);
            return DH_ERR_ARG;
        }
    }
    PhotonCccArgs a;
    cone_directions(*model, *beam, sp, a);
    a.sp = sp;
    a.mu_cal = model->mu_cal;
    a.mu = beam->mu;
    a.scale = scale;
    a.accumulate = accumulate ? 1 : 0;
    const PGrid pg = make_pgrid(g);
    const BrickGrid bg = brick_grid(make_grid(g));
    const unsigned nb = brick_blocks(bg);
    hipLaunchKernelGGL(k_ph_ccc, dim3(nb), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<const float4 *>(records), terma, electron, dose, pg, make_geom(beam->geom), a,
                       bg, nb);
    return check_launch("k_ph_ccc");
}
