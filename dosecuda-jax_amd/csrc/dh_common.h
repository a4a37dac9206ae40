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
    return r;
}

// Voxel bricks: 4x4x4 voxels, one wave64 per brick, lane = bx + 4*by + 16*bz.
constexpr int BRICK = 4;
constexpr int WAVE = 64;

struct BrickGrid {
    int bx, by, bz;      // bricks per axis
    __host__ __device__ long long count() const { return (long long)bx * by * bz; }
};

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
// wave64 helpers
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

}  // namespace dh
