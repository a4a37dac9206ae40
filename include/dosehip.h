/*
 * dosehip.h - C ABI of the gfx950 proton pencil-beam and photon
 * collapsed-cone dose library (libdosehip.so).
 *
 * Replaces the proton half of the reference extension `dose_kernels`
 * (DoseCUDA/dose_kernels/dosemodule.cu:104-374, IMPTClasses.cu:274-340):
 *   proton_raytrace_cuda  ->  dh_rlsp_pad / dh_rlsp_from_hu, dh_wet_raymarch, dh_wet_smooth
 *   proton_spot_cuda      ->  dh_pb_forward
 *   (jax.grad of DoseCUDA/Jax/impt_jax.py, no CUDA counterpart) -> dh_pb_adjoint
 *
 * Conventions
 *  - All array arguments are DEVICE pointers (HIP global memory on the current
 *    device), borrowed for the duration of the stream work; the library never
 *    allocates, frees or synchronises.  Work is enqueued on `stream`
 *    (a hipStream_t passed as void*, NULL = default stream).
 *  - Volumes are float32, indexed v = i + nx*(j + ny*k) with i along x
 *    (numpy axis 2 of a (z, y, x) array), as in CudaClasses.cuh:189-191.
 *  - The ray-march reads RLSP in a "quad" layout: element (i, j, k) with
 *    i < nx, j < ny, k <= nz, index i + nx*(j + ny*k), is the float4
 *    (a, b - a, c - a, (d - c) - (b - a)) of a = rho(i,j,k), b = rho(i+1,j,k),
 *    c = rho(i,j+1,k), d = rho(i+1,j+1,k), rho = 0 past the high end of each
 *    axis (the bilinear form of the slice-k cell face); 4*nx*ny*(nz+1)
 *    floats, 16-byte aligned, followed by a table of B =
 *    ceil(nx/4)*ceil(ny/4)*ceil(nz/4) float pairs, one per 4x4x4 block of
 *    trilinear cells (index bx + nbx*(by + nby*bz)): the rho every corner of
 *    the block's cells holds when they are all equal, NaN otherwise, and for
 *    such a uniform block its run radius r (0..15, a float): every block of the
 *    grid within Chebyshev distance r of it is uniform with the same rho (the
 *    ray-march adds the run of samples that stays in that cube at once); then
 *    2*B ints of scratch.  dh_rlsp_from_hu / dh_rlsp_pad write all of it;
 *    dh_padded_len gives the total.
 *  - Return value 0 = success; negative = error, message via dh_last_error().
 *  - Entry points are re-entrant; per-thread error text.
 */
#ifndef DOSEHIP_H
#define DOSEHIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DH_OK            0
#define DH_ERR_ARG      -1   /* invalid argument (sizes, null pointers, workspace too small) */
#define DH_ERR_HIP      -2   /* HIP runtime error from a launch */

/* Dose grid: sizes along x (numpy axis 2), y (axis 1), z (axis 0); isotropic
 * spacing in mm (only spacing[0] is used, dosemodule.cu:146). */
typedef struct dh_grid {
    int32_t nx, ny, nz;
    float spacing;
} dh_grid;

/* Beam geometry (CudaBeam, CudaClasses.cu:17-51), all float32:
 *  iso      isocentre minus grid origin (mm)            dosemodule.cu:160-164
 *  src      source position relative to the isocentre  = R_couch(R_gantry((0, SAD, 0)))
 *  sin_g/cos_g  of the adjusted gantry angle (gantry + 180 deg, dosemodule.cu:152)
 *  sin_c/cos_c  of the couch angle
 *  vsadx/vsady  virtual source-axis distances (mm) */
typedef struct dh_beam_geom {
    float iso[3];
    float src[3];
    float sin_g, cos_g, sin_c, cos_c;
    float vsadx, vsady;
} dh_beam_geom;

/* One spot, sorted by energy id (Spot, IMPTClasses.cuh:8-15). */
typedef struct dh_spot {
    float x, y;      /* mm at the isocentre plane, beam's-eye view */
    float mu;        /* weight */
    int32_t energy_id;
} dh_spot;

/* One energy layer = a run of equal energy ids in the sorted spot array
 * (Layer, IMPTClasses.cuh:17-26; built like importLayers, IMPTClasses.cu:37-67).
 * The kernels use a layer only if its energy id is inside the model, its
 * range [spot_start, spot_start + n_spots) is non-empty and inside the spot
 * array, and that range intersects no earlier usable-range layer of the
 * array; other layers add no dose, and spots in no usable layer get zero
 * gradients (importLayers only builds disjoint ascending runs). */
typedef struct dh_layer {
    int32_t spot_start;
    int32_t n_spots;
    int32_t energy_id;
    float r80;       /* mm */
    float energy;    /* MeV */
} dh_layer;

/* Beam-model lookup tables on the device (IMPTBeamModel, plan_impt.py:63-140).
 *  idd_sigma  float2 [n_energies][lut_len] = (IDD, sigma_ms) interleaved
 *  depths     float  [n_energies][lut_len] (mm), ascending per energy
 *  dvp        float  [n_energies][dvp_len] = (E, R80, A, B, C, ...)
 *  depth_uniform != 0 asserts depths[e][k] == depth0 + k*ddepth for every e, k
 *  (true of the shipped tables: 0..399 mm); the lookup then indexes directly
 *  instead of bisecting (lowerBound, CudaClasses.cuh:130-144). */
typedef struct dh_beam_model {
    const float *idd_sigma;
    const float *depths;
    const float *dvp;
    int32_t n_energies;
    int32_t lut_len;
    int32_t dvp_len;
    int32_t depth_uniform;
    float depth0;
    float ddepth;
} dh_beam_model;

/* Spots of one beam on the device. */
typedef struct dh_spot_set {
    const dh_spot *spots;    /* [n_spots], sorted by energy id */
    const dh_layer *layers;  /* [n_layers] */
    int32_t n_spots;
    int32_t n_layers;
} dh_spot_set;

/* Library version, e.g. 100 for 0.1.0. */
int dh_version(void);

/* Copy the last error message of the calling thread into buf (NUL-terminated,
 * truncated to len).  Returns the length of the full message. */
int dh_last_error(char *buf, size_t len);

/* Number of floats of a quad RLSP volume with its block table and scratch:
 * 4*nx*ny*(nz+1) + 4*ceil(nx/4)*ceil(ny/4)*ceil(nz/4). */
int64_t dh_padded_len(dh_grid g);

/* HU -> RLSP, piecewise linear through (hu_knots, rlsp_knots), clamped at both
 * ends (np.interp, plan_impt.py:158-168), written in the quad layout.
 * hu: [nz*ny*nx] device; knots: HOST float64 arrays (as parsed from
 * HU_RLSP.csv), strictly ascending hu, 2 <= n_knots <= 32.  The interpolation
 * runs in float64 and is rounded to float32, like the reference's
 * np.interp(...).astype(np.single). */
int dh_rlsp_from_hu(const float *hu, float *rlsp_padded, dh_grid g,
                    const double *hu_knots, const double *rlsp_knots, int32_t n_knots,
                    void *stream);

/* Copy a plain RLSP volume (nz*ny*nx floats) into the quad layout. */
int dh_rlsp_pad(const float *rlsp, float *rlsp_padded, dh_grid g, void *stream);

/* Raw water-equivalent thickness (cm) of every voxel: 1 mm steps from the voxel
 * centre toward the source, zero-padded trilinear RLSP, samples with continuous
 * index outside [0, n) dropped, WET = -0.05 + 0.1 * sum(max(rlsp, 0))
 * (rayTraceKernel, CudaClasses.cu:180-228; bounds as impt_jax.py:262-277). */
int dh_wet_raymarch(const float *rlsp_padded, float *wet_raw, dh_grid g,
                    const dh_beam_geom *geom, void *stream);

/* Lateral smoothing of raw WET: mean of the voxel and the nearest voxels
 * (round half to even) at dr = 1..9 mm, dr < 10*WET, along 6 beam's-eye-view
 * directions k*60 deg that fall inside the grid (smoothRayKernel,
 * IMPTClasses.cu:139-202; impt_jax.py:317-418). */
int dh_wet_smooth(const float *wet_raw, float *wet, dh_grid g,
                  const dh_beam_geom *geom, void *stream);

/* Bytes of device workspace dh_pb_forward / dh_pb_adjoint need for n_spots
 * spots in n_layers layers (256-byte aligned pointer required). */
int64_t dh_pb_workspace_size(int32_t n_spots, int32_t n_layers, dh_grid g);

/* Double-Gaussian pencil-beam dose of one beam (pencilBeamKernel,
 * IMPTClasses.cu:204-271; impt_jax.py:425-541).  wet: smoothed WET (cm).
 * dose: overwritten (accumulate == 0) or added to (accumulate != 0).
 * cutoff_sigma > 0 skips (voxel brick, spot) pairs whose beam's-eye-view
 * distance exceeds cutoff_sigma * (largest halo sigma of the brick); each
 * skipped term is below exp(-cutoff_sigma^2 / 2) of that spot's peak.
 * cutoff_sigma <= 0 evaluates every pair as the reference does. */
int dh_pb_forward(const float *wet, float *dose, dh_grid g, const dh_beam_geom *geom,
                  const dh_beam_model *model, const dh_spot_set *spots,
                  float cutoff_sigma, int32_t accumulate,
                  void *workspace, int64_t workspace_bytes, void *stream);

/* Adjoint of dh_pb_forward: given grad_dose = dL/dDose (per voxel), writes
 *   grad_mu[s]      = dL/dmu_s
 *   grad_xy[2s+0/1] = dL/dx_s, dL/dy_s
 * for the sorted spots.  WET-derived per-voxel terms are recomputed, not stored.
 * Deterministic (fixed-order two-pass reduction, no atomics). */
int dh_pb_adjoint(const float *wet, const float *grad_dose, dh_grid g,
                  const dh_beam_geom *geom, const dh_beam_model *model,
                  const dh_spot_set *spots, float cutoff_sigma,
                  float *grad_mu, float *grad_xy,
                  void *workspace, int64_t workspace_bytes, void *stream);

/* dh_pb_adjoint without a workspace argument.  No counterpart in the
 * reference: this entry point is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  It allocates nothing: the beam's workspace (spot and
 * layer records, adjoint partial sums, brick summaries) lives in an array of
 * DH_NOALLOC_SCRATCH_BYTES that the library holds on every device it runs on
 * until the process ends.  Same kernels and results as dh_pb_adjoint (bit for
 * bit when dh_pb_adjoint's layout fits the array; otherwise fewer brick slabs,
 * so the partial sums add in another order).  Returns DH_ERR_ARG, with a
 * message, when even one slab does not fit (about
 * 48 * n_spots + 64 * n_layers + 8 * bricks bytes, bricks = grid voxels / 64).
 * Every call on a device shares that device's array: calls that overlap on
 * one device (other streams, other threads) race and must be serialised by
 * the caller. */
#define DH_NOALLOC_SCRATCH_BYTES (64LL << 20)
int dh_pb_adjoint_noalloc(const float *wet, const float *grad_dose, dh_grid g,
                          const dh_beam_geom *geom, const dh_beam_model *model,
                          const dh_spot_set *spots, float cutoff_sigma,
                          float *grad_mu, float *grad_xy, void *stream);

/* Work accounting for dh_pb_forward / dh_pb_adjoint with the same arguments
 * (no dose or gradient is computed).  Adds to the DEVICE array counts[4]:
 *   counts[0]  forward pairs the reference evaluates: voxels with
 *              wet <= 1.1 R80 of a layer x that layer's spots
 *   counts[1]  forward pairs evaluated here after the cutoff_sigma cull
 *   counts[2]  adjoint pairs of the same kind over voxels with grad_dose != 0
 *   counts[3]  adjoint pairs evaluated here after the cull
 * grad_dose may be NULL (forward counts only). */
int dh_pb_count_pairs(const float *wet, const float *grad_dose, dh_grid g,
                      const dh_beam_geom *geom, const dh_beam_model *model,
                      const dh_spot_set *spots, float cutoff_sigma,
                      unsigned long long *counts,
                      void *workspace, int64_t workspace_bytes, void *stream);

/* Spot weights and positions from arrays in the caller's (input) row order:
 * for i < n, with r = order[i] (device int64, a permutation of 0..n-1:
 * sorted position -> input row),
 *   spots[i].mu = mu[r]; spots[i].x = xy[2r]; spots[i].y = xy[2r + 1].
 * mu or xy may be NULL (that field is kept); energy ids are never changed.
 * An entry with r outside [0, n) is left as it is.  One launch per call; this
 * is the per-call AoS fill of proton_spot (dosemodule.cu:204-225) for spots
 * already sorted on the device, so moving spots needs no host round trip. */
int dh_spot_params(dh_spot *spots, const int64_t *order, const float *mu, const float *xy, int32_t n,
                   void *stream);

/* ------------------------------------------------------------------------
 * Batched plans.  Many plans (each its own dose grid, 1..n beams, any spot
 * counts) computed together: every phase is one launch over the work items of
 * all beams (bricks, spot chunks x brick slabs), so small and large plans fill
 * the GPU together.  The reference computes one beam per call
 * (computeIMPTPlan, plan_impt.py:170-206; proton_spot_cuda, IMPTClasses.cu:301-340);
 * a batch gives the same per-plan dose (beam doses added in beam order), and
 * the same gradients as dh_pb_adjoint per beam.
 *
 * Use: dh_batch_workspace_size -> allocate -> dh_batch_setup (uploads the
 * descriptor tables, synchronises `stream`, not capturable) -> any number of
 * dh_batch_wet / dh_batch_forward / dh_batch_adjoint calls.  Those three only
 * enqueue kernels that read the tables in the workspace: no allocation, copy,
 * memset or synchronisation, so they may be captured in a HIP graph and
 * replayed.  Spot weights and positions are read from the device spot arrays
 * by every forward / adjoint call, so updating them between replays is
 * enough; sizes, pointers and geometry are fixed at setup.
 */

/* A plan of a batch. */
typedef struct dh_batch_plan {
    dh_grid grid;
    const float *rlsp_padded;  /* quad RLSP (dh_padded_len floats, 16-byte aligned); dh_batch_wet input */
    float *dose;               /* [nvox] plan dose, written by dh_batch_forward */
    const float *grad_dose;    /* [nvox] dL/dDose, read by dh_batch_adjoint (may be NULL if unused) */
    int32_t beam_start;        /* beams [beam_start, beam_start + n_beams) of the beam array; */
    int32_t n_beams;           /* plans list their beams in order, without gaps */
} dh_batch_plan;

/* A beam of a batch. */
typedef struct dh_batch_beam {
    dh_beam_geom geom;
    dh_beam_model model;       /* device lookup tables (may be shared by beams) */
    dh_spot_set spots;         /* device spot and layer records, as for dh_pb_forward */
    float *wet;                /* [nvox] smoothed WET (cm): written by dh_batch_wet, read by forward/adjoint */
    float *wet_raw;            /* [nvox] raw WET, or NULL for workspace scratch */
    float *dose;               /* [nvox] this beam's dose, or NULL: the plan dose for a one-beam plan,
                                  workspace scratch otherwise */
    float *grad_mu;            /* [n_spots] dL/dmu (sorted spot order) */
    float *grad_xy;            /* [2 n_spots] dL/dx, dL/dy */
} dh_batch_beam;

/* Schedules, chosen per batch (e.g. by an autotuner):
 *   adj_cs          spots per adjoint workgroup: 32, 64, 128 (packed: up to 128 spots of a layer, each
 *                   brick's reachable spots compacted onto the lanes), or 0 (128 with a cutoff, else 32
 *                   or 64 from the average layer size)
 *   adj_wg_target   workgroups per beam to aim for (bricks are cut into slabs to reach it)
 *   adj_min_bricks  fewest bricks per slab
 *   adj_max_slabs   most slabs per beam (<= 65535)
 *   fwd_blocks      forward bricks: 0 = one wave per brick, n > 0 = a persistent grid of n 256-thread
 *                   blocks striding over the bricks of all beams
 *   ray_blocks      the same for the ray-march
 *   ray_shape       the ray-march's voxel patch per wave (a workgroup takes the 16x4x4 voxels of four
 *                   consecutive bricks of a row): 0 = 4x4x4, 1 = 8x4x2, 2 = 16x4x1, 3 = 16x2x2,
 *                   4 = 16x1x4, 5 = 8x2x4; bricks not four in one row go one brick per wave
 * Zero fields take the defaults (0, 32768, 64, 1024, 0, 0, 0).  Results do not depend on fwd_blocks,
 * ray_blocks and ray_shape; the adjoint fields change only the order of partial sums. */
typedef struct dh_batch_tune {
    int32_t adj_cs;
    int32_t adj_wg_target;
    int32_t adj_min_bricks;
    int32_t adj_max_slabs;
    int32_t fwd_blocks;
    int32_t ray_blocks;
    int32_t ray_shape;         /* ray-march voxel patch of a wave: 0 = 4x4x4 (one brick), 1 = 8x4x2,
                                  2 = 16x4x1 (four bricks of a row, one z-layer), 3 = 16x2x2, 4 = 16x1x4,
                                  5 = 8x2x4 */
} dh_batch_tune;

/* Filled by dh_batch_setup; plain data owned by the caller, read-only to it.
 * The fields describe the workspace layout and the item counts of each phase. */
typedef struct dh_batch {
    void *ws;
    int64_t ws_bytes;
    int32_t n_plans, n_beams;
    float cutoff_sigma;
    int32_t magic;
    int64_t n_bricks;          /* ray-march / forward items (bricks of all beams) */
    int64_t n_smooth_blocks;   /* smoothing items of all beams: 32x4x8-voxel tiles, or 256-voxel blocks
                                  when a beam's stencil halo is too wide for the tiled kernel */
    int64_t n_spots, n_layers; /* prepared spot and layer records */
    int64_t n_sum_blocks;      /* 1024-voxel blocks of plans with several beams */
    int64_t n_partials;        /* adjoint partial records */
    int64_t n_adj32, n_adj64;  /* adjoint workgroups of 32- and 64-spot chunks */
    int64_t n_adj128;          /* adjoint workgroups of packed chunks (up to 128 spots, lanes filled per brick) */
    int64_t n_reduce_blocks;   /* 64-spot reduction blocks */
    int64_t tab[16];           /* byte offsets of the device tables in ws; tab[13] = end of the tables
                                  uploaded at setup (the rest of ws is scratch the phases rewrite);
                                  tab[14] = flags, not an offset: bit 3 = tiled smoothing, whose LDS
                                  box (floats) is in bits 32 and up */
    int32_t fwd_blocks, ray_blocks;   /* brick-kernel schedules from dh_batch_tune */
    int32_t ray_shape, reserved;      /* ray-march patch shape from dh_batch_tune; reserved = 0 */
    int32_t prepared;          /* spot / layer records prepared by a phase call (DH_BATCH_REUSE_PREP needs it) */
    int32_t n_summary_rows;    /* brick-summary items: rows of bricks along x of the beams with spots */
    void *const *events;       /* kernel timing events (dh_batch_timing), NULL = off */
} dh_batch;

/* Workspace bytes for a batch (negative on invalid descriptors; see dh_last_error). */
int64_t dh_batch_workspace_size(const dh_batch_plan *plans, int32_t n_plans,
                                const dh_batch_beam *beams, int32_t n_beams,
                                const dh_batch_tune *tune);

/* Validate the descriptors, lay out the workspace and upload the per-beam
 * tables (ray-march, smoothing stencil, pencil-beam and adjoint contexts).
 * cutoff_sigma as for dh_pb_forward.  tune may be NULL (defaults). */
int dh_batch_setup(dh_batch *batch, const dh_batch_plan *plans, int32_t n_plans,
                   const dh_batch_beam *beams, int32_t n_beams, const dh_batch_tune *tune,
                   float cutoff_sigma, void *workspace, int64_t workspace_bytes, void *stream);

/* Flags of the phase calls.
 * DH_BATCH_REUSE_PREP  reuse the spot and layer records the previous phase call
 *                      prepared (spot arrays unchanged since) instead of
 *                      preparing them again.  An error (DH_ERR_ARG) until some
 *                      call on this batch has prepared them: dh_batch_forward,
 *                      dh_batch_adjoint, or dh_batch_wet with
 *                      DH_BATCH_WET_REACH, each without this flag.
 * DH_BATCH_WET_REACH   (dh_batch_wet, cutoff_sigma > 0) compute WET only in
 *                      bricks that some layer of the beam can reach with the
 *                      CURRENT spot positions, and where smoothing reads raw
 *                      WET for them; elsewhere WET = 1e30 (beyond every
 *                      range).  Dose and gradients of forward / adjoint calls
 *                      with the same spot positions equal those from the full
 *                      WET bit for bit; after spot positions move, call
 *                      dh_batch_wet again.  The volume is not the reference's
 *                      WET output (use the full pass for that). */
#define DH_BATCH_REUSE_PREP 1
#define DH_BATCH_WET_REACH  2

/* WET of every beam: ray-march then smoothing (dh_wet_raymarch + dh_wet_smooth per beam). */
int dh_batch_wet(dh_batch *batch, int32_t flags, void *stream);

/* Kernel timing for eager calls (not for graph capture).  events is a HOST
 * array of 2*DH_TIME_N hipEvent_t (created by the caller; NULL entries are
 * skipped) that stays valid while timing is on: events[2k] and events[2k+1]
 * are recorded on the stream right before and right after the launches of
 * timed kernel k.  events == NULL turns timing off. */
#define DH_TIME_RAYMARCH 0  /* k_wet_raymarch_batch */
#define DH_TIME_SMOOTH   1  /* k_wet_smooth_tile_batch (k_wet_smooth_batch for wide halos) */
#define DH_TIME_FORWARD  2  /* k_pb_forward_batch */
#define DH_TIME_ADJOINT  3  /* k_pb_adjoint_batch (both chunk sizes) */
#define DH_TIME_NEED     4  /* k_wet_need_batch + k_wet_dilate_batch */
#define DH_TIME_N        5
int dh_batch_timing(dh_batch *batch, void *const *events);

/* Dose of every plan (dh_pb_forward per beam, beam doses added in beam order). */
int dh_batch_forward(dh_batch *batch, int32_t flags, void *stream);

/* Spot gradients of every beam from its plan's grad_dose (dh_pb_adjoint per beam).
 * The three phase calls update batch->prepared (host side), hence the non-const batch. */
int dh_batch_adjoint(dh_batch *batch, int32_t flags, void *stream);

/* ------------------------------------------------------------------------
 * Photon collapsed-cone convolution (CCC) dose of one control point.
 * Replaces the photon half of `dose_kernels`:
 *   photon_dose_cuda (dosemodule.cu:377-557, photon_dose_cuda
 *   IMRTClasses.cu:353-405)  ->  dh_photon_density[_from_hu], dh_photon_raytrace,
 *                                dh_photon_terma, dh_photon_terma_records,
 *                                dh_photon_ccc
 *
 * Density and TERMA are read through a "record" volume in place of the
 * reference's two border-addressed 3-D textures (TextureClasses.cu:14-41),
 * over the grid padded by one voxel on every side: P = (nx+2)*(ny+2)*(nz+2)
 * cells, index p = i' + (nx+2)*(j' + (ny+2)*k').  Floats 8p .. 8p+7 hold the
 * bilinear form of TERMA on the slice-k' face of cell p, then that of density;
 * floats 8P + 4p .. 8P + 4p + 3 hold the density form again (the ray-trace
 * reads only these).  The form of field f is (a, b-a, c-a, (d-c)-(b-a)) with
 * a = f(i',j',k'), b = f(i'+1,j',k'), c = f(i',j'+1,k'), d = f(i'+1,j'+1,k'),
 * where f(i',j',k') is the voxel value at (i'-1, j'-1, k'-1) inside the grid
 * and the texture's border value outside it (TERMA 0, density AIR_DENSITY =
 * 0.0012, IMRTClasses.cuh:6).  A sample at continuous voxel index c (texture
 * coordinate c + 1/2) is then the blend of two cells' forms at c + 1, with
 * the border values wherever the texture would read only border texels.
 * After them, from float 12 P on, a block table of the density: one float
 * pair per block of 4x4x4 sample cells (padded cells 4b .. 4b+3 per axis,
 * blocks (n+4)/4 per axis, x fastest), the density every corner of those
 * cells holds (NaN where they differ) and the block's run radius, then two
 * ints per block of scratch; the density entry points write it, and the
 * ray-trace adds the samples that stay in a uniform cube of blocks at once.
 * dh_photon_record_len gives the floats of a record volume (12 P + 4 blocks;
 * 32-byte aligned, below 2^32 bytes).
 * ---------------------------------------------------------------------- */

/* MLC leaf pair (MLCPair, IMRTClasses.cuh:8-15), mm at the isocentre plane. */
typedef struct dh_mlc_pair {
    float x1, x2;        /* leaf tips along x */
    float y_offset;      /* pair centre along y */
    float y_width;
} dh_mlc_pair;

/* Photon beam model (IMRTBeam::Model, IMRTClasses.cuh:22-58).  Tables are
 * device pointers; the six collapsed cones are passed by value. */
typedef struct dh_photon_model {
    const float *profile;    /* [3][n_profile]: radius (mm, ascending), intensity, softening */
    const float *spectrum;   /* [3][n_spectrum]: attenuation (cm^2/g), primary weight, scatter weight */
    int32_t n_profile;
    int32_t n_spectrum;
    float kernel[6][6];      /* per cone: theta (deg), Am, am, Bm, bm, ray length (cm) */
    float mu_cal;
    float primary_src_dist, scatter_src_dist;   /* mm */
    float primary_src_size, scatter_src_size;   /* Gaussian sigma, mm */
    float mlc_distance;                         /* mm from the isocentre toward the source */
    float scatter_src_weight;
    float electron_attenuation, electron_src_weight, electron_fitted_dmax;
} dh_photon_model;

/* One control point (IMRTBeam, IMRTClasses.cuh:20-83; dosemodule.cu:493-536).
 * geom: as for protons with the source at primary_src_dist (vsadx / vsady
 * unused); sin_col / cos_col of the collimator angle; mlc: device array. */
typedef struct dh_photon_beam {
    dh_beam_geom geom;
    float sin_col, cos_col;
    const dh_mlc_pair *mlc;
    int32_t n_pairs;
    float mu;
} dh_photon_beam;

/* Floats of a record volume for grid g. */
int64_t dh_photon_record_len(dh_grid g);

/* Density half of the records and its block table from a density volume
 * (nz*ny*nx floats; the reference's DensityTexture, IMRTClasses.cu:394). */
int dh_photon_density(const float *density, float *records, dh_grid g, void *stream);

/* The same from HU, density = np.interp(HU, hu_knots, density_knots) rounded
 * to float32 (IMRTDoseGrid.DensityFromHU, plan_imrt.py:114-124); n_knots <= 32. */
int dh_photon_density_from_hu(const float *hu, float *records, dh_grid g, const double *hu_knots,
                              const double *density_knots, int32_t n_knots, void *stream);

/* Radiological depth (g/cm^2) of every voxel through the density
 * (rayTraceKernel, CudaClasses.cu:180-228): samples 1 mm apart from the voxel
 * toward the source while the texture coordinate is inside the grid,
 * -0.05 + 0.1 * sum max(density, 0). */
int dh_photon_raytrace(const float *records, float *depth, dh_grid g, const dh_photon_beam *beam, void *stream);

/* TERMA and electron contamination of every voxel (termaKernel,
 * IMRTClasses.cu:209-250): off-axis profile, Gaussian primary and scatter
 * sources seen through the leaf pairs, spectrum attenuation at the depth. */
int dh_photon_terma(const float *depth, float *terma, float *electron, dh_grid g, const dh_photon_model *model,
                    const dh_photon_beam *beam, void *stream);

/* TERMA half of the records from a TERMA volume (the reference's TERMATexture,
 * IMRTClasses.cu:399). */
int dh_photon_terma_records(const float *terma, float *records, dh_grid g, void *stream);

/* Collapsed-cone dose (cccKernel, IMRTClasses.cu:252-350): per voxel with
 * TERMA > 0.01, 6 cones x 12 azimuths tilted toward the source, each ray
 * marched from its length down to the voxel in steps of one spacing;
 * dose = mu_cal * mu * (sum + electron) where that is a non-negative number,
 * else 0.  Writes dose[v] = scale * that (accumulate == 0) or adds it
 * (accumulate != 0). */
int dh_photon_ccc(const float *records, const float *terma, const float *electron, float *dose, dh_grid g,
                  const dh_photon_model *model, const dh_photon_beam *beam, float scale, int32_t accumulate,
                  void *stream);

#ifdef __cplusplus
}
#endif
#endif /* DOSEHIP_H */
