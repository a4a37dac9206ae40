// Batched plans: descriptor tables in the workspace and the phase launches.
//
// The reference computes one beam per call and one plan per computeIMPTPlan
// (plan_impt.py:170-206, proton_spot_cuda IMPTClasses.cu:301-340).  Here a
// batch of plans with uneven grids and spot counts runs each phase as one
// launch over the items of every beam:
//   wet      (reach mode: the need mask and its dilation, one lane per brick)
//            ray-march bricks of all beams, then smoothing tiles of all beams
//            (256-voxel blocks when a stencil halo is too wide for LDS tiles)
//   forward  spot records, layer records, dose bricks of all beams, then the
//            fixed-order sum of the beam doses where beams write volumes of
//            their own (plans with several beams)
//   adjoint  brick summaries (one workgroup per row of bricks), (chunk, slab)
//            workgroups of all beams (one launch per chunk size: 32, 64, or
//            packed 128), per-spot reduction
// A work item finds its beam by binary search in a table of item offsets and
// reads the beam's context (pointers, geometry, schedule) from a device table,
// so the launches need no host-side state beyond item counts and can be
// captured in a HIP graph.
#include <algorithm>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "dh_common.h"

namespace dh {

static constexpr int kMagic = 0x44484231;  // "DHB1"
static constexpr long long kAlign = 256;
static long long align_up(long long x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Device tables, as byte offsets into the workspace (dh_batch.tab).
enum Table {
    T_CTX = 0,      // BeamCtx [n_beams]
    T_PLANS,        // PlanCtx [n_plans]
    T_BRICK_OFF,    // long long [n_beams]: ray-march / forward / need-mask items
    T_SMOOTH_OFF,   // smoothing tiles (F_SMOOTH_TILE) or 256-voxel blocks
    T_SPOT_OFF,     // spot records
    T_LAYER_OFF,    // layer records
    T_ADJ128_OFF,   // adjoint workgroups, packed 128-spot chunks
    T_ADJ32_OFF,    // adjoint workgroups, 32-spot chunks
    T_ADJ64_OFF,    // adjoint workgroups, 64-spot chunks
    T_REDUCE_OFF,   // reduction blocks
    T_SUM_OFF,      // long long [n_plans]: plan-sum blocks
    T_STENCIL,      // Stencil [n_beams]
    T_TOUCH,        // int [n_beams][MAX_TOUCH]: brick offsets smoothing reads
    T_HEADER_END,   // first byte after the tables (uploaded at setup)
    T_FLAGS,        // not an offset: capability bits below; bits 32+: tiled-smoothing box (floats)
    T_ROW_OFF,      // long long [n_beams]: brick-summary rows (rows of beams with spots)
    N_TABLES
};
static_assert(N_TABLES <= 16, "dh_batch.tab holds 16 entries");
static_assert(T_HEADER_END == 13, "dh_batch.tab[13] is documented as the end of the uploaded tables");

enum Flags { F_RLSP = 1, F_GRAD_DOSE = 2, F_GRAD_OUT = 4, F_SMOOTH_TILE = 8 };

static long long bricks_of(const dh_grid &g) {
    return (long long)((g.nx + 3) / 4) * ((g.ny + 3) / 4) * ((g.nz + 3) / 4);
}
static long long nvox_of(const dh_grid &g) { return (long long)g.nx * g.ny * g.nz; }

static AdjTune tune_from(const dh_batch_tune *t) {
    AdjTune r = default_tune();
    if (t) {
        if (t->adj_cs) r.cs = t->adj_cs;
        if (t->adj_wg_target) r.wg_target = t->adj_wg_target;
        if (t->adj_min_bricks) r.min_bricks = t->adj_min_bricks;
        if (t->adj_max_slabs) r.max_slabs = t->adj_max_slabs;
    }
    return r;
}

// Layout of a batch: host copies of the tables plus the workspace size.
struct Layout {
    std::vector<BeamCtx> ctx;
    std::vector<PlanCtx> plans;
    std::vector<long long> brick, smooth, smooth_tiles, spot, layer, adj128, adj32, adj64, reduce, sum, rows;
    std::vector<Stencil> stencil;
    std::vector<int> touch;
    long long tab[N_TABLES];
    long long n_bricks = 0, n_smooth = 0, n_spots = 0, n_layers = 0, n_sum = 0, n_part = 0, n_adj32 = 0,
              n_adj64 = 0, n_adj128 = 0, n_reduce = 0, n_rows = 0, n_tiles = 0;
    int box_max = 0;  // largest tiled-smoothing box of the batch (floats)
    long long bytes = 0;
    int flags = 0;
};

// Validate the descriptors and lay out the workspace at `base` (nullptr: sizes
// only; device pointers into the workspace are then offsets from 0).
static bool plan_layout(const dh_batch_plan *plans, int n_plans, const dh_batch_beam *beams, int n_beams,
                        const AdjTune &tune, float cutoff_sigma, char *base, Layout &L) {
    if (n_plans <= 0 || !plans || n_beams < 0 || (n_beams > 0 && !beams)) {
        set_error("dh_batch: need at least one plan (n_plans=%d, n_beams=%d)", n_plans, n_beams);
        return false;
    }
    if (!tune_ok(tune)) {
        set_error("dh_batch: invalid schedule (adj_cs=%d, wg_target=%d, min_bricks=%d, max_slabs=%d)", tune.cs,
                  tune.wg_target, tune.min_bricks, tune.max_slabs);
        return false;
    }
    int flags = F_RLSP | F_GRAD_DOSE | F_GRAD_OUT;
    int next = 0;
    for (int p = 0; p < n_plans; ++p) {
        const dh_batch_plan &P = plans[p];
        if (P.beam_start != next || P.n_beams < 0 || P.beam_start + (long long)P.n_beams > n_beams) {
            set_error("dh_batch: plan %d lists beams [%d, %d); plans must list beams 0..%d in order without gaps",
                      p, P.beam_start, P.beam_start + P.n_beams, n_beams);
            return false;
        }
        next += P.n_beams;
        if (!pencil_grid_ok(P.grid) || !grid_ok(P.grid)) return false;
        if (P.grid.spacing < 0.2f) {
            set_error("dh_batch: plan %d spacing %g mm below 0.2 mm", p, (double)P.grid.spacing);
            return false;
        }
        if (!P.dose) {
            set_error("dh_batch: plan %d has no dose volume", p);
            return false;
        }
        if (!P.rlsp_padded) flags &= ~F_RLSP;
        else if (((uintptr_t)P.rlsp_padded & 15u) != 0) {
            set_error("dh_batch: plan %d quad RLSP volume not 16-byte aligned", p);
            return false;
        }
        if (!P.grad_dose) flags &= ~F_GRAD_DOSE;
    }
    if (next != n_beams) {
        set_error("dh_batch: plans list %d beams, %d given", next, n_beams);
        return false;
    }

    // device address of workspace byte `o` (an offset when sizing only)
    auto at = [base](long long o) { return reinterpret_cast<char *>(reinterpret_cast<uintptr_t>(base) + o); };
    L.ctx.assign(n_beams, BeamCtx{});
    L.plans.resize(n_plans);
    L.stencil.resize(n_beams);
    // tables first, then per-beam pencil-beam workspaces, then scratch volumes
    long long off = 0;
    L.tab[T_CTX] = off;
    off += align_up((long long)n_beams * sizeof(BeamCtx));
    L.tab[T_PLANS] = off;
    off += align_up((long long)n_plans * sizeof(PlanCtx));
    for (int t = T_BRICK_OFF; t <= T_REDUCE_OFF; ++t) {
        L.tab[t] = off;
        off += align_up((long long)(n_beams > 0 ? n_beams : 1) * 8);
    }
    L.tab[T_SUM_OFF] = off;
    off += align_up((long long)n_plans * 8);
    L.tab[T_STENCIL] = off;
    off += align_up((long long)n_beams * sizeof(Stencil));
    L.tab[T_TOUCH] = off;
    off += align_up((long long)n_beams * MAX_TOUCH * 4);
    L.tab[T_ROW_OFF] = off;
    off += align_up((long long)(n_beams > 0 ? n_beams : 1) * 8);
    L.tab[T_HEADER_END] = off;
    L.touch.assign((size_t)n_beams * MAX_TOUCH, 0);

    for (int p = 0; p < n_plans; ++p) {
        const dh_batch_plan &P = plans[p];
        const long long nvox = nvox_of(P.grid);
        PlanCtx &pc = L.plans[p];
        pc.dose = P.dose;
        pc.nvox = nvox;
        pc.beam0 = P.beam_start;
        pc.n_beams = P.n_beams;
        bool direct = P.n_beams == 1;  // the only beam writes the plan dose itself
        for (int b = P.beam_start; b < P.beam_start + P.n_beams; ++b) {
            const dh_batch_beam &B = beams[b];
            Model m;
            if (!fill_model(&B.model, m)) return false;
            if (!spot_set_ok(&B.spots)) return false;
            if (!B.wet) {
                set_error("dh_batch: beam %d has no WET volume", b);
                return false;
            }
            const int S = B.spots.n_spots, NL = S > 0 ? B.spots.n_layers : 0;
            if (S > 0 && (!B.grad_mu || !B.grad_xy)) flags &= ~F_GRAD_OUT;
            const long long nb = bricks_of(P.grid);
            BeamWs w = {};
            w.nslab = 1;
            w.cs = 64;
            if (S > 0) {
                w = carve_beam(at(off), nb, S, NL, tune, cutoff_sigma > 0.0f);
                off += w.bytes;
            }
            BeamCtx &c = L.ctx[b];
            fill_pencil_ctx(c, P.grid, B.geom, m, B.spots, w, cutoff_sigma);
            c.rlsp_q = reinterpret_cast<const float4 *>(P.rlsp_padded);
            c.wet = B.wet;
            c.gdose = P.grad_dose;
            c.grad_mu = B.grad_mu;
            c.grad_xy = B.grad_xy;
            make_stencil(B.geom, (double)P.grid.spacing, L.stencil[b]);
            c.stencil = reinterpret_cast<const Stencil *>(at(L.tab[T_STENCIL] + (long long)b * sizeof(Stencil)));
            c.n_touch = make_touch(L.stencil[b], L.touch.data() + (size_t)b * MAX_TOUCH, MAX_TOUCH);
            L.box_max = std::max(L.box_max, smooth_box(L.stencil[b], c.slo, c.sbox));
            c.touch = reinterpret_cast<const int *>(at(L.tab[T_TOUCH] + (long long)b * MAX_TOUCH * 4));
            c.dose = B.dose;
            if (c.dose && c.dose != P.dose) direct = false;
            if (!c.dose && P.n_beams == 1) c.dose = P.dose;
            if (!c.dose) {
                c.dose = reinterpret_cast<float *>(at(off));
                off += align_up(nvox * 4);
            }
            c.wet_raw = B.wet_raw;
            if (!c.wet_raw) {
                c.wet_raw = reinterpret_cast<float *>(at(off));
                off += align_up(nvox * 4);
            }
            c.need = reinterpret_cast<unsigned char *>(at(off));
            off += align_up(nb);
            L.brick.push_back(L.n_bricks);
            L.n_bricks += nb;
            L.smooth.push_back(L.n_smooth);
            L.n_smooth += (nvox + 255) / 256;
            L.smooth_tiles.push_back(L.n_tiles);
            L.n_tiles += (long long)((P.grid.nx + STX - 1) / STX) * ((P.grid.ny + STY - 1) / STY) *
                         ((P.grid.nz + STZ - 1) / STZ);
            L.spot.push_back(L.n_spots);
            L.n_spots += S;
            L.layer.push_back(L.n_layers);
            L.n_layers += NL;
            L.n_part += S > 0 ? (long long)w.nslab * S : 0;
            const long long items = S > 0 ? adjoint_chunks(S, NL, w.cs) * w.nslab : 0;
            L.adj32.push_back(L.n_adj32);
            L.adj64.push_back(L.n_adj64);
            L.adj128.push_back(L.n_adj128);
            if (w.cs == 32) L.n_adj32 += items;
            else if (w.cs == 64) L.n_adj64 += items;
            else L.n_adj128 += items;
            L.reduce.push_back(L.n_reduce);
            L.n_reduce += (S + 63) / 64;
            // brick summary: one workgroup per row of bricks along x; a beam
            // without spots has no adjoint work and gets none
            L.rows.push_back(L.n_rows);
            if (S > 0) L.n_rows += (long long)c.bg.by * c.bg.bz;
        }
        // plans whose dose is not written by their beam directly (several
        // beams, none, or a beam with a volume of its own) get the sum pass
        L.sum.push_back(L.n_sum);
        if (!direct) L.n_sum += (nvox + PLAN_SUM_VOXELS - 1) / PLAN_SUM_VOXELS;
    }
    // smoothing items: tiles when every box fits the LDS budget, else blocks
    if (L.box_max <= SMOOTH_BOX_MAX) {
        L.smooth = L.smooth_tiles;
        L.n_smooth = L.n_tiles;
        flags |= F_SMOOTH_TILE;
    }
    if (L.n_rows >= (1LL << 31)) {
        set_error("dh_batch: %lld brick rows, at most 2^31 - 1", L.n_rows);
        return false;
    }
    L.bytes = off;
    L.flags = flags;
    return true;
}

static bool batch_ok(const dh_batch *b) {
    if (!b || b->magic != kMagic || !b->ws) {
        set_error("dh_batch: not set up (call dh_batch_setup first)");
        return false;
    }
    return true;
}

template <class T>
static const T *table(const dh_batch *b, int t) {
    return reinterpret_cast<const T *>(reinterpret_cast<const char *>(b->ws) + b->tab[t]);
}

}  // namespace dh

using namespace dh;

extern "C" int64_t dh_batch_workspace_size(const dh_batch_plan *plans, int32_t n_plans, const dh_batch_beam *beams,
                                           int32_t n_beams, const dh_batch_tune *tune) {
    // the adjoint schedule (and so the partial records) depends on whether the
    // cutoff culls, which this query does not know: the larger layout
    Layout L, Lc;
    if (!plan_layout(plans, n_plans, beams, n_beams, tune_from(tune), 0.0f, nullptr, L)) return DH_ERR_ARG;
    if (!plan_layout(plans, n_plans, beams, n_beams, tune_from(tune), 1.0f, nullptr, Lc)) return DH_ERR_ARG;
    return L.bytes > Lc.bytes ? L.bytes : Lc.bytes;
}

extern "C" int dh_batch_setup(dh_batch *batch, const dh_batch_plan *plans, int32_t n_plans,
                              const dh_batch_beam *beams, int32_t n_beams, const dh_batch_tune *tune,
                              float cutoff_sigma, void *workspace, int64_t workspace_bytes, void *stream) {
    if (!batch) {
        set_error("dh_batch_setup: null batch");
        return DH_ERR_ARG;
    }
    memset(batch, 0, sizeof *batch);
    if (tune && (tune->fwd_blocks < 0 || tune->ray_blocks < 0 || tune->fwd_blocks > (1 << 20) ||
                 tune->ray_blocks > (1 << 20) || tune->ray_shape < 0 || tune->ray_shape > 5)) {
        set_error("dh_batch_setup: invalid brick schedule (fwd_blocks=%d, ray_blocks=%d, ray_shape=%d)",
                  tune->fwd_blocks, tune->ray_blocks, tune->ray_shape);
        return DH_ERR_ARG;
    }
    const AdjTune t = tune_from(tune);
    Layout L;
    if (!plan_layout(plans, n_plans, beams, n_beams, t, cutoff_sigma, nullptr, L)) return DH_ERR_ARG;
    if (!workspace || workspace_bytes < L.bytes || ((uintptr_t)workspace % kAlign) != 0) {
        set_error("dh_batch_setup: workspace too small or misaligned (%lld bytes given, %lld needed, 256-byte "
                  "alignment)",
                  (long long)workspace_bytes, L.bytes);
        return DH_ERR_ARG;
    }
    Layout D;  // the same layout with real device addresses
    plan_layout(plans, n_plans, beams, n_beams, t, cutoff_sigma, reinterpret_cast<char *>(workspace), D);

    // host image of the table region, uploaded in one copy
    std::vector<char> img((size_t)D.tab[T_HEADER_END], 0);
    auto put = [&](int tab, const void *src, size_t bytes) {
        if (bytes) memcpy(img.data() + D.tab[tab], src, bytes);
    };
    put(T_CTX, D.ctx.data(), D.ctx.size() * sizeof(BeamCtx));
    put(T_PLANS, D.plans.data(), D.plans.size() * sizeof(PlanCtx));
    put(T_BRICK_OFF, D.brick.data(), D.brick.size() * 8);
    put(T_SMOOTH_OFF, D.smooth.data(), D.smooth.size() * 8);
    put(T_SPOT_OFF, D.spot.data(), D.spot.size() * 8);
    put(T_LAYER_OFF, D.layer.data(), D.layer.size() * 8);
    put(T_ADJ128_OFF, D.adj128.data(), D.adj128.size() * 8);
    put(T_ADJ32_OFF, D.adj32.data(), D.adj32.size() * 8);
    put(T_ADJ64_OFF, D.adj64.data(), D.adj64.size() * 8);
    put(T_REDUCE_OFF, D.reduce.data(), D.reduce.size() * 8);
    put(T_SUM_OFF, D.sum.data(), D.sum.size() * 8);
    put(T_STENCIL, D.stencil.data(), D.stencil.size() * sizeof(Stencil));
    put(T_TOUCH, D.touch.data(), D.touch.size() * 4);
    put(T_ROW_OFF, D.rows.data(), D.rows.size() * 8);
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(workspace, img.data(), img.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        set_error("dh_batch_setup: table upload: %s", hipGetErrorString(e));
        return DH_ERR_HIP;
    }

    batch->ws = workspace;
    batch->ws_bytes = workspace_bytes;
    batch->n_plans = n_plans;
    batch->n_beams = n_beams;
    batch->cutoff_sigma = cutoff_sigma > 0.0f ? cutoff_sigma : 0.0f;
    batch->n_bricks = D.n_bricks;
    batch->n_smooth_blocks = D.n_smooth;
    batch->n_spots = D.n_spots;
    batch->n_layers = D.n_layers;
    batch->n_sum_blocks = D.n_sum;
    batch->n_partials = D.n_part;
    batch->n_adj32 = D.n_adj32;
    batch->n_adj64 = D.n_adj64;
    batch->n_adj128 = D.n_adj128;
    batch->n_reduce_blocks = D.n_reduce;
    for (int i = 0; i < N_TABLES; ++i) batch->tab[i] = D.tab[i];
    batch->tab[T_FLAGS] = D.flags | (D.flags & F_SMOOTH_TILE ? (long long)D.box_max << 32 : 0LL);
    batch->n_summary_rows = (int32_t)D.n_rows;
    batch->fwd_blocks = tune && tune->fwd_blocks > 0 ? tune->fwd_blocks : 0;
    batch->ray_blocks = tune && tune->ray_blocks > 0 ? tune->ray_blocks : 0;
    batch->ray_shape = tune ? tune->ray_shape : 0;
    batch->magic = kMagic;
    return DH_OK;
}

// Spot and layer records: prepared now, or (flags & DH_BATCH_REUSE_PREP) those
// an earlier phase call on this batch prepared.  Reusing records no call has
// prepared would hand the kernels uninitialised workspace (spot ranges used as
// indices), so it is an error.
static int prepare_records(dh_batch *b, int32_t flags, hipStream_t st, const char *who) {
    if (flags & DH_BATCH_REUSE_PREP) {
        if (!b->prepared) {
            set_error("%s: DH_BATCH_REUSE_PREP before any call prepared the spot records of this batch", who);
            return DH_ERR_ARG;
        }
        return DH_OK;
    }
    if (b->n_beams == 0) return DH_OK;
    const int rc = launch_prep_batch(table<BeamCtx>(b, T_CTX), table<long long>(b, T_SPOT_OFF),
                                     table<long long>(b, T_LAYER_OFF), b->n_beams, b->n_spots, b->n_layers, st);
    if (!rc) b->prepared = 1;
    return rc;
}

extern "C" int dh_batch_wet(dh_batch *b, int32_t flags, void *stream) {
    if (!batch_ok(b)) return DH_ERR_ARG;
    if (!(b->tab[T_FLAGS] & F_RLSP)) {
        set_error("dh_batch_wet: a plan of the batch has no quad RLSP volume");
        return DH_ERR_ARG;
    }
    if (b->ray_shape < 0 || b->ray_shape > 5) {
        set_error("dh_batch_wet: ray_shape %d is not in 0 .. 5", b->ray_shape);
        return DH_ERR_ARG;
    }
    if (b->n_beams == 0) return DH_OK;
    hipStream_t st = (hipStream_t)stream;
    const bool reach = (flags & DH_BATCH_WET_REACH) != 0;
    if (reach) {
        int rc = prepare_records(b, flags, st, "dh_batch_wet");
        if (!rc) rc = launch_wet_need_batch(table<BeamCtx>(b, T_CTX), table<long long>(b, T_BRICK_OFF), b->n_beams,
                                            b->n_bricks, st, b->events);
        if (rc) return rc;
    }
    const int box = (b->tab[T_FLAGS] & F_SMOOTH_TILE) ? (int)(b->tab[T_FLAGS] >> 32) : 0;
    return launch_wet_batch(table<BeamCtx>(b, T_CTX), table<long long>(b, T_BRICK_OFF),
                            table<long long>(b, T_SMOOTH_OFF), b->n_beams, b->n_bricks, b->n_smooth_blocks, reach,
                            b->ray_blocks, b->ray_shape, box, st, b->events);
}

extern "C" int dh_batch_forward(dh_batch *b, int32_t flags, void *stream) {
    if (!batch_ok(b)) return DH_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const int rc = prepare_records(b, flags, st, "dh_batch_forward");
    if (rc) return rc;
    return launch_forward_batch(table<BeamCtx>(b, T_CTX), table<long long>(b, T_BRICK_OFF), b->n_beams,
                                b->n_bricks, table<PlanCtx>(b, T_PLANS), table<long long>(b, T_SUM_OFF), b->n_plans,
                                b->n_sum_blocks, b->cutoff_sigma > 0.0f, b->fwd_blocks, st, b->events);
}

extern "C" int dh_batch_adjoint(dh_batch *b, int32_t flags, void *stream) {
    if (!batch_ok(b)) return DH_ERR_ARG;
    if (b->n_spots == 0) return DH_OK;
    if (!(b->tab[T_FLAGS] & F_GRAD_DOSE) || !(b->tab[T_FLAGS] & F_GRAD_OUT)) {
        set_error("dh_batch_adjoint: a plan has no grad_dose or a beam has no gradient outputs");
        return DH_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    const int rc = prepare_records(b, flags, st, "dh_batch_adjoint");
    if (rc) return rc;
    return launch_adjoint_batch(table<BeamCtx>(b, T_CTX), b->n_beams, table<long long>(b, T_ROW_OFF),
                                b->n_summary_rows, table<long long>(b, T_ADJ32_OFF), b->n_adj32,
                                table<long long>(b, T_ADJ64_OFF), b->n_adj64, table<long long>(b, T_ADJ128_OFF),
                                b->n_adj128, table<long long>(b, T_REDUCE_OFF),
                                b->n_reduce_blocks, b->cutoff_sigma > 0.0f, st, b->events);
}

extern "C" int dh_batch_timing(dh_batch *b, void *const *events) {
    if (!batch_ok(b)) return DH_ERR_ARG;
    b->events = events;
    return DH_OK;
}
