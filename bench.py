#!/usr/bin/env python3
"""Forward + adjoint of the 10k-spot head phantom on one MI355X: batched vs plain.

Workload (synthetic, seed 2): 160 x 200 x 200 (z, y, x) HU grid at 2 mm --
ellipsoidal soft tissue N(40, 20) clipped to [-100, 100], 8 mm bone shell
U(700, 1200), two 15 mm air spheres, air outside; 3 beams at gantry 0/120/240;
40 energy layers (RS0 ids 20..59) x 81 spots (3 mm grid in a 15 mm disk) per
beam = 9720 spots.  Plan copy 0 has the seed-2 weights U(0, 1); copies 1..K-1
draw their own weights U(0, 1).

One step (batched path) = ONE HIP-graph replay over a batch of K such plans:
reach-limited WET (ray-march + smoothing where the spots can reach), plan dose
(forward), dL/dD of L = mean_target (D - 2)^2 + 0.1 mean_body D^2 (target:
25 mm sphere at the isocentre, body: HU > -500), adjoint (dL/dweight,
dL/d(x, y)) for every beam of every plan.  The timed graph runs the plans as
--streams (default 2) sub-batches on concurrent streams (StreamedBatches:
results bit-identical to one batch); the one-batch, one-stream graph is
timed beside it, and the per-kernel times come from eager steps of the one
batch (each kernel with the GPU to itself).

Plain path (per plan, for the ratio): per beam eager calls of full WET
(ray-march + smoothing), dense forward (every spot for every in-range voxel,
cutoff 0) summed over beams, the same dL/dD, dense adjoint per beam.  Two
more lines separate the gains: the same eager per-beam path with the 6-sigma
cutoff (culling alone), and one plan on the batched graph path (K = 1:
batching of its beams, graph replay and reach-limited WET, without other
plans to share the launches).

value = spot.voxel evaluations per second counted as the reference does: per
plan, every spot of a layer for every voxel with WET <= 1.1 R80 (forward) plus
the same over voxels with dL/dD != 0 of the plain path (adjoint), times K, per
second of the timed batched steps.

Prints ONE JSON line.
"""

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "dosecuda-jax_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "fwd+adjoint ms & spot·voxel evals/s, 10k-spot head phantom; batched vs plain"
FWD_FLOP_PER_PAIR = 20.0   # r_x, r_y 2; |r|^2 4; t.r 4; d2 3; 2 scales + factor 4; sum 1; accumulate 2
ADJ_FLOP_PER_PAIR = 41.0
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def masks(g, dev):
    import numpy as np
    import torch
    from DoseCUDA.workloads import voxel_centres
    z, y, x = voxel_centres(g)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    target = (X ** 2 + Y ** 2 + Z ** 2) <= 25.0 ** 2
    body = g.HU > -500.0
    return torch.from_numpy(target).to(dev), torch.from_numpy(body).to(dev)


def loss_grad(dose, tmask, bmask, n_t, n_b, out=None):
    """dL/dD of mean_target (D - 2)^2 + 0.1 mean_body D^2 (element-wise ops only)."""
    import torch
    g = torch.where(tmask, (dose - 2.0) * (2.0 / n_t), torch.zeros_like(dose))
    g = g + torch.where(bmask, dose * (0.2 / n_b), torch.zeros_like(dose))
    if out is not None:
        out.copy_(g)
        return out
    return g


class LossGrad:
    """The same dL/dD as loss_grad as one fused element-wise kernel:
    dL/dD = a * D + b with a = 2/n_t [target] + 0.2/n_b [body], b = -4/n_t [target]."""

    def __init__(self, tmask, bmask, n_t, n_b):
        self.a = tmask.float() * (2.0 / n_t) + bmask.float() * (0.2 / n_b)
        self.b = tmask.float() * (-4.0 / n_t)

    def __call__(self, dose, out=None):
        import torch
        if out is None:
            out = torch.empty_like(dose)
        return torch.addcmul(self.b, self.a, dose, out=out)


def plan_copies(K):
    """K head-phantom plans: copy 0 with the seed-2 weights, the others with weights of their own."""
    import numpy as np
    from DoseCUDA.workloads import head_plan
    g, plan = head_plan(seed=2)
    plans = [plan]
    for k in range(1, K):
        _, p = head_plan(seed=2)
        rng = np.random.default_rng(100 + k)
        for b in p.beam_list:
            b.spot_list[:, 2] = rng.uniform(0.0, 1.0, size=b.spot_list.shape[0]).astype(np.float32)
        plans.append(p)
    return g, plans


def cpu_baseline(g, plan, eng, gvol_np, n_sample, seed=0):
    """Oracle float32 build ('port') on a voxel sample of the same workload:
    ray-march of the sampled voxels, then forward dose and adjoint at them."""
    import numpy as np
    from oracle import impt_oracle as O
    O.lib("f32")
    rng = np.random.default_rng(seed)
    nvox = int(np.prod(g.HU.shape))
    idx = np.sort(rng.choice(nvox, size=min(n_sample, nvox), replace=False)).astype(np.int64)
    machine_dir = os.path.join(PKG, "DoseCUDA", "lookuptables", "protons", plan.machine_name)
    rlsp = O.rlsp_from_hu(g.HU, machine_dir)
    s = eng.spacing
    t_total = 0.0
    pairs = 0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    for k, b in enumerate(eng.beams):
        model = b.model
        geom = O.beam_geometry(np.asarray(b.beam.iso, np.float32), eng.origin, b.beam.gantry_angle,
                               b.beam.couch_angle, model.VSADX, model.VSADY)
        _, spots, layers, r80 = O.spots_and_layers(b.beam.spot_list, model.divergence_params)
        marr = O.ModelArrays(model)
        wet_full = b.wet.detach().cpu().numpy().astype(np.float64)     # same WET the GPU used
        wet_s = np.zeros_like(wet_full)
        wet_s.reshape(-1)[idx] = wet_full.reshape(-1)[idx]
        gs = np.zeros_like(wet_full)
        gs.reshape(-1)[idx] = gvol_np.reshape(-1)[idx]
        raw_full = b.wet_raw.detach().cpu().numpy() if b.wet_raw is not None else wet_full
        t0 = time.perf_counter()
        O.raymarch(rlsp, s, geom, idx=idx, precision="f32")
        O.smooth(raw_full, s, geom, idx=idx, precision="f32")
        O.pb_dose(wet_s, s, geom, spots, layers, r80, marr, idx=idx, precision="f32")
        t1 = time.perf_counter()
        O.pb_grad(wet_s, gs, s, geom, spots, layers, r80, marr, precision="f32")
        t2 = time.perf_counter()
        t_total += t2 - t0
        w_mm = wet_full.reshape(-1)[idx] * 10.0
        gnz = gvol_np.reshape(-1)[idx] != 0
        for (st, ns, eid), R in zip(layers, r80):
            inr = w_mm <= 1.1 * R
            pairs += int(inr.sum()) * int(ns) + int((inr & gnz).sum()) * int(ns)
        log("cpu baseline beam %d: fwd %.2fs adj %.2fs" % (k, t1 - t0, t2 - t1))
    model_name = cpu_model()
    return {"value": pairs / t_total, "unit": "spot·voxel evals/s", "cores": threads, "kind": "port",
            "cpu_model": model_name,
            "sample": "%d random voxels of the head phantom (plan copy 0), 3 beams: ray-march of the sampled "
                      "voxels, smoothing of them (54 neighbour reads each), forward dose and adjoint at them "
                      "(float32 oracle build, OpenMP on %d threads of %s, dense: every spot); %d spot.voxel pairs "
                      "in %.1f s" % (len(idx), threads, model_name, pairs, t_total)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def load_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k.get("hbm_bytes_per_launch")


def _graph_ms(step, dev, warmup, steps):
    import torch
    from DoseCUDA.batch import GraphedStep
    gs = GraphedStep(step, dev, warmup=max(1, warmup))
    for _ in range(warmup):
        gs.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        gs.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3, gs


def _target_masks(g, dev, radius=25.0):
    import numpy as np
    import torch
    from DoseCUDA.workloads import voxel_centres
    z, y, x = voxel_centres(g)
    r2 = z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2
    t = torch.from_numpy(np.ascontiguousarray(r2 <= radius * radius)).to(dev)
    return t


def run_uneven(args):
    """64 plans of uneven size (cube side U{64..256} @ 2 mm, water + 20 mm HU = 1000 slab at 60 mm, 1-4 beams
    at random gantry, 500-40k spots log-uniform, seed 4): one batched graph step vs per-plan paths."""
    import numpy as np
    import torch
    from DoseCUDA import device as D
    from DoseCUDA.batch import BatchedDose, StreamedBatches, tune_dict
    from DoseCUDA.engine import IMPTDoseEngine
    from DoseCUDA.workloads import uneven_plans
    n_plans = args.uneven_plans
    log("building %d uneven plans" % n_plans)
    plans = uneven_plans(n_plans, seed=4)
    engines = [IMPTDoseEngine(g, p, gpu_id=0, cutoff_sigma=args.cutoff_sigma) for g, p in plans]
    dev = engines[0].device
    masks = [_target_masks(g, dev) for g, _ in plans]
    counts = [(int(m.sum().item()), int(np.prod(e.shape))) for m, e in zip(masks, engines)]

    def grad_into(d, out, m, nt, nb):
        g = torch.where(m, (d - 2.0) * (2.0 / max(nt, 1)), torch.zeros_like(d))
        out.copy_(g + d * (0.2 / nb))

    def make_step(bd, idx):
        def step():
            doses = bd.forward()
            for d, gd, k in zip(doses, bd.grad_doses, idx):
                grad_into(d, gd, masks[k], counts[k][0], counts[k][1])
            return doses, bd.adjoint(None, reuse_prep=True, input_order=False)
        return step

    def tune_grad(doses):
        out = [torch.empty_like(d) for d in doses]
        for k, (d, g) in enumerate(zip(doses, out)):
            grad_into(d, g, masks[k], counts[k][0], counts[k][1])
        return out

    t0 = time.perf_counter()
    bd = BatchedDose(engines, wet_mode="reach", tune=None if args.no_tune else "auto", tune_grad_fn=tune_grad,
                     tune_grad_key="uneven-loss")
    t_setup = time.perf_counter() - t0
    log("batch items %s, schedule %s, set-up %.1f s" % (bd.item_counts(), tune_dict(bd.tune), t_setup))
    ms_batch, gs = _graph_ms(make_step(bd, list(range(n_plans))), dev, args.warmup, args.steps)
    del gs, bd
    torch.cuda.empty_cache()
    log("batched step %.2f ms for %d plans" % (ms_batch, n_plans))
    # the same step as sub-batches on concurrent streams
    ms_streamed = None
    if args.streams > 1:
        sb = StreamedBatches(engines, n_streams=args.streams,
                             grad_fn=lambda p, d, out: grad_into(d, out, masks[p], counts[p][0], counts[p][1]),
                             wet_mode="reach", tune=None if args.no_tune else "auto", tune_grad_key="uneven-loss")
        log("%d sub-batches of %s plans" % (len(sb.subs), [len(x) for x in sb.groups]))
        ms_streamed, gs = _graph_ms(lambda: sb.step(input_order=False), dev, args.warmup, args.steps)
        del gs, sb
        torch.cuda.empty_cache()
        log("streamed step %.2f ms for %d plans" % (ms_streamed, n_plans))
    # each plan alone, same batched path (its own graph, default schedule)
    single = []
    for k, e in enumerate(engines):
        b1 = BatchedDose([e], wet_mode="reach")
        ms, gs = _graph_ms(make_step(b1, [k]), dev, 1, 3)
        single.append(ms)
        del gs, b1
    # plain path per plan: eager per-beam calls, full WET, dense forward and adjoint
    plain = []
    for k, e in enumerate(engines):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for b in e.beams:
            b.wet_raw, b.wet = D.wet_from_padded(e.rlsp_padded, e.grid, b.geo, dev, raw_out=b.wet_raw, out=b.wet)
        dose = torch.empty(e.shape, dtype=torch.float32, device=dev)
        if not e.beams:
            dose.zero_()
        for j, b in enumerate(e.beams):
            D.pb_forward(b.wet, dose, e.grid, b.geo, b.dmodel, b.spots, b.ws, b.ws_bytes, cutoff_sigma=0.0,
                         accumulate=j > 0)
        gd = torch.empty_like(dose)
        grad_into(dose, gd, masks[k], counts[k][0], counts[k][1])
        for b in e.beams:
            D.pb_adjoint(b.wet, gd, e.grid, b.geo, b.dmodel, b.spots, b.ws, b.ws_bytes, cutoff_sigma=0.0)
        torch.cuda.synchronize()
        plain.append((time.perf_counter() - t0) * 1e3)
        for b in e.beams:
            b._ws = None   # free the single-beam workspace again
        torch.cuda.empty_cache()
    single, plain = np.array(single), np.array(plain)
    sides = [int(g.HU.shape[0]) for g, _ in plans]
    spots = [sum(b.spot_list.shape[0] for b in p.beam_list) for _, p in plans]
    rec = {
        "metric": "batched fwd+adjoint of uneven plans: ms per plan (batch) and per-plan latency",
        "workload": "%d plans, cube side U{64..256} @ 2 mm (sides %d..%d), water + 20 mm HU=1000 slab at 60 mm, "
                    "1-4 beams at random gantry, %d..%d spots per plan (total %d), seed 4; step = reach-limited "
                    "WET, forward, dL/dD (25 mm target, 2 Gy), adjoint" % (n_plans, min(sides), max(sides),
                                                                          min(spots), max(spots), sum(spots)),
        "batch_ms_per_step": ms_batch,
        "batch_ms_per_plan": ms_batch / n_plans,
        "streams": args.streams,
        "streamed_batch_ms_per_step": ms_streamed,
        "alone_batched_ms": {"p50": float(np.percentile(single, 50)), "p99": float(np.percentile(single, 99)),
                             "sum": float(single.sum())},
        "plain_ms": {"p50": float(np.percentile(plain, 50)), "p99": float(np.percentile(plain, 99)),
                     "sum": float(plain.sum())},
        "speedup_batch_vs_plain_per_plan": float(plain.sum() / ms_batch),
        "speedup_batch_vs_alone": float(single.sum() / ms_batch),
        "speedup_streamed_vs_plain_per_plan": float(plain.sum() / ms_streamed) if ms_streamed else None,
        "setup_and_autotune_s": t_setup,
    }
    print(json.dumps(rec), flush=True)


def run_body(args):
    """Large body-size optimisation loop: 300x512x512 @ 1 mm, 4 beams, ~63k spots; WET once, then
    `--iters` projected-gradient weight steps w <- max(0, w - eta dL/dw), each one batched graph replay
    (forward, dL/dD, adjoint, update); final dose vs the float64 oracle on a voxel sample."""
    import numpy as np
    import torch
    from DoseCUDA.batch import BatchedDose, GraphedStep, tune_dict
    from DoseCUDA.engine import IMPTDoseEngine
    from DoseCUDA.workloads import body_plan
    log("building the body phantom")
    g, plan, target = body_plan()
    t_wall0 = time.perf_counter()
    eng = IMPTDoseEngine(g, plan, gpu_id=0, cutoff_sigma=args.cutoff_sigma)
    dev = eng.device
    tmask = torch.from_numpy(target).to(dev)
    body = torch.from_numpy(g.HU > -500.0).to(dev)
    n_t, n_b = float(tmask.sum().item()), float(body.sum().item())
    bd = BatchedDose([eng], wet_mode="full", tune=None if args.no_tune else "auto")
    bd.compute_wet()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_wall0
    log("set-up (HU->RLSP, WET, batch, autotune) %.1f s, schedule %s" % (t_setup, tune_dict(bd.tune)))
    mus = [b.mu0.clone() for b in eng.beams]
    lr = torch.zeros(1, dtype=torch.float32, device=dev)
    loss_t = torch.zeros(1, dtype=torch.float32, device=dev)

    def step():
        doses = bd.forward(mus, None)
        d = doses[0]
        diff = torch.where(tmask, d - 2.0, torch.zeros_like(d))
        gd = bd.grad_doses[0]
        gd.copy_(diff * (2.0 / n_t) + torch.where(body, d * (0.2 / n_b), torch.zeros_like(d)))
        grads = bd.adjoint(None, reuse_prep=True)
        for m, (gm, _) in zip(mus, grads):
            m.sub_(lr * gm).clamp_(min=0.0)
        return doses

    # step size from the first gradient: eta = 0.05 max(w) / max |dL/dw|
    d0 = bd.forward(mus, None)[0]
    loss0 = float(((d0[tmask] - 2.0) ** 2).mean() + 0.1 * (d0[body] ** 2).mean())
    gd = bd.grad_doses[0]
    gd.copy_(torch.where(tmask, d0 - 2.0, torch.zeros_like(d0)) * (2.0 / n_t)
             + torch.where(body, d0 * (0.2 / n_b), torch.zeros_like(d0)))
    g0 = bd.adjoint(None)
    gmax = max(float(gm.abs().max()) for gm, _ in g0)
    lr.fill_(0.05 * max(float(m.max()) for m in mus) / gmax)
    gs = GraphedStep(lambda: step(), dev, warmup=1)   # warm-up and capture run 2 iterations
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in range(args.iters):
        gs.replay()
        if (it + 1) % 10 == 0:
            torch.cuda.synchronize()
            log("iteration %d" % (it + 1))
    torch.cuda.synchronize()
    t_loop = time.perf_counter() - t0
    d = bd.forward(mus, None)[0]
    loss1 = float(((d[tmask] - 2.0) ** 2).mean() + 0.1 * (d[body] ** 2).mean())
    eng.compute_wet(force=True)   # single-beam path: raw and smoothed WET of every voxel (same kernels)
    # accuracy: float64 oracle at sampled voxels (half in the 10%-of-max region) on the engine's WET
    from oracle import impt_oracle as O
    dn = d.detach().cpu().numpy()
    rng = np.random.default_rng(0)
    hi = np.flatnonzero(dn.reshape(-1) > 0.1 * dn.max())
    idx = np.unique(np.concatenate([rng.choice(hi, size=min(2000, hi.size), replace=False),
                                    rng.choice(dn.size, size=2000, replace=False)])).astype(np.int64)
    s = eng.spacing
    ref = np.zeros(idx.size)
    wet_err = 0.0
    rlsp = O.rlsp_from_hu(g.HU, os.path.join(PKG, "DoseCUDA", "lookuptables", "protons", plan.machine_name))
    for b, m in zip(eng.beams, mus):
        model = b.model
        geom = O.beam_geometry(np.asarray(b.beam.iso, np.float32), eng.origin, b.beam.gantry_angle,
                               b.beam.couch_angle, model.VSADX, model.VSADY)
        sl = np.array(b.beam.spot_list, dtype=np.float32)
        sl[:, 2] = m.detach().cpu().numpy()
        _, spots, layers, r80 = O.spots_and_layers(sl, model.divergence_params)
        wet_gpu = b.wet.detach().cpu().numpy().astype(np.float64)
        ref += O.pb_dose(wet_gpu, s, geom, spots, layers, r80, O.ModelArrays(model), idx=idx)
        raw_o = O.raymarch(rlsp, s, geom, idx=idx)
        raw_g = b.wet_raw.detach().cpu().numpy().reshape(-1)[idx] if b.wet_raw is not None else None
        if raw_g is not None:
            wet_err = max(wet_err, float(np.abs(raw_g - raw_o).max()))
    got = dn.reshape(-1)[idx]
    rec = {
        "metric": "large body optimisation loop: wall time and final dose error vs float64 oracle",
        "workload": "300x512x512 @ 1 mm body phantom (seed 5), 4 beams (gantry 0/90/180/270), %d spots, "
                    "%d projected-gradient weight iterations, each one HIP-graph replay of forward + dL/dD + "
                    "adjoint + update; WET computed once" % (sum(b.spots.n_spots for b in eng.beams), args.iters),
        "setup_s": t_setup,
        "loop_s": t_loop,
        "ms_per_iteration": t_loop / args.iters * 1e3,
        "loss_first": loss0,
        "loss_last": loss1,
        "final_dose_vs_oracle_max_rel_to_dmax": float(np.abs(got - ref).max() / max(dn.max(), 1e-30)),
        "raw_wet_vs_oracle_max_cm": wet_err,
        "oracle_sample_voxels": int(idx.size),
        "schedule": tune_dict(bd.tune),
    }
    print(json.dumps(rec), flush=True)


def run_small(args):
    """The two small configurations on the 100^3 @ 2.5 mm water + slab grid: the 320-spot plan (1 beam,
    5 layers; forward dose: WET + spot sum) and the 10k-spot target plan (forward + adjoint of
    mean_target (D - 2)^2 + 0.1 mean_shell D^2). Per plan: the drop-in computeIMPTPlan (host arrays, every
    spot), the plain device path (per-beam calls; dense, then with the 6-sigma cutoff) and the batched
    HIP-graph step (reach-limited WET, cutoff 6); the float32 oracle on every voxel (OpenMP) beside them."""
    import numpy as np
    import torch
    from DoseCUDA.batch import BatchedDose, GraphedStep
    from DoseCUDA.engine import IMPTDoseEngine
    from DoseCUDA.workloads import water_slab_plan, target_plan
    from oracle import impt_oracle as O
    reps = max(3, args.steps)

    def timed(fn, n=reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n * 1e3

    def cpu_oracle(g, plan, grad=None):
        """float32 oracle on every voxel: ray-march, smoothing, dose (and the adjoint for grad)."""
        O.lib("f32")
        machine_dir = os.path.join(PKG, "DoseCUDA", "lookuptables", "protons", plan.machine_name)
        rlsp = O.rlsp_from_hu(g.HU, machine_dir)
        sp = float(np.float32(g.spacing[0]))
        t0 = time.perf_counter()
        for beam in plan.beam_list:
            model = O.beam_model_for(plan, beam)
            geom = O.beam_geometry(np.asarray(beam.iso, np.float32), np.asarray(g.origin, np.float32),
                                   beam.gantry_angle, beam.couch_angle, model.VSADX, model.VSADY)
            _, spots, layers, r80 = O.spots_and_layers(beam.spot_list, model.divergence_params)
            marr = O.ModelArrays(model)
            raw = O.raymarch(rlsp, sp, geom, precision="f32")
            w = O.smooth(raw, sp, geom, precision="f32")
            O.pb_dose(w, sp, geom, spots, layers, r80, marr, precision="f32")
            if grad is not None:
                O.pb_grad(w, grad, sp, geom, spots, layers, r80, marr, precision="f32")
        return (time.perf_counter() - t0) * 1e3

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {"metric": "small configurations: ms per plan, plain and batched paths and the CPU oracle",
           "unit": "ms", "n_gpus": 1, "higher_is_better": False, "dtype": "f32", "data": "synthetic",
           "cpu": {"kind": "port", "threads": threads, "cpu_model": cpu_model()}}
    # 320-spot water + slab plan: forward
    log("water + slab plan (320 spots)")
    g, plan = water_slab_plan()
    g.computeIMPTPlan(plan)
    t0 = time.perf_counter()
    for _ in range(reps):
        g.computeIMPTPlan(plan)
    dropin = (time.perf_counter() - t0) / reps * 1e3
    e0 = IMPTDoseEngine(g, plan, gpu_id=0, cutoff_sigma=0.0)
    e6 = IMPTDoseEngine(g, plan, gpu_id=0, cutoff_sigma=6.0)
    plain0 = timed(lambda: (e0.compute_wet(force=True), e0.dose()))
    plain6 = timed(lambda: (e6.compute_wet(force=True), e6.dose()))
    bd = BatchedDose([IMPTDoseEngine(g, plan, gpu_id=0, cutoff_sigma=6.0)], wet_mode="reach", tune=None)
    gs = GraphedStep(lambda: bd.forward(), bd.device)
    batched = timed(gs.replay)
    out["water_slab_320_spots_forward"] = {
        "dropin_host_arrays_ms": dropin, "plain_dense_ms": plain0, "plain_culled_ms": plain6,
        "batched_graph_ms": batched, "cpu_oracle_ms": cpu_oracle(g, plan)}
    # 10k-spot target plan: forward + adjoint
    log("target plan (10k spots)")
    g, plan, target, shell = target_plan()
    dev = e0.device
    tm = torch.from_numpy(target).to(dev)
    sm = torch.from_numpy(shell).to(dev)
    n_t, n_s = float(target.sum()), float(shell.sum())

    def grad_of(d):
        return torch.where(tm, (d - 2.0) * (2.0 / n_t), torch.zeros_like(d)) + \
            torch.where(sm, d * (0.2 / n_s), torch.zeros_like(d))

    res = {}
    dense_dose = None
    for name, cut in (("plain_dense_ms", 0.0), ("plain_culled_ms", 6.0)):
        e = IMPTDoseEngine(g, plan, gpu_id=0, cutoff_sigma=cut)

        def fa(e=e):
            e.compute_wet(force=True)
            e.adjoint(grad_of(e.dose()))
        res[name] = timed(fa)
        if dense_dose is None:
            dense_dose = e.dose()
    bd = BatchedDose([IMPTDoseEngine(g, plan, gpu_id=0, cutoff_sigma=6.0)], wet_mode="reach", tune=None)

    def step():
        d = bd.forward()[0]
        bd.grad_doses[0].copy_(grad_of(d))
        return bd.adjoint(None, reuse_prep=True, input_order=False)
    gs = GraphedStep(step, bd.device)
    res["batched_graph_ms"] = timed(gs.replay)
    # the oracle's adjoint takes the dL/dD of the dense GPU dose (an input here, not a result)
    gvol = grad_of(dense_dose).cpu().numpy().astype(np.float64)
    res["cpu_oracle_ms"] = cpu_oracle(g, plan, gvol)
    res["spots"] = int(sum(len(b.spot_list) for b in plan.beam_list))
    out["target_10k_spots_forward_adjoint"] = res
    print(json.dumps(out), flush=True)


def run_photon(args):
    """Photon collapsed-cone plan on the head phantom: 7 beams x 5 MLC segments (35 control points); step =
    the whole plan's dose on the device (per control point: radiological depth, TERMA, TERMA records,
    collapsed-cone sum accumulated with its output factor)."""
    print(json.dumps(photon_record(args, args.steps, args.warmup, args.photon_cpu_sample)), flush=True)


def photon_record(args, steps, warmup, cpu_sample):
    """The photon plan's measurement record (run_photon); cpu_sample 0 leaves out the CPU baseline."""
    import numpy as np
    import torch
    from DoseCUDA import photon as PH
    from DoseCUDA.workloads import imrt_head_plan
    log("building the photon workload")
    g, plan = imrt_head_plan()
    t0 = time.perf_counter()
    eng = PH.PhotonDoseEngine(g, plan, gpu_id=0)
    dev = eng.device
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    n_cp = eng.control_points()
    s = float(np.float32(g.spacing[0]))
    per_voxel = PH.samples_per_voxel(plan.beam_models[0], s)
    # work accounting (outside the timed region): voxels with TERMA > 0.01 per control point
    in_field = []
    for dmodel, cps in eng.beams:
        for cpc, _, _ in cps:
            PH.raytrace(eng.ws, cpc)
            PH.terma(eng.ws, dmodel, cpc)
            in_field.append(int((eng.ws.terma > 0.01).sum().item()))
    samples = per_voxel * sum(in_field)
    steps = max(1, steps)

    def plan_ms(pipeline):
        for _ in range(max(1, warmup)):
            eng.compute(pipeline)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.compute(pipeline)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps * 1e3

    # one stream (every kernel in order), then the timed form: depth and TERMA of the next
    # control point on a side stream beside the current cone sum
    ms_one = plan_ms(False)
    ms = plan_ms(True)
    log("photon plan: %.2f ms per plan (%d control points; %.2f ms on one stream)" % (ms, n_cp, ms_one))
    # per-kernel HIP events on the launch stream (torch's current stream), one extra plan on one stream
    # (each kernel alone on the GPU)
    st = torch.cuda.current_stream(dev)
    names = ("k_ph_raytrace", "k_ph_terma+k_ph_records_from", "k_ph_ccc")
    acc = {k: 0.0 for k in names}
    ev = []
    last = None
    for (dmodel, cps), bd in zip(eng.beams, eng.beam_doses):
        for cpc, of, key in cps:
            e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
            e[0].record(st)
            if key != last:   # as PhotonDoseEngine.compute: one ray-trace per beam geometry
                PH.raytrace(eng.ws, cpc)
                last = key
            e[1].record(st)
            PH.terma(eng.ws, dmodel, cpc)
            e[2].record(st)
            PH.ccc(eng.ws, dmodel, cpc, bd, of, True)
            e[3].record(st)
            ev.append(e)
    torch.cuda.synchronize()
    for e in ev:
        for k, name in enumerate(names):
            acc[name] += e[k].elapsed_time(e[k + 1])
    n_trace = eng.depth_traces()
    ccc_ms = acc["k_ph_ccc"] / n_cp
    flop = 64.0 * samples / n_cp   # per launch: 64 FLOP per cone-ray sample (two trilinear fields, recursion)
    nvox = int(np.prod(g.HU.shape))
    # CPU baseline: float32 oracle on the in-field voxels of the first control points (each with the GPU's
    # TERMA volume as input), up to cpu_sample voxels in all
    cpu = None
    if cpu_sample > 0:
        from oracle import imrt_oracle as OR
        rho = OR.density_from_hu(g.HU, os.path.join(PKG, "DoseCUDA", "lookuptables", "photons", plan.machine_name))
        rng = np.random.default_rng(0)
        labels = list(np.asarray(plan.dicom_energy_label).astype(str))
        n_vox = n_samp = 0
        tc = 0.0
        used = 0
        for (dmodel, cps), beam in zip(eng.beams, plan.beam_list):
            for (cpc, _, _), cp in zip(cps, beam.cp_list):
                if n_vox >= cpu_sample:
                    break
                PH.raytrace(eng.ws, cpc)
                PH.terma(eng.ws, dmodel, cpc)
                t32 = eng.ws.terma.cpu().numpy()
                el = eng.ws.electron.cpu().numpy().astype(np.float64)
                model = plan.beam_models[labels.index(str(beam.dicom_energy_label))]
                geom = OR.photon_geometry(np.asarray(cp.iso, np.float32), np.asarray(g.origin, np.float32), cp.ga,
                                          cp.ta, cp.ca, np.float32(model.primary_source_distance))
                hot = np.flatnonzero(t32.reshape(-1) > 0.01)
                idx = np.sort(rng.choice(hot, size=min(cpu_sample - n_vox, hot.size), replace=False)).astype(np.int64)
                tc0 = time.perf_counter()
                dep = OR.raytrace(rho, s, geom, idx=idx, precision="f32")
                OR.terma(dep, rho.shape, s, geom, model, OR.mlc_rows(cp.mlc), idx=idx, precision="f32")
                OR.ccc(t32, rho, el.reshape(-1)[idx], s, geom, model.kernel, np.float32(model.mu_calibration),
                       np.float32(cp.mu), idx=idx, precision="f32")
                tc += time.perf_counter() - tc0
                n_vox += idx.size
                n_samp += idx.size * per_voxel
                used += 1
        cpu = {"value": n_samp / tc / (samples / n_cp), "unit": "control points/s", "cores": os.cpu_count(),
               "kind": "port", "cpu_model": cpu_model(),
               "sample": "%d in-field voxels of the first %d control points: depth, TERMA and the cone sum at "
                         "them (float32 oracle build, OpenMP on %d threads), %.3g cone samples in %.1f s, scaled "
                         "to the plan's %.3g samples per control point" %
                         (n_vox, used, os.cpu_count(), n_samp, tc, samples / n_cp)}
    rec = {
        "metric": "photon collapsed-cone plan dose: ms per plan and control points/s",
        "value": n_cp / (ms * 1e-3),
        "unit": "control points/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "head phantom 160x200x200 @ 2 mm on an IMRTDoseGrid, VarianTrueBeamHF 6 MV, "
                               "7 beams (gantry 0..308.6 deg) x 5 step-and-shoot MLC segments = %d control points; "
                               "step = the plan's dose on the device (density from HU once, at set-up; depth "
                               "and TERMA of the next control point on a side stream beside the cone sum)" % n_cp,
                   "voxels": nvox, "setup_s": t_setup},
        "one_stream_ms_per_plan": ms_one,
        "in_field_voxels_per_cp": float(np.mean(in_field)),
        "samples_per_in_field_voxel": per_voxel,
        "cone_samples_per_s": samples / (ms * 1e-3),
        "kernels": {"k_ph_raytrace": {"avg_ms_per_launch": acc["k_ph_raytrace"] / n_trace,
                                      "launches_per_plan": n_trace},
                    "k_ph_terma+k_ph_records_from": {"avg_ms_per_cp": acc["k_ph_terma+k_ph_records_from"] / n_cp},
                    "k_ph_ccc": {"avg_ms_per_cp": acc["k_ph_ccc"] / n_cp}},
        "roofline": {"kernel": "k_ph_ccc", "bound": "valu", "achieved": flop / (ccc_ms * 1e-3) / 1e12,
                     "peak": 157.3, "unit": "TFLOP/s", "frac": flop / (ccc_ms * 1e-3) / 1e12 / 157.3,
                     "traffic": load_traffic("k_ph_ccc"), "avg_launch_ms": ccc_ms, "flop_per_sample": 64.0,
                     "samples_per_launch": samples / n_cp},
        "cpu_baseline": cpu,
    }
    return rec


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plans", type=int, default=4, help="head-phantom plans per batch (K)")
    ap.add_argument("--streams", type=int, default=2,
                    help="head workload: the timed step runs the plans as this many sub-batches on concurrent "
                         "streams (1 = one batch on one stream)")
    ap.add_argument("--cutoff-sigma", type=float, default=6.0)
    ap.add_argument("--plain-steps", type=int, default=2, help="steps of the plain (dense, per-beam) path; 0 = skip")
    ap.add_argument("--cpu-sample", type=int, default=800000,
                    help="voxels in the CPU-baseline sample (about 10-15 s of CPU work on 16 cores); 0 = skip")
    ap.add_argument("--no-tune", action="store_true", help="library default schedules instead of the autotuner's")
    ap.add_argument("--workload", choices=("head", "uneven", "body", "photon", "small"), default="head",
                    help="head: the headline metric (default); uneven: 64 uneven plans batched; body: 512x512x300 "
                         "optimisation loop (each prints its own JSON line)")
    ap.add_argument("--iters", type=int, default=50, help="optimisation iterations of the body workload")
    ap.add_argument("--uneven-plans", type=int, default=64, help="plans of the uneven workload")
    ap.add_argument("--photon-cpu-sample", type=int, default=6000000,
                    help="photon workload: in-field voxels of the CPU baseline's sample (the first control "
                         "points' in-field voxels, about 10 s of CPU work on 16 cores)")
    ap.add_argument("--no-photon", action="store_true",
                    help="head workload: leave out the photon plan measurement appended to its record")
    ap.add_argument("--eager-only", action="store_true",
                    help="time eager batched steps only (no graph), e.g. under a profiler")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run 3 device-to-device copies of this many MiB (calibrates FETCH_SIZE/WRITE_SIZE "
                         "counters against a copy of known size) and exit")
    args = ap.parse_args()
    if args.copy_probe_mib > 0:
        import torch
        x = torch.ones(args.copy_probe_mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": x.numel() * 4}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the dose of one plan is computed "
              "on one device; see the sharding design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from DoseCUDA import _native as N
    from DoseCUDA import device as D
    from DoseCUDA.batch import BatchedDose, StreamedBatches
    from DoseCUDA.engine import IMPTDoseEngine

    assert torch.cuda.is_available(), "bench.py needs a HIP GPU"
    if args.workload == "uneven":
        return run_uneven(args)
    if args.workload == "photon":
        return run_photon(args)
    if args.workload == "body":
        return run_body(args)
    if args.workload == "small":
        return run_small(args)
    K = args.plans
    log("building workload: %d head-phantom plans" % K)
    g, plans = plan_copies(K)
    engines = [IMPTDoseEngine(g, p, gpu_id=0, cutoff_sigma=args.cutoff_sigma) for p in plans]
    dev = engines[0].device
    tmask, bmask = masks(g, dev)
    n_t = float(tmask.sum().item())
    n_b = float(bmask.sum().item())
    lgrad = LossGrad(tmask, bmask, n_t, n_b)
    nvox = int(np.prod(engines[0].shape))
    nb = len(engines[0].beams)
    n_spots = sum(b.spots.n_spots for b in engines[0].beams)

    # raw-WET volumes of the engines' own (rather than workspace scratch), so the voxels the reach pass
    # marches can be counted
    for e in engines:
        for b in e.beams:
            b.wet_raw = torch.zeros(e.shape, dtype=torch.float32, device=dev)
    t_tune = time.perf_counter()
    bd = BatchedDose(engines, wet_mode="reach", tune=None if args.no_tune else "auto",
                     tune_grad_fn=lambda doses: [lgrad(d) for d in doses], tune_grad_key="head-loss")
    t_tune = time.perf_counter() - t_tune
    from DoseCUDA.batch import tune_dict
    log("schedule %s (set-up incl. autotuning %.1f s)" % (tune_dict(bd.tune), t_tune))
    mus = [b.mu0.clone() for e in engines for b in e.beams]
    xys = [b.xy0.clone() for e in engines for b in e.beams]

    def step():
        doses = bd.forward(mus, xys)
        for d, gd in zip(doses, bd.grad_doses):
            lgrad(d, out=gd)
        return doses, bd.adjoint(None, reuse_prep=True, input_order=False)

    log("batch items %s" % bd.item_counts())
    ms_graph = ms_graph_1 = None
    n_streams = max(1, min(args.streams, K))
    if not args.eager_only:
        log("capturing the batched step (one batch, one stream)")
        ms_graph_1, gs = _graph_ms(step, dev, args.warmup, args.steps)
        del gs
        ms_graph = ms_graph_1
        if n_streams > 1:
            # the same step as sub-batches of the plans on concurrent streams (StreamedBatches)
            t_s = time.perf_counter()
            sb = StreamedBatches(engines, n_streams=n_streams, grad_fn=lambda p, d, out: lgrad(d, out=out),
                                 wet_mode="reach", tune=None if args.no_tune else "auto", tune_grad_key="head-loss")
            t_tune += time.perf_counter() - t_s
            log("%d sub-batches %s on %d streams, schedules %s" %
                (len(sb.subs), sb.groups, n_streams, [tune_dict(x.tune) for x in sb.subs]))

            def step_streamed():
                return sb.step(mus, xys, input_order=False)

            log("timing %d graph replays of the streamed step" % args.steps)
            ms_graph, gs = _graph_ms(step_streamed, dev, args.warmup, args.steps)
            del gs, sb
            torch.cuda.empty_cache()

    # eager batched steps with per-kernel HIP events (same kernels, same stream)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * N.DH_TIME_N)]
    for e in ev:
        e.record()
    torch.cuda.synchronize()
    handles = (ctypes.c_void_p * len(ev))(*[e.cuda_event for e in ev])
    N.check(bd.lib.dh_batch_timing(ctypes.byref(bd.batch), handles), "dh_batch_timing")
    for _ in range(args.warmup if args.eager_only else 1):
        step()
    torch.cuda.synchronize()
    kt = {k: [] for k in range(N.DH_TIME_N)}
    n_eager = args.steps if args.eager_only else max(3, min(args.steps, 10))
    log("timing %d eager batched steps with kernel events" % n_eager)
    t0 = time.perf_counter()
    for _ in range(n_eager):
        step()
        for k in range(N.DH_TIME_N):
            kt[k].append((ev[2 * k], ev[2 * k + 1]))
        torch.cuda.synchronize()   # events are re-recorded by the next step
        for k in range(N.DH_TIME_N):
            a, b = kt[k][-1]
            kt[k][-1] = a.elapsed_time(b)
    ms_eager = (time.perf_counter() - t0) / n_eager * 1e3
    N.check(bd.lib.dh_batch_timing(ctypes.byref(bd.batch), None), "dh_batch_timing")
    kms = {N.TIMED_KERNELS[k]: float(np.mean(v)) for k, v in kt.items()}
    ms_step = ms_graph if ms_graph is not None else ms_eager

    # voxels the reach-limited ray-march computes per step: raw WET poisoned with NaN, one step, count
    for e in engines:
        for b in e.beams:
            b.wet_raw.fill_(float("nan"))
    step()
    torch.cuda.synchronize()
    marched = int(sum(int(torch.isfinite(b.wet_raw).sum().item()) for e in engines for b in e.beams))
    log("reach pass marched %d of %d voxel-beams (%.1f%%)" % (marched, nvox * nb * K, 100.0 * marched / (nvox * nb * K)))

    # K = 1: one plan on the batched graph path (its own tuned schedule)
    k1_ms = None
    if not args.eager_only:
        b1 = BatchedDose([engines[0]], wet_mode="reach", tune=None if args.no_tune else "auto",
                         tune_grad_fn=lambda doses: [lgrad(d) for d in doses], tune_grad_key="head-loss")
        mus1 = [b.mu0.clone() for b in engines[0].beams]
        xys1 = [b.xy0.clone() for b in engines[0].beams]

        def step1():
            doses = b1.forward(mus1, xys1)
            lgrad(doses[0], out=b1.grad_doses[0])
            return doses, b1.adjoint(None, reuse_prep=True, input_order=False)

        log("timing %d graph replays of one plan (K = 1)" % args.steps)
        k1_ms, gs1 = _graph_ms(step1, dev, args.warmup, args.steps)
        del gs1, b1
        torch.cuda.empty_cache()

    # work counts per plan (plan copy 0): evaluated pairs with the culled path, reference pairs with the
    # plain path's dL/dD; both on the full WET (the dose kernels see the same values where they evaluate)
    e0 = engines[0]
    step()
    gd_cull = bd.grad_doses[0].clone()
    e0.compute_wet(force=True)      # full WET into plan 0's volumes (the batch's next step redoes its own)
    counts_cull = torch.zeros(4, dtype=torch.int64, device=dev)
    for b in e0.beams:
        D.pb_count_pairs(b.wet, gd_cull, e0.grid, b.geo, b.dmodel, b.spots, b.ws, b.ws_bytes,
                         cutoff_sigma=args.cutoff_sigma, counts=counts_cull)

    # plain path: one plan, eager per-beam calls, full WET every step, dense forward + separate adjoint
    plain_ms = None
    gd_dense = None

    def plain_step(cutoff=0.0):
        for b in e0.beams:
            b.wet_raw, b.wet = D.wet_from_padded(e0.rlsp_padded, e0.grid, b.geo, dev, raw_out=b.wet_raw, out=b.wet)
        dose = torch.empty(e0.shape, dtype=torch.float32, device=dev)
        for k, b in enumerate(e0.beams):
            b.spots.set_weights_positions(b.mu0, b.xy0)
            D.pb_forward(b.wet, dose, e0.grid, b.geo, b.dmodel, b.spots, b.ws, b.ws_bytes, cutoff_sigma=cutoff,
                         accumulate=k > 0)
        gd = lgrad(dose)
        for b in e0.beams:
            D.pb_adjoint(b.wet, gd, e0.grid, b.geo, b.dmodel, b.spots, b.ws, b.ws_bytes, cutoff_sigma=cutoff)
        return dose, gd

    def plain_time(cutoff):
        plain_step(cutoff)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        for _ in range(args.plain_steps):
            out = plain_step(cutoff)
        torch.cuda.synchronize()
        return (time.perf_counter() - tp) / args.plain_steps * 1e3, out

    plain_culled_ms = None
    if args.plain_steps > 0:
        log("plain path, culled (cutoff %g), %d steps" % (args.cutoff_sigma, args.plain_steps))
        plain_culled_ms, _ = plain_time(args.cutoff_sigma)
        log("plain path, dense, %d steps" % args.plain_steps)
        plain_ms, (_, gd_dense) = plain_time(0.0)
    else:
        # the plain path's dL/dD: dose of every pair (cutoff 0), as the plain steps would compute it
        e0.compute_wet(force=True)
        cut0 = e0.cutoff_sigma
        e0.cutoff_sigma = 0.0
        dose_d = e0.dose()
        e0.cutoff_sigma = cut0
        gd_dense = lgrad(dose_d)
    counts_ref = torch.zeros(4, dtype=torch.int64, device=dev)
    for b in e0.beams:
        D.pb_count_pairs(b.wet, gd_dense, e0.grid, b.geo, b.dmodel, b.spots, b.ws, b.ws_bytes,
                         cutoff_sigma=args.cutoff_sigma, counts=counts_ref)
    # drop-in path with host arrays in and out (PCIe transfers included): computeIMPTPlan of plan copy 0,
    # forward only (the reference's call has no gradient)
    dropin_ms = None
    if args.plain_steps > 0:
        log("drop-in computeIMPTPlan (host arrays, PCIe included) %d calls" % args.plain_steps)
        g.computeIMPTPlan(plans[0])
        td = time.perf_counter()
        for _ in range(args.plain_steps):
            g.computeIMPTPlan(plans[0])
        dropin_ms = (time.perf_counter() - td) / args.plain_steps * 1e3
    cc = [int(v) for v in counts_cull.cpu().tolist()]
    cr = [int(v) for v in counts_ref.cpu().tolist()]
    ref_pairs_plan = cr[0] + cr[2]           # reference-equivalent pairs per plan
    fwd_eval, adj_eval = cc[1] * K, cc[3] * K  # evaluated per batched step

    # roofline of the dominant kernel (largest time per batched step)
    rows = {
        "k_pb_adjoint_batch": ("valu", adj_eval * ADJ_FLOP_PER_PAIR, ADJ_FLOP_PER_PAIR, adj_eval),
        "k_pb_forward_batch": ("valu", fwd_eval * FWD_FLOP_PER_PAIR, FWD_FLOP_PER_PAIR, fwd_eval),
    }
    dom_name = max(kms, key=lambda k: kms[k])
    if dom_name in rows:
        bound, flop, fpp, pairs = rows[dom_name]
        achieved = flop / (kms[dom_name] * 1e-3) / 1e12
        roofline = {"kernel": dom_name, "bound": bound, "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                    "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": load_traffic(dom_name),
                    "avg_launch_ms": kms[dom_name], "flop_per_pair": fpp, "pairs_per_launch": pairs}
    else:
        # WET kernels, algorithmic bytes per voxel and beam: ray-march 8 (RLSP read 4 + raw WET write 4) over
        # the voxels the reach pass marched; smoothing 8 (raw WET read + smoothed write) over every voxel
        if dom_name == "k_wet_raymarch_batch":
            units, per_vox, what = marched, 8.0, "marched voxel-beams (reach pass)"
        else:
            units, per_vox, what = nvox * nb * K, 8.0, "voxel-beams"
        byts = per_vox * units
        achieved = byts / (kms[dom_name] * 1e-3) / 1e9
        roofline = {"kernel": dom_name, "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                    "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": load_traffic(dom_name),
                    "avg_launch_ms": kms[dom_name], "bytes_per_launch": byts, "bytes_per_unit": per_vox,
                    "units_per_launch": units, "unit_kind": what}
    kernels = {}
    for name, ms in kms.items():
        r = {"avg_ms_per_step": ms}
        if name in rows:
            _, flop, _, pairs = rows[name]
            r.update({"pairs": pairs, "tflops": flop / (ms * 1e-3) / 1e12,
                      "frac_fp32_peak": flop / (ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS})
        kernels[name] = r

    cpu = None
    if args.cpu_sample > 0:
        log("cpu baseline on %d voxels" % args.cpu_sample)
        cpu = cpu_baseline(g, plans[0], e0, gd_dense.detach().cpu().numpy(), args.cpu_sample)

    ms_plan = ms_step / K
    rec = {
        "metric": METRIC,
        "value": ref_pairs_plan * K / (ms_step * 1e-3),
        "unit": "spot·voxel evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "batch of %d head-phantom plans (160x200x200 @ 2 mm, 3 beams at gantry 0/120/240, "
                        "40 layers x 81 spots = 9720 spots each); step = one HIP-graph replay: reach-limited "
                        "WET, plan doses, dL/dD, adjoint (weights and positions) of every plan" % K,
            "plans": K, "voxels_per_plan": nvox, "spots_per_plan": n_spots, "beams_per_plan": nb,
            "cutoff_sigma": args.cutoff_sigma, "path": "graph" if ms_graph is not None else "eager",
            "streams": n_streams if ms_graph is not None else 1,
            "schedule": tune_dict(bd.tune), "setup_and_autotune_s": t_tune,
        },
        "one_stream_graph_ms_per_step": ms_graph_1,
        "ms_per_plan": ms_plan,
        "plain_ms_per_plan": plain_ms,
        "plain_culled_ms_per_plan": plain_culled_ms,
        "batched_k1_ms_per_plan": k1_ms,
        "speedup_vs_plain_culled_per_plan": (plain_culled_ms / ms_plan) if plain_culled_ms else None,
        "speedup_k1_vs_plain_culled": (plain_culled_ms / k1_ms) if (plain_culled_ms and k1_ms) else None,
        "marched_voxel_beams_per_step": marched,
        "dropin_forward_ms_per_plan_host_arrays": dropin_ms,
        "speedup_vs_plain_per_plan": (plain_ms / ms_plan) if plain_ms else None,
        "eager_ms_per_step": ms_eager,
        "kernels": kernels,
        "pairs_per_plan": {"forward_reference": cr[0], "forward_evaluated": cc[1],
                           "adjoint_reference": cr[2], "adjoint_evaluated": cc[3]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if not (args.no_photon or args.eager_only):
        # the photon collapsed-cone plan (the --workload photon measurement, 5 plans, no CPU baseline),
        # after the timed proton steps; left out of profiling runs (--eager-only)
        try:
            ph = photon_record(args, 5, 1, 0)
            rec["photon"] = {k: ph[k] for k in ("metric", "value", "unit", "ms_per_step", "one_stream_ms_per_plan",
                                                "cone_samples_per_s", "kernels", "roofline")}
            rec["photon"]["workload"] = ph["config"]["workload"]
        except Exception as e:   # the proton record is printed whatever happens here
            rec["photon"] = {"error": repr(e)}
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
