#!/usr/bin/env python3
"""Forward + adjoint of the 10k-spot head phantom on one MI355X.

Workload (synthetic, seed 2): 160 x 200 x 200 (z, y, x) HU grid at 2 mm --
ellipsoidal soft tissue N(40, 20) clipped to [-100, 100], 8 mm bone shell
U(700, 1200), two 15 mm air spheres, air outside; 3 beams at gantry 0/120/240;
40 energy layers (RS0 ids 20..59) x 81 spots (3 mm grid in a 15 mm disk) per
beam = 9720 spots, weights U(0, 1).

One step = per beam: WET ray-march + smoothing; plan dose (forward); dL/dD of
L = mean_target (D - 2)^2 + 0.1 mean_body D^2 (target: 25 mm sphere at the
isocentre, body: HU > -500); adjoint (dL/dweight, dL/d(x, y)) per beam.

value = spot.voxel evaluations per second, counted as the reference does
(every spot of a layer for every voxel with WET <= 1.1 R80, forward plus
adjoint over voxels with dL/dD != 0), per step, over the timed steps.

Prints ONE JSON line.
"""

import argparse
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


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def build_problem(engine_cls, cutoff):
    import numpy as np
    import torch
    from DoseCUDA.workloads import head_plan, voxel_centres
    g, plan = head_plan(seed=2)
    eng = engine_cls(g, plan, gpu_id=0, cutoff_sigma=cutoff)
    z, y, x = voxel_centres(g)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    target = (X ** 2 + Y ** 2 + Z ** 2) <= 25.0 ** 2
    body = g.HU > -500.0
    dev = eng.device
    tmask = torch.from_numpy(target).to(dev)
    bmask = torch.from_numpy(body).to(dev)
    return g, plan, eng, tmask, bmask


def loss_grad(dose, tmask, bmask, n_t, n_b):
    """dL/dD of mean_target (D - 2)^2 + 0.1 mean_body D^2."""
    import torch
    g = torch.where(tmask, (dose - 2.0) * (2.0 / n_t), torch.zeros_like(dose))
    return g + torch.where(bmask, dose * (0.2 / n_b), torch.zeros_like(dose))


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
        t0 = time.perf_counter()
        O.raymarch(rlsp, s, geom, idx=idx, precision="f32")
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
    return {"value": pairs / t_total, "unit": "spot·voxel evals/s", "cores": threads, "kind": "port",
            "sample": "%d random voxels of the same phantom, 3 beams: ray-march of the sampled voxels + "
                      "forward dose + adjoint at them (float32 oracle build, OpenMP); smoothing omitted; "
                      "%d spot.voxel pairs in %.1f s" % (len(idx), pairs, t_total)}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cutoff-sigma", type=float, default=6.0)
    ap.add_argument("--plain-steps", type=int, default=2, help="steps of the dense (cutoff 0) path; 0 = skip")
    ap.add_argument("--cpu-sample", type=int, default=60000, help="voxels in the CPU-baseline sample; 0 = skip")
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
    from DoseCUDA.engine import IMPTDoseEngine
    from DoseCUDA import device as D

    assert torch.cuda.is_available(), "bench.py needs a HIP GPU"
    log("building workload")
    g, plan, eng, tmask, bmask = build_problem(IMPTDoseEngine, args.cutoff_sigma)
    n_t = float(tmask.sum().item())
    n_b = float(bmask.sum().item())
    dev = eng.device
    mus = [b.mu0.clone() for b in eng.beams]
    xys = [b.xy0.clone() for b in eng.beams]
    nb = len(eng.beams)
    n_spots = sum(b.spots.n_spots for b in eng.beams)
    nvox = int(np.prod(eng.shape))

    ev = {k: [] for k in ("wet", "fwd", "adj")}

    def step(record):
        def rec():
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            return e
        for b in eng.beams:
            e0 = rec() if record else None
            b.wet_raw, b.wet = D.wet_from_padded(eng.rlsp_padded, eng.grid, b.geo, dev, raw_out=b.wet_raw,
                                                 out=b.wet)
            if record:
                ev["wet"].append((e0, rec()))
        dose = torch.empty(eng.shape, dtype=torch.float32, device=dev)
        for k, b in enumerate(eng.beams):
            b.spots.set_weights_positions(mus[k], xys[k])
            e0 = rec() if record else None
            D.pb_forward(b.wet, dose, eng.grid, b.geo, b.dmodel, b.spots, b.ws, b.ws_bytes,
                         cutoff_sigma=eng.cutoff_sigma, accumulate=k > 0)
            if record:
                ev["fwd"].append((e0, rec()))
        gd = loss_grad(dose, tmask, bmask, n_t, n_b)
        out = []
        for k, b in enumerate(eng.beams):
            e0 = rec() if record else None
            gm, gx = D.pb_adjoint(b.wet, gd, eng.grid, b.geo, b.dmodel, b.spots, b.ws, b.ws_bytes,
                                  cutoff_sigma=eng.cutoff_sigma)
            if record:
                ev["adj"].append((e0, rec()))
            out.append((gm, gx))
        return dose, gd, out

    log("warmup %d" % args.warmup)
    for _ in range(args.warmup):
        dose, gd, _ = step(False)
    torch.cuda.synchronize()

    # work counts (reference-equivalent and evaluated pairs) for this step
    counts = torch.zeros(4, dtype=torch.int64, device=dev)
    for b in eng.beams:
        D.pb_count_pairs(b.wet, gd, eng.grid, b.geo, b.dmodel, b.spots, b.ws, b.ws_bytes,
                         cutoff_sigma=eng.cutoff_sigma, counts=counts)
    counts = [int(v) for v in counts.cpu().tolist()]
    dense_pairs = counts[0] + counts[2]

    log("timing %d steps" % args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dose, gd, grads = step(True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed / args.steps * 1e3

    def avg_ms(key):
        return float(np.mean([a.elapsed_time(b) for a, b in ev[key]]))

    t_wet, t_fwd, t_adj = avg_ms("wet"), avg_ms("fwd"), avg_ms("adj")
    fwd_tflops = counts[1] / nb * FWD_FLOP_PER_PAIR / (t_fwd * 1e-3) / 1e12
    adj_tflops = counts[3] / nb * ADJ_FLOP_PER_PAIR / (t_adj * 1e-3) / 1e12
    if t_fwd >= t_adj:
        dom = {"kernel": "k_pb_forward", "bound": "valu", "achieved": fwd_tflops, "peak": PEAK_FP32_TFLOPS,
               "unit": "TFLOP/s", "frac": fwd_tflops / PEAK_FP32_TFLOPS, "traffic": None,
               "avg_launch_ms": t_fwd, "flop_per_pair": FWD_FLOP_PER_PAIR, "pairs_per_launch": counts[1] / nb}
    else:
        dom = {"kernel": "k_pb_adjoint", "bound": "valu", "achieved": adj_tflops, "peak": PEAK_FP32_TFLOPS,
               "unit": "TFLOP/s", "frac": adj_tflops / PEAK_FP32_TFLOPS, "traffic": None,
               "avg_launch_ms": t_adj, "flop_per_pair": ADJ_FLOP_PER_PAIR, "pairs_per_launch": counts[3] / nb}
    # ray-march + smoothing: compulsory bytes per voxel per beam: RLSP read 4 + raw WET write 4 + read 4 +
    # smoothed write 4
    wet_gbs = 16.0 * nvox / (t_wet * 1e-3) / 1e9
    roof_wet = {"kernel": "k_wet_raymarch+k_wet_smooth", "bound": "hbm", "achieved": wet_gbs,
                "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": wet_gbs / PEAK_HBM_GBS, "traffic": None,
                "avg_launch_ms": t_wet}

    # plain path: every spot for every in-range voxel (cutoff 0), same step
    plain_ms = None
    if args.plain_steps > 0:
        log("plain (dense) path %d steps" % args.plain_steps)
        eng.cutoff_sigma = 0.0
        step(False)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        for _ in range(args.plain_steps):
            step(False)
        torch.cuda.synchronize()
        plain_ms = (time.perf_counter() - tp) / args.plain_steps * 1e3
        eng.cutoff_sigma = args.cutoff_sigma

    cpu = None
    if args.cpu_sample > 0:
        log("cpu baseline on %d voxels" % args.cpu_sample)
        cpu = cpu_baseline(g, plan, eng, gd.detach().cpu().numpy(), args.cpu_sample)

    rec = {
        "metric": METRIC,
        "value": dense_pairs / (ms_step * 1e-3),
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
            "workload": "head phantom 160x200x200 @ 2 mm, 3 beams (gantry 0/120/240), 40 layers x 81 spots = "
                        "9720 spots; step = WET ray-march + smoothing, plan dose, dL/dD, adjoint (weights and "
                        "positions), all beams",
            "voxels": nvox, "spots": n_spots, "beams": nb, "cutoff_sigma": args.cutoff_sigma,
            "path": "plain per-plan (one plan per launch sequence; no batching yet)",
        },
        "breakdown_ms": {"wet_per_beam": t_wet, "forward_per_beam": t_fwd, "adjoint_per_beam": t_adj},
        "pairs_per_step": {"forward_reference": counts[0], "forward_evaluated": counts[1],
                           "adjoint_reference": counts[2], "adjoint_evaluated": counts[3]},
        "dense_path_ms_per_step": plain_ms,
        "roofline": dom,
        "roofline_wet": roof_wet,
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
