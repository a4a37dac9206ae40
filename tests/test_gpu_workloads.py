"""The benchmark workloads on the GPU vs the FP64 oracle.

Each test runs the path bench.py times -- ``BatchedDose`` with reach-limited
WET and a 6-sigma cutoff -- and, where stated, the full-WET dense variant
(cutoff 0, every pair as the reference evaluates it) and the drop-in
``computeIMPTPlan`` (default: dense).  The batch is compared with the oracle
directly, not through the single-beam path:

* small plans in one batch (1 mm, 2 mm and 2.5 mm grids; gantries 0/120/240 at
  2 mm, where the 60-degree smoothing directions meet exact half-voxel ties,
  impt_jax.py:387-389; couch 30; n_fractions 2): the full-volume oracle
  (``O.compute_plan``, ``O.pb_grad``) -- raw and smoothed WET, dose, gamma,
  gradients;
* head phantom, 160x200x200 @ 2 mm, gantries 0/120/240, 9720 spots;
* 64 uneven plans (cube sides 64..256 @ 2 mm, 1-4 beams, 500-40k spots);
* body phantom, 300x512x512 @ 1 mm, 4 beams, 63k spots, after projected-
  gradient weight iterations;

the large ones at sampled voxels (``oracle_sample.OraclePlan``: the oracle
marches the rays each sampled voxel's smoothed WET needs): voxels above 10% of
D_max, uniform voxels and the dose maximum.  Gradients are checked for a dL/dD
that is zero off the sample, so the oracle's sum over voxels is exact there.

Tolerances (float32 kernels vs float64 oracle):
  * dose: max |dD| / D_max <= 1e-3 over the compared voxels (the bar is 5e-3);
    2%/2 mm global gamma at the reference voxels >= 0.995 (searching the full
    GPU volume);
  * smoothed and raw WET: <= 2e-4 cm where the pass computed them; a voxel the
    reach pass skipped must get < 1e-6 D_max of that beam's oracle dose;
  * gradients: max |dg| / max |g| <= 1e-3 for dL/dmu and for dL/d(x, y).
Measured values are printed with the case name.
"""

import gc

import numpy as np
import pytest
import torch

from oracle import impt_oracle as O
from oracle_sample import OraclePlan, gamma_at, rel_err, sample_voxels

pytestmark = pytest.mark.gpu

DOSE_TOL = 1e-3
GAMMA_MIN = 0.995
WET_TOL = 2e-4
GRAD_TOL = 1e-3


def _say(msg):
    print("[gpu-workloads] " + msg, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


def _free():
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _np(t):
    return t.detach().cpu().numpy()


def _check_wet(case, bd_engine, op, idx, dmax_ref, reach):
    """GPU WET of each beam at the sample vs the oracle.  In reach mode a voxel
    whose WET was not computed (1e30) must get no dose from that beam."""
    worst = 0.0
    for k, b in enumerate(bd_engine.beams):
        w = _np(b.wet).reshape(-1)[idx]
        done = w < 1e29
        if reach:
            skipped = ~done
            if skipped.any():
                bd_ref = op.beam_dose(k)[skipped] * op.n_fx
                assert bd_ref.max() < 1e-6 * dmax_ref, (case, k, float(bd_ref.max() / dmax_ref))
        else:
            assert done.all()
            raw = _np(b.wet_raw).reshape(-1)[idx]
            worst = max(worst, float(np.abs(raw - op.raw[k]).max()))
        if done.any():
            worst = max(worst, float(np.abs(w[done] - op.wet[k][done]).max()))
    return worst


def _raw_buffers(engines):
    """Raw-WET volumes of the engines' own, so a batch writes its ray-march
    output there (rather than into workspace scratch) for the WET checks."""
    for e in engines:
        for b in e.beams:
            if b.wet_raw is None:
                b.wet_raw = torch.zeros(e.shape, dtype=torch.float32, device=e.device)


def _gpu_grads(bd, shapes, idxs, gvals, dev):
    grads = []
    for sh, ix, gv in zip(shapes, idxs, gvals):
        gvol = torch.zeros(int(np.prod(sh)), dtype=torch.float32, device=dev)
        gvol[torch.from_numpy(ix).to(dev)] = torch.from_numpy(gv.astype(np.float32)).to(dev)
        grads.append(gvol.view(sh))
    res = bd.adjoint(grads)
    torch.cuda.synchronize()
    return [(_np(gm).astype(np.float64), _np(gxy).astype(np.float64)) for gm, gxy in res]


def _grad_errs(got, want):
    em = max(rel_err(g[0], w[0]) for g, w in zip(got, want))
    ex = max(rel_err(g[1], w[1]) for g, w in zip(got, want))
    return em, ex


# ---------------------------------------------------------------------------
# small plans in one batch vs the full-volume oracle
def _small_plans():
    from DoseCUDA import IMPTPlan
    from DoseCUDA.workloads import _beam, _grid, grid_spots, voxel_centres
    out = []
    # 2 mm head-like phantom: ellipsoid of soft tissue with a 4 mm bone shell and
    # an air cavity, air outside; gantries 0/120/240
    shape = (34, 40, 38)
    g = _grid(shape, 2.0)
    rng = np.random.default_rng(12)
    z, y, x = voxel_centres(g)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    r_out = (X / 32.0) ** 2 + (Y / 34.0) ** 2 + (Z / 30.0) ** 2
    r_in = (X / 28.0) ** 2 + (Y / 30.0) ** 2 + (Z / 26.0) ** 2
    hu = np.full(shape, -1000.0, np.float32)
    hu[r_out <= 1.0] = np.clip(rng.normal(40.0, 20.0, size=int((r_out <= 1.0).sum())), -100.0, 100.0)
    shell = (r_out <= 1.0) & (r_in > 1.0)
    hu[shell] = rng.uniform(700.0, 1200.0, size=int(shell.sum()))
    hu[(X - 8.0) ** 2 + (Y - 6.0) ** 2 + (Z - 4.0) ** 2 <= 7.0 ** 2] = -1000.0
    g.HU = hu
    plan = IMPTPlan()
    c = np.arange(-9.0, 9.1, 3.0)
    for ga in (0.0, 120.0, 240.0):
        sl = grid_spots([10, 16, 22], c, c, disk_radius=9.0)
        sl[:, 2] = rng.uniform(0.2, 1.0, size=sl.shape[0]).astype(np.float32)
        plan.addBeam(_beam(ga, sl))
    out.append((g, plan))
    # 2.5 mm heterogeneous box, oblique beams with couch, two fractions
    shape = (26, 30, 28)
    g = _grid(shape, 2.5)
    rng = np.random.default_rng(31)
    hu = rng.normal(30.0, 25.0, size=shape).astype(np.float32)
    hu[:, 6:9, :] = 1000.0
    hu[8:14, 12:18, 10:16] = -1000.0
    g.HU = hu
    plan = IMPTPlan()
    plan.n_fractions = 2
    c = np.arange(-10.0, 10.1, 4.0)
    for ga, couch in ((120.0, 30.0), (240.0, 0.0)):
        sl = grid_spots([20, 28, 34], c, c)
        sl[:, 2] = rng.uniform(0.2, 1.0, size=sl.shape[0]).astype(np.float32)
        plan.addBeam(_beam(ga, sl, couch=couch))
    out.append((g, plan))
    # 1 mm water + slab, one beam at gantry 90
    shape = (30, 36, 34)
    g = _grid(shape, 1.0)
    hu = np.zeros(shape, np.float32)
    hu[:, :, 8:12] = 900.0
    g.HU = hu
    plan = IMPTPlan()
    c = np.arange(-4.0, 4.1, 2.0)
    sl = grid_spots([2, 5], c, c)
    plan.addBeam(_beam(90.0, sl))
    out.append((g, plan))
    return out


@pytest.mark.parametrize("wet_mode,cutoff", [("reach", 6.0), ("full", 0.0)])
def test_batched_small_plans_vs_full_oracle(dev, machine_dir, wet_mode, cutoff):
    from DoseCUDA.batch import BatchedDose
    from DoseCUDA.engine import IMPTDoseEngine
    plans = _small_plans()
    _say("small batch (%d plans) %s cutoff %g" % (len(plans), wet_mode, cutoff))
    engines = [IMPTDoseEngine(g, p, gpu_id=0, cutoff_sigma=cutoff) for g, p in plans]
    _raw_buffers(engines)
    bd = BatchedDose(engines, wet_mode=wet_mode)
    doses = [_np(d).copy() for d in bd.forward()]
    torch.cuda.synchronize()
    rng = np.random.default_rng(17)
    for (g, plan), e, d in zip(plans, engines, doses):
        ref, beams, wets = O.compute_plan(g, plan, machine_dir, return_wet=True)
        s = float(np.float32(g.spacing[0]))
        m = O.compare(d, ref)
        gam = O.gamma_pass_rate(d, ref, s)
        # WET where the pass computed it (everywhere in full mode)
        dw = 0.0
        rlsp = O.rlsp_from_hu(g.HU, machine_dir)
        for k, b in enumerate(e.beams):
            wg = _np(b.wet)
            done = wg < 1e29
            assert done.all() or wet_mode == "reach"
            dw = max(dw, float(np.abs(wg[done] - wets[k][done]).max()))
            if wet_mode == "full":
                raw_o = O.raymarch(rlsp, s, O.beam_geometry(np.asarray(b.beam.iso, np.float32),
                                                            np.asarray(g.origin, np.float32), b.beam.gantry_angle,
                                                            b.beam.couch_angle, b.model.VSADX, b.model.VSADY))
                dw = max(dw, float(np.abs(_np(b.wet_raw) - raw_o).max()))
        _say("  grid %s @ %g mm, %d beams: max|dD|/Dmax %.3e gamma %.4f max|dWET| %.2e cm"
             % (g.HU.shape, s, len(plan.beam_list), m["max_rel_to_dmax"], gam, dw))
        assert m["max_rel_to_dmax"] <= DOSE_TOL and gam >= GAMMA_MIN and dw <= WET_TOL
    # gradients of L = sum_v g_v D_v with a dense random dL/dD on every voxel with dose
    gvols = [(rng.normal(size=d.shape) * (d > 0)).astype(np.float32) for d in doses]
    res = bd.adjoint([torch.from_numpy(gv).to(dev) for gv in gvols])
    torch.cuda.synchronize()
    k = 0
    for (g, plan), e, gv in zip(plans, engines, gvols):
        ref, _, wets = O.compute_plan(g, plan, machine_dir, return_wet=True)
        s = float(np.float32(g.spacing[0]))
        for j, beam in enumerate(plan.beam_list):
            model = O.beam_model_for(plan, beam)
            geom = O.beam_geometry(np.asarray(beam.iso, np.float32), np.asarray(g.origin, np.float32),
                                   beam.gantry_angle, beam.couch_angle, model.VSADX, model.VSADY)
            order, spots, layers, r80 = O.spots_and_layers(beam.spot_list, model.divergence_params)
            gm_s, gxy_s = O.pb_grad(wets[j], gv.astype(np.float64) * plan.n_fractions, s, geom, spots, layers, r80,
                                    O.ModelArrays(model))
            gm_w = np.zeros(order.size)
            gxy_w = np.zeros((order.size, 2))
            gm_w[order], gxy_w[order] = gm_s, gxy_s
            gm, gxy = (_np(t).astype(np.float64) for t in res[k])
            em, ex = rel_err(gm, gm_w), rel_err(gxy, gxy_w)
            _say("  grid %s beam %d: grad rel err mu %.2e xy %.2e" % (g.HU.shape, j, em, ex))
            assert em <= GRAD_TOL and ex <= GRAD_TOL
            k += 1
    del bd, engines
    _free()


# ---------------------------------------------------------------------------
# head phantom (160 x 200 x 200 @ 2 mm, 3 beams, 9720 spots)
@pytest.fixture(scope="module")
def head(dev, machine_dir):
    from DoseCUDA.batch import BatchedDose
    from DoseCUDA.engine import IMPTDoseEngine
    from DoseCUDA.workloads import head_plan
    _say("head phantom: building plan and reach/cutoff-6 batch")
    g, plan = head_plan(seed=2)
    eng = IMPTDoseEngine(g, plan, gpu_id=0, cutoff_sigma=6.0)
    bd = BatchedDose([eng], wet_mode="reach")
    dose = _np(bd.forward()[0]).copy()
    torch.cuda.synchronize()
    idx = sample_voxels(dose, 2000, 2000, seed=22)
    _say("head phantom: oracle at %d voxels" % idx.size)
    op = OraclePlan(g, plan, machine_dir, idx)
    ref = op.dose()
    yield {"g": g, "plan": plan, "eng": eng, "bd": bd, "dose": dose, "idx": idx, "op": op, "ref": ref}
    _free()


def _report_dose(case, got_vol, idx, ref, spacing):
    got = got_vol.reshape(-1)[idx]
    dmax = float(ref.max())
    err = float(np.abs(got - ref).max() / dmax)
    gam = gamma_at(ref, idx, got_vol, spacing, dmax=dmax)
    _say("%s: max|dD|/Dmax %.3e (D_max %.4g, %d voxels) gamma 2%%/2mm %.4f" % (case, err, dmax, idx.size, gam))
    return err, gam


def test_head_phantom_reach_cutoff6(head, dev):
    h = head
    err, gam = _report_dose("head reach cutoff 6", h["dose"], h["idx"], h["ref"], 2.0)
    dw = _check_wet("head", h["eng"], h["op"], h["idx"], float(h["ref"].max()), reach=True)
    _say("head reach: max|dWET| %.2e cm at computed voxels" % dw)
    rng = np.random.default_rng(5)
    gv = rng.normal(size=h["idx"].size)
    got = _gpu_grads(h["bd"], [h["dose"].shape], [h["idx"]], [gv], dev)
    want = h["op"].grad(gv)
    em, ex = _grad_errs(got, want)
    _say("head reach: grad rel err mu %.2e xy %.2e" % (em, ex))
    assert err <= DOSE_TOL and gam >= GAMMA_MIN and dw <= WET_TOL
    assert em <= GRAD_TOL and ex <= GRAD_TOL


def test_head_phantom_full_dense(head, dev):
    from DoseCUDA.batch import BatchedDose
    from DoseCUDA.engine import IMPTDoseEngine
    h = head
    _say("head phantom: full WET, cutoff 0 (dense) batch")
    eng = IMPTDoseEngine(h["g"], h["plan"], gpu_id=0, cutoff_sigma=0.0)
    _raw_buffers([eng])
    bd = BatchedDose([eng], wet_mode="full")
    dose = _np(bd.forward()[0]).copy()
    torch.cuda.synchronize()
    err, gam = _report_dose("head full cutoff 0", dose, h["idx"], h["ref"], 2.0)
    dw = _check_wet("head full", eng, h["op"], h["idx"], float(h["ref"].max()), reach=False)
    _say("head full: max|dWET| (raw and smoothed) %.2e cm" % dw)
    gv = np.random.default_rng(6).normal(size=h["idx"].size)
    got = _gpu_grads(bd, [dose.shape], [h["idx"]], [gv], dev)
    em, ex = _grad_errs(got, h["op"].grad(gv))
    _say("head full: grad rel err mu %.2e xy %.2e" % (em, ex))
    assert err <= DOSE_TOL and gam >= GAMMA_MIN and dw <= WET_TOL
    assert em <= GRAD_TOL and ex <= GRAD_TOL
    del bd, eng
    _free()


def test_head_phantom_dropin_default(head):
    """IMPTDoseGrid.computeIMPTPlan with host arrays, at its default (every spot
    for every in-range voxel, as the reference)."""
    h = head
    _say("head phantom: drop-in computeIMPTPlan")
    g = h["g"]
    g.computeIMPTPlan(h["plan"])
    err, gam = _report_dose("head drop-in", np.asarray(g.dose), h["idx"], h["ref"], 2.0)
    assert err <= DOSE_TOL and gam >= GAMMA_MIN
    _free()


# ---------------------------------------------------------------------------
# 64 uneven plans in one batch
def test_uneven_batch_vs_oracle(dev, machine_dir):
    from DoseCUDA.batch import BatchedDose
    from DoseCUDA.engine import IMPTDoseEngine
    from DoseCUDA.workloads import uneven_plans
    _say("uneven batch: building 64 plans")
    plans = uneven_plans(64, seed=4)
    engines = [IMPTDoseEngine(g, p, gpu_id=0, cutoff_sigma=6.0) for g, p in plans]
    bd = BatchedDose(engines, wet_mode="reach")
    _say("uneven batch: forward (%d beams)" % bd.n_beams)
    doses = [_np(d).copy() for d in bd.forward()]
    torch.cuda.synchronize()
    idxs, ops, gvs, errs, gams = [], [], [], [], []
    rng = np.random.default_rng(44)
    for p, ((g, plan), d) in enumerate(zip(plans, doses)):
        idx = sample_voxels(d, 60, 60, seed=100 + p)
        op = OraclePlan(g, plan, machine_dir, idx)
        ref = op.dose()
        got = d.reshape(-1)[idx]
        dmax = float(ref.max())
        errs.append(float(np.abs(got - ref).max() / dmax) if dmax > 0 else float(np.abs(got).max()))
        gams.append(gamma_at(ref, idx, d, 2.0, dmax=dmax) if dmax > 0 else 1.0)
        idxs.append(idx)
        ops.append(op)
        gvs.append(rng.normal(size=idx.size))
    worst = int(np.argmax(errs))
    _say("uneven batch: max|dD|/Dmax worst %.3e (plan %d, side %d), median %.3e; gamma min %.4f"
         % (errs[worst], worst, plans[worst][0].HU.shape[0], float(np.median(errs)), min(gams)))
    got = _gpu_grads(bd, [d.shape for d in doses], idxs, gvs, dev)
    want = []
    for op, gv in zip(ops, gvs):
        want += op.grad(gv)
    em, ex = _grad_errs(got, want)
    _say("uneven batch: grad rel err (worst beam) mu %.2e xy %.2e over %d beams" % (em, ex, len(want)))
    assert max(errs) <= DOSE_TOL and min(gams) >= GAMMA_MIN
    assert em <= GRAD_TOL and ex <= GRAD_TOL
    del bd, engines
    _free()


# ---------------------------------------------------------------------------
# body phantom optimisation loop
def test_body_loop_vs_oracle(dev, machine_dir):
    from DoseCUDA.batch import BatchedDose
    from DoseCUDA.engine import IMPTDoseEngine
    from DoseCUDA.workloads import body_plan
    iters = 3
    _say("body phantom: building")
    g, plan, target = body_plan()
    eng = IMPTDoseEngine(g, plan, gpu_id=0, cutoff_sigma=6.0)
    tmask = torch.from_numpy(target).to(dev)
    body = torch.from_numpy(g.HU > -500.0).to(dev)
    n_t, n_b = float(tmask.sum().item()), float(body.sum().item())
    bd = BatchedDose([eng], wet_mode="reach")
    mus = [b.mu0.clone() for b in eng.beams]
    lr = None
    losses = []
    _say("body phantom: %d projected-gradient iterations" % iters)
    for it in range(iters):
        d = bd.forward(mus, None)[0]
        losses.append(float(((d[tmask] - 2.0) ** 2).mean() + 0.1 * (d[body] ** 2).mean()))
        gd = torch.where(tmask, d - 2.0, torch.zeros_like(d)) * (2.0 / n_t) \
            + torch.where(body, d * (0.2 / n_b), torch.zeros_like(d))
        grads = bd.adjoint([gd])
        if lr is None:
            lr = 0.05 * max(float(m.max()) for m in mus) / max(float(gm.abs().max()) for gm, _ in grads)
        for m, (gm, _) in zip(mus, grads):
            m.sub_(lr * gm).clamp_(min=0.0)
    d = _np(bd.forward(mus, None)[0]).copy()
    torch.cuda.synchronize()
    losses.append(float(((torch.from_numpy(d).to(dev)[tmask] - 2.0) ** 2).mean()
                        + 0.1 * (torch.from_numpy(d).to(dev)[body] ** 2).mean()))
    _say("body phantom: loss %s" % ", ".join("%.4g" % v for v in losses))
    assert losses[-1] < losses[0]
    idx = sample_voxels(d, 2000, 2000, seed=55)
    _say("body phantom: oracle at %d voxels" % idx.size)
    op = OraclePlan(g, plan, machine_dir, idx)
    mus_np = [_np(m).astype(np.float64) for m in mus]
    ref = op.dose(mus_np)
    err, gam = _report_dose("body after %d iterations" % iters, d, idx, ref, 1.0)
    # the reach pass ran for the final forward: compare its WET where computed
    dw = _check_wet_with(op, eng, idx, ref, mus_np)
    _say("body: max|dWET| %.2e cm at computed voxels" % dw)
    gv = np.random.default_rng(9).normal(size=idx.size)
    got = _gpu_grads(bd, [d.shape], [idx], [gv], dev)
    em, ex = _grad_errs(got, op.grad(gv, mus_np))
    _say("body: grad rel err mu %.2e xy %.2e" % (em, ex))
    assert err <= DOSE_TOL and gam >= GAMMA_MIN and dw <= WET_TOL
    assert em <= GRAD_TOL and ex <= GRAD_TOL
    del bd, eng
    _free()


def _check_wet_with(op, eng, idx, ref, mus_np):
    worst = 0.0
    dmax = float(ref.max())
    for k, b in enumerate(eng.beams):
        w = _np(b.wet).reshape(-1)[idx]
        done = w < 1e29
        if (~done).any():
            bd_ref = op.beam_dose(k, mus_np[k])[~done] * op.n_fx
            assert bd_ref.max() < 1e-6 * dmax, (k, float(bd_ref.max() / dmax))
        if done.any():
            worst = max(worst, float(np.abs(w[done] - op.wet[k][done]).max()))
    return worst
