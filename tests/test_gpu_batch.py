"""Batched plans (dh_batch_*) vs the single-beam path, and HIP-graph replay (GPU).

The batched kernels run the same per-item device code as the single-beam ones
and add beam doses in the same order, so WET, dose and gradients must agree
BIT FOR BIT with the single-beam path, which the oracle tests pin
(test_gpu_parity.py).  Other adjoint schedules (chunk size, slab count) change
only the order of the partial sums: <= 1e-5 relative.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-batch] " + msg, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


def _phantom(shape, spacing, seed):
    from DoseCUDA.workloads import _grid
    g = _grid(shape, spacing)
    rng = np.random.default_rng(seed)
    hu = rng.uniform(-200.0, 300.0, size=shape).astype(np.float32)
    hu[:, shape[1] // 5: shape[1] // 5 + 4, :] = 1200.0
    hu[3:6, :, 4:7] = -1000.0
    g.HU = hu
    return g


def _plan(gantries, couches, layers, radius=12.0, pitch=3.0, seed=5, n_fractions=1):
    from DoseCUDA import IMPTPlan
    from DoseCUDA.workloads import _beam, grid_spots
    rng = np.random.default_rng(seed)
    plan = IMPTPlan()
    plan.n_fractions = n_fractions
    c = np.arange(-radius, radius + 1e-6, pitch)
    for ga, ca in zip(gantries, couches):
        sl = grid_spots(layers, c, c, disk_radius=radius)
        sl[:, 2] = rng.uniform(0.2, 1.0, size=sl.shape[0]).astype(np.float32)
        plan.addBeam(_beam(ga, sl, couch=ca))
    return plan


def _engines(dev, cutoff=6.0):
    """Four plans of different grids: one beam (water slab), two oblique beams
    with n_fractions = 2, three beams on a small grid, and a plan without beams."""
    from DoseCUDA import IMPTPlan
    from DoseCUDA.engine import IMPTDoseEngine
    from DoseCUDA.workloads import water_slab_plan
    out = []
    g0, p0 = water_slab_plan(n=40, spacing=2.5)
    out.append(IMPTDoseEngine(g0, p0, cutoff_sigma=cutoff))
    g1 = _phantom((40, 56, 48), 2.5, 3)
    out.append(IMPTDoseEngine(g1, _plan([120.0, 240.0], [30.0, 0.0], [22, 27, 31], n_fractions=2),
                              cutoff_sigma=cutoff))
    g2 = _phantom((24, 28, 20), 2.0, 7)
    out.append(IMPTDoseEngine(g2, _plan([0.0, 90.0, 200.0], [0.0, 0.0, 15.0], [20, 24], radius=6.0, seed=9),
                              cutoff_sigma=cutoff))
    g3 = _phantom((16, 20, 24), 2.0, 1)
    out.append(IMPTDoseEngine(g3, IMPTPlan(), cutoff_sigma=cutoff))
    return out


def _grads(engines, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for e in engines:
        g = rng.normal(size=e.shape).astype(np.float32)
        g[rng.uniform(size=e.shape) < 0.3] = 0.0
        out.append(torch.from_numpy(g).to(e.device))
    return out


def _single_results(engines, grads):
    wets, doses, adj = [], [], []
    for e, g in zip(engines, grads):
        e.compute_wet(force=True)
        wets.append([b.wet.clone() for b in e.beams])
        doses.append(e.dose().clone())
        adj.append([(m.clone(), x.clone()) for m, x in e.adjoint(g)])
    torch.cuda.synchronize()
    return wets, doses, adj


@pytest.mark.parametrize("cutoff", [6.0, 0.0])
def test_batch_matches_single_path(dev, cutoff):
    from DoseCUDA.batch import BatchedDose
    _say("batch vs single path, cutoff %g" % cutoff)
    engines = _engines(dev, cutoff)
    grads = _grads(engines)
    wets, doses, adj = _single_results(engines, grads)
    bd = BatchedDose(engines)
    _say("items %s" % bd.item_counts())
    bd.compute_wet(force=True)
    got = bd.forward()
    gadj = bd.adjoint(grads)
    torch.cuda.synchronize()
    k = 0
    for p, e in enumerate(engines):
        for j, b in enumerate(e.beams):
            assert torch.equal(b.wet, wets[p][j]), "WET plan %d beam %d" % (p, j)
            assert torch.equal(gadj[k][0], adj[p][j][0]), "dL/dmu plan %d beam %d" % (p, j)
            assert torch.equal(gadj[k][1], adj[p][j][1]), "dL/dxy plan %d beam %d" % (p, j)
            k += 1
        assert torch.equal(got[p], doses[p]), "dose plan %d" % p
    assert float(got[3].abs().max()) == 0.0      # plan without beams
    assert float(got[1].max()) > 0.0


@pytest.mark.parametrize("wet_mode", ["full", "reach"])
def test_batch_smoothing_kernels(dev, wet_mode):
    """Both batched smoothing kernels equal the single-beam smoothing bit for
    bit: the tiled (LDS) kernel on 2-2.5 mm grids, the gather kernel for a
    batch with a 1 mm grid and an oblique beam (halo box above 64 KB)."""
    from DoseCUDA.batch import BatchedDose
    from DoseCUDA.engine import IMPTDoseEngine
    _say("smoothing kernels, %s WET" % wet_mode)
    g = _phantom((20, 40, 36), 1.0, 11)
    wide = [IMPTDoseEngine(g, _plan([120.0, 0.0], [30.0, 0.0], [20, 24], radius=6.0, seed=3), cutoff_sigma=6.0)]
    for engines, tiled in ((wide, False), (_engines(dev)[:3], True)):
        want = []
        for e in engines:
            e.compute_wet(force=True)
            want.append([b.wet.clone() for b in e.beams])
        for e in engines:
            for b in e.beams:
                b.wet.fill_(float("nan"))
        bd = BatchedDose(engines, wet_mode=wet_mode)
        assert bd.smooth_tiled == tiled
        if wet_mode == "full":
            bd.compute_wet(force=True)
        else:
            bd.forward()
        torch.cuda.synchronize()
        for e, ws in zip(engines, want):
            for b, w in zip(e.beams, ws):
                keep = ~(b.wet >= 1e29) if wet_mode == "reach" else torch.ones_like(w, dtype=torch.bool)
                assert torch.equal(b.wet[keep], w[keep]), "tiled=%s" % tiled
                assert int(keep.sum()) > 0


def test_batch_beam_doses(dev):
    from DoseCUDA.batch import BatchedDose
    _say("per-beam dose volumes")
    engines = _engines(dev)[:3]
    bd = BatchedDose(engines, keep_beam_doses=True)
    doses = bd.forward()
    torch.cuda.synchronize()
    k = 0
    for p, e in enumerate(engines):
        acc = None
        for j in range(len(e.beams)):
            single = e.beam_dose(j)
            assert torch.equal(bd.beam_doses[k], single), "beam dose plan %d beam %d" % (p, j)
            acc = single.clone() if acc is None else acc + single
            k += 1
        assert torch.equal(doses[p], acc * float(e.n_fractions))


def test_batch_graph_replay(dev):
    """One HIP graph: forward -> element-wise loss gradient -> adjoint.  Replays
    with new weights between them (and eager work in between) match eager
    batched steps bit for bit."""
    from DoseCUDA.batch import BatchedDose, GraphedStep
    _say("graph capture / replay")
    engines = _engines(dev)[:3]
    bd = BatchedDose(engines)
    bd.compute_wet()
    mus = [b.mu0.clone() for e in engines for b in e.beams]
    xys = [b.xy0.clone() for e in engines for b in e.beams]

    def step():
        doses = bd.forward(mus, xys)
        for d, gd in zip(doses, bd.grad_doses):
            torch.mul(d - 1.0, 2.0 / d.numel(), out=gd)
        return doses, bd.adjoint(None)

    gs = GraphedStep(step, dev)
    rng = np.random.default_rng(11)
    for rep in range(3):
        for m, x in zip(mus, xys):
            m.copy_(torch.from_numpy(rng.uniform(0.1, 1.0, size=m.shape[0]).astype(np.float32)))
            x.add_(torch.from_numpy(rng.normal(scale=0.3, size=tuple(x.shape)).astype(np.float32)).to(dev))
        doses_g, adj_g = gs.replay()
        doses_g = [d.clone() for d in doses_g]
        adj_g = [(a.clone(), b.clone()) for a, b in adj_g]
        torch.cuda.synchronize()
        _ = torch.randn(1000, device=dev).sum().item()   # eager work between replays
        doses_e, adj_e = step()
        torch.cuda.synchronize()
        for a, b in zip(doses_g, doses_e):
            assert torch.equal(a, b), "replay %d dose" % rep
        for (a0, a1), (b0, b1) in zip(adj_g, adj_e):
            assert torch.equal(a0, b0) and torch.equal(a1, b1), "replay %d gradients" % rep
        _say("replay %d ok" % rep)


@pytest.mark.parametrize("tune,cutoff", [((32, 0, 0, 0), 6.0), ((64, 0, 0, 0), 6.0), ((0, 512, 16, 8), 6.0),
                                         ((32, 200000, 8, 4000), 6.0), ((0, 0, 0, 0, 7, 3), 6.0),
                                         ((128, 0, 0, 0), 6.0), ((128, 512, 16, 8), 6.0), ((128, 0, 0, 0), 0.0),
                                         ((64, 0, 0, 0), 0.0)])
def test_batch_adjoint_schedules(dev, tune, cutoff):
    from DoseCUDA.batch import BatchedDose, make_tune
    _say("adjoint schedule %s cutoff %g" % (tune, cutoff))
    engines = _engines(dev, cutoff=cutoff)[:3]
    grads = _grads(engines, seed=3)
    ref = BatchedDose(engines)
    ref.compute_wet()
    want = [(a.clone(), b.clone()) for a, b in ref.adjoint(grads)]
    bd = BatchedDose(engines, tune=make_tune(*tune))
    bd.compute_wet()
    got = bd.adjoint(grads)
    torch.cuda.synchronize()
    for (a0, a1), (b0, b1) in zip(got, want):
        if b0.numel() == 0:
            continue
        assert float((a0 - b0).abs().max()) <= 1e-5 * float(b0.abs().max())
        assert float((a1 - b1).abs().max()) <= 1e-5 * float(b1.abs().max())


@pytest.mark.parametrize("cutoff", [6.0, 0.0])
def test_batch_reach_limited_wet(dev, cutoff):
    """WET only where the current spots reach: dose and gradients equal the
    full-WET batch bit for bit, also after spot positions move (graph replays
    recompute the reach-limited WET), and the skipped bricks hold the marker."""
    from DoseCUDA.batch import BatchedDose, GraphedStep
    _say("reach-limited WET, cutoff %g" % cutoff)
    engines = _engines(dev, cutoff)[:3]
    full = BatchedDose(engines)
    full.compute_wet()
    mus = [b.mu0.clone() for e in engines for b in e.beams]
    xys = [b.xy0.clone() for e in engines for b in e.beams]
    grads = _grads(engines, seed=5)

    def run_full():
        d = [x.clone() for x in full.forward(mus, xys)]
        a = [(m.clone(), x.clone()) for m, x in full.adjoint(grads)]
        return d, a

    # separate engine objects so the two batches keep separate WET volumes
    engines_r = _engines(dev, cutoff)[:3]
    reach = BatchedDose(engines_r, wet_mode="reach")
    reach.poison_scratch()     # raw WET the reach pass skips must never be read
    for e in engines_r:
        for b in e.beams:
            b.wet.fill_(float("nan"))

    def step_reach():
        d = reach.forward(mus, xys)
        for gd, g in zip(reach.grad_doses, grads):
            gd.copy_(g)
        return d, reach.adjoint(None)

    gs = GraphedStep(step_reach, dev)
    rng = np.random.default_rng(21)
    for rep in range(3):
        if rep:
            for x in xys:
                x.add_(torch.from_numpy(rng.normal(scale=4.0, size=tuple(x.shape)).astype(np.float32)).to(dev))
        want_d, want_a = run_full()
        got_d, got_a = gs.replay()
        torch.cuda.synchronize()
        for a, b in zip(got_d, want_d):
            assert torch.equal(a, b), "replay %d dose" % rep
        for (a0, a1), (b0, b1) in zip(got_a, want_a):
            assert torch.equal(a0, b0) and torch.equal(a1, b1), "replay %d gradients" % rep
        # every voxel holds the marker or exactly the full WET (no raw WET read unwritten)
        for e_f, e_r in zip(engines, engines_r):
            for bf, br in zip(e_f.beams, e_r.beams):
                keep = ~(br.wet >= 1e29)      # NaN (poisoned scratch read) stays in the comparison
                assert torch.equal(br.wet[keep], bf.wet[keep]), "replay %d: reach WET != full WET" % rep
        skipped = sum(int((b.wet >= 1e29).sum()) for e in engines_r for b in e.beams)
        total = sum(b.wet.numel() for e in engines_r for b in e.beams)
        _say("replay %d ok, WET skipped in %d of %d voxels" % (rep, skipped, total))
        if cutoff == 0.0:
            assert skipped == 0


def test_batch_persistent_brick_schedules(dev):
    """Persistent forward / ray-march grids and the ray-march patch shapes
    compute every brick once: WET, dose and gradients equal the default
    schedule (one wave per brick, 16x4 patches) bit for bit."""
    from DoseCUDA.batch import BatchedDose, make_tune
    _say("persistent brick schedules")
    grads = _grads(_engines(dev)[:3], seed=8)
    out = []
    for tune in (make_tune(), make_tune(fwd_blocks=5, ray_blocks=3), make_tune(fwd_blocks=4096, ray_blocks=4096),
                 make_tune(ray_shape=1), make_tune(ray_shape=2), make_tune(ray_shape=3), make_tune(ray_shape=4),
                 make_tune(ray_shape=5), make_tune(ray_blocks=3, ray_shape=1)):
        engines = _engines(dev)[:3]
        bd = BatchedDose(engines, tune=tune, wet_mode="reach")
        d = [x.clone() for x in bd.forward()]
        a = [(m.clone(), x.clone()) for m, x in bd.adjoint(grads)]
        w = [b.wet.clone() for e in engines for b in e.beams]
        out.append((d, a, w))
    torch.cuda.synchronize()
    for d, a, w in out[1:]:
        for x, y in zip(d, out[0][0]):
            assert torch.equal(x, y)
        for (x0, x1), (y0, y1) in zip(a, out[0][1]):
            assert torch.equal(x0, y0) and torch.equal(x1, y1)
        for x, y in zip(w, out[0][2]):
            assert torch.equal(x, y)


def test_autotuner_picks_and_caches(dev, tmp_path, monkeypatch):
    """The autotuner returns a valid schedule, stores it, and the next batch of
    the same shape takes it from the cache without tuning; results with the
    tuned schedule match the default one (bit for bit where only brick
    schedules change, 1e-5 for the adjoint)."""
    from DoseCUDA import autotune as A
    from DoseCUDA.batch import BatchedDose, tune_dict
    _say("autotuner")
    monkeypatch.setenv("DOSEHIP_TUNE_CACHE", str(tmp_path / "tune.json"))
    engines = _engines(dev)[:3]
    bd = BatchedDose(engines, tune="auto", wet_mode="reach")
    t = tune_dict(bd.tune)
    _say("tuned %s" % t)
    assert t["adj_cs"] in (32, 64, 128) and t["fwd_blocks"] in A.FWD_BLOCKS and t["ray_blocks"] in A.RAY_BLOCKS
    assert t["ray_shape"] in A.RAY_SHAPES
    assert A.lookup(A.signature(engines, 6.0, "reach")) == t

    def boom(*a, **k):
        raise AssertionError("cache miss")
    monkeypatch.setattr(A, "tune", boom)
    bd2 = BatchedDose(_engines(dev)[:3], tune="auto", wet_mode="reach")
    assert tune_dict(bd2.tune) == t
    grads = _grads(engines, seed=4)
    ref = BatchedDose(_engines(dev)[:3], wet_mode="reach")
    d0 = [x.clone() for x in ref.forward()]
    a0 = [(m.clone(), x.clone()) for m, x in ref.adjoint(grads)]
    d1 = bd.forward()
    a1 = bd.adjoint(grads)
    torch.cuda.synchronize()
    for x, y in zip(d1, d0):
        assert torch.equal(x, y)
    for (x0, x1), (y0, y1) in zip(a1, a0):
        if y0.numel():
            assert float((x0 - y0).abs().max()) <= 1e-5 * float(y0.abs().max())
            assert float((x1 - y1).abs().max()) <= 1e-5 * float(y1.abs().max())


def test_autotuner_leaves_engine_wet_cache_valid(dev, tmp_path, monkeypatch):
    """Tuning a reach-mode batch writes reach-limited WET (1e30 outside the
    spots' reach) into the engines' own WET volumes; the engines must not
    reuse it as their full WET afterwards.  compute_wet, tune, then engine
    doses with spots moved 8 mm must equal a fresh engine's bit for bit."""
    from DoseCUDA.batch import BatchedDose
    _say("autotuner vs the engines' WET cache")
    monkeypatch.setenv("DOSEHIP_TUNE_CACHE", str(tmp_path / "tune.json"))
    engines = _engines(dev)[:3]
    for e in engines:
        e.compute_wet()
    BatchedDose(engines, tune="auto", wet_mode="reach")
    fresh = _engines(dev)[:3]
    for e, f in zip(engines, fresh):
        xys = [b.xy0 + 8.0 for b in e.beams]
        got = e.dose(None, xys).clone()
        want = f.dose(None, [b.xy0 + 8.0 for b in f.beams])
        torch.cuda.synchronize()
        assert torch.equal(got, want)
        for b, c in zip(e.beams, f.beams):
            assert torch.equal(b.wet, c.wet)


def test_reuse_prep_requires_prepared_records(dev):
    """DH_BATCH_REUSE_PREP on a batch whose spot records no call has prepared is
    an argument error (raised before any kernel runs); after a forward it works."""
    from DoseCUDA.batch import BatchedDose
    _say("reuse-prep guard")
    engines = _engines(dev)[:2]
    bd = BatchedDose(engines, wet_mode="full")
    bd.compute_wet()                       # full WET prepares no spot records
    assert bd.batch.prepared == 0
    with pytest.raises(ValueError, match="REUSE_PREP"):
        bd.adjoint(_grads(engines), reuse_prep=True)
    bd.forward()
    assert bd.batch.prepared == 1
    a = bd.adjoint(_grads(engines), reuse_prep=True)
    torch.cuda.synchronize()
    assert all(torch.isfinite(g).all() for g, _ in a)


@pytest.mark.parametrize("cutoff", [6.0, 0.0])
def test_batch_workspace_covers_both_schedules(dev, cutoff):
    """The workspace size query does not know the cutoff, which picks the
    adjoint chunking (packed 128-spot chunks when culling, 32 / 64 otherwise)
    and with it the number of partial records.  On a 160^3 grid with 40 small
    layers the two layouts differ; set-up must accept the queried size for
    either cutoff, and the gradients must match the single-beam path."""
    from DoseCUDA.batch import BatchedDose
    from DoseCUDA.engine import IMPTDoseEngine
    _say("workspace size vs schedule, cutoff %g" % cutoff)
    g = _phantom((160, 160, 160), 2.0, 12)
    eng = IMPTDoseEngine(g, _plan([0.0], [0.0], list(range(20, 60)), radius=4.0, pitch=3.0), cutoff_sigma=cutoff)
    grads = _grads([eng], seed=8)
    eng.compute_wet(force=True)
    want = [(m.clone(), x.clone()) for m, x in eng.adjoint(grads[0])]
    bd = BatchedDose([eng])
    _say("items %s" % bd.item_counts())
    bd.compute_wet(force=True)
    got = bd.adjoint(grads)
    torch.cuda.synchronize()
    assert float(want[0][0].abs().max()) > 0.0
    for (a0, a1), (b0, b1) in zip(got, want):
        assert torch.equal(a0, b0) and torch.equal(a1, b1)


@pytest.mark.parametrize("wet_mode,n_streams", [("reach", 2), ("reach", 3), ("full", 2)])
def test_streamed_batches(dev, wet_mode, n_streams):
    """Sub-batches on concurrent streams (StreamedBatches) vs one BatchedDose
    over the same plans: doses and gradients bit-identical, eager (forward,
    adjoint, step) and in HIP-graph replays with moving spots, with eager work
    between the replays."""
    from DoseCUDA.batch import BatchedDose, GraphedStep, StreamedBatches
    _say("streamed batches: %s WET, %d streams" % (wet_mode, n_streams))
    engines = _engines(dev)
    grads = _grads(engines, seed=3)
    mus = [b.mu0.clone() for e in engines for b in e.beams]
    xys = [b.xy0.clone() for e in engines for b in e.beams]
    one = BatchedDose(engines, wet_mode=wet_mode)
    sb = StreamedBatches(engines, n_streams=n_streams, wet_mode=wet_mode)
    assert len(sb.subs) == min(n_streams, len(engines))
    assert [i for g in sb.groups for i in g] == list(range(len(engines)))

    def ref_step():
        doses = [d.clone() for d in one.forward(mus, xys)]
        adj = [(a.clone(), b.clone()) for a, b in one.adjoint(grads, reuse_prep=wet_mode == "reach")]
        return doses, adj

    def check(doses, adj, want, what):
        assert len(doses) == len(want[0]) and len(adj) == len(want[1])
        for p, (a, b) in enumerate(zip(doses, want[0])):
            assert torch.equal(a, b), "%s: plan %d dose" % (what, p)
        for k, ((a0, a1), (b0, b1)) in enumerate(zip(adj, want[1])):
            assert torch.equal(a0, b0) and torch.equal(a1, b1), "%s: beam %d gradients" % (what, k)

    want = ref_step()
    doses = [d.clone() for d in sb.forward(mus, xys)]
    adj = sb.adjoint(grads, reuse_prep=wet_mode == "reach")
    torch.cuda.synchronize()
    check(doses, adj, want, "forward + adjoint")

    gsrc = [g.clone() for g in grads]

    def grad_fn(p, dose, out):
        out.copy_(gsrc[p])

    d2, a2 = sb.step(mus, xys, grad_fn=grad_fn)
    torch.cuda.synchronize()
    check(d2, a2, want, "step")

    gs = GraphedStep(lambda: sb.step(mus, xys, grad_fn=grad_fn, input_order=False), dev)
    rng = np.random.default_rng(17)
    for rep in range(3):
        for m, x in zip(mus, xys):
            m.copy_(torch.from_numpy(rng.uniform(0.1, 1.0, size=m.shape[0]).astype(np.float32)))
            x.add_(torch.from_numpy(rng.normal(scale=0.3, size=tuple(x.shape)).astype(np.float32)).to(dev))
        doses_g, adj_g = gs.replay()
        doses_g = [d.clone() for d in doses_g]
        adj_g = sb._input_order()
        adj_g = [(a.clone(), b.clone()) for a, b in adj_g]
        torch.cuda.synchronize()
        _ = torch.randn(1000, device=dev).sum().item()   # eager work between replays
        check(doses_g, adj_g, ref_step(), "replay %d" % rep)
        _say("replay %d ok" % rep)
