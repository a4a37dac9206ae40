"""Edge cases of the HIP path vs the FP64 oracle (GPU).

* degenerate grids (1 or 2 voxels along an axis, 1x1x1): WET within 2e-4 cm
  of the oracle, dose within 1e-3 of D_max (or 1e-6 absolute when there is
  hardly any dose), with and without culling;
* a beam whose energy ids are all outside the model (no usable layer) and a
  beam with a negative id (which blocks every layer, as importLayers does):
  zero dose and zero gradients, like the oracle;
* spots far outside the volume: dose equals the oracle's (about zero) and
  the gradients are finite;
* beams on the range-shifter model (negative sigma_ms, an energy whose R80 is
  negative): dose within 1e-3 of D_max of the oracle;
* a batch whose plans include a beam without spots, and a batch of the
  degenerate grids: WET, dose and gradients bit-identical to the single-beam
  path.
"""

import numpy as np
import pytest
import torch

from oracle import impt_oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-edges] " + msg, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


def _grid_plan(shape, spacing, spots, gantry=30.0, couch=20.0, seed=3):
    from DoseCUDA import IMPTPlan
    from DoseCUDA.workloads import _beam, _grid
    g = _grid(shape, spacing)
    g.HU = np.random.default_rng(seed).uniform(-300.0, 400.0, size=shape).astype(np.float32)
    plan = IMPTPlan()
    plan.addBeam(_beam(gantry, np.asarray(spots, np.float32), couch=couch))
    return g, plan


def _spots(ids, coords=(-3.0, 0.0, 3.0), mu=0.7):
    return [(x, y, mu, e) for e in ids for y in coords for x in coords]


def _oracle_dose_wet(g, plan, machine_dir):
    total, _, wets = O.compute_plan(g, plan, machine_dir, return_wet=True)
    return total, wets


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 6, 3), (2, 2, 2), (5, 1, 7), (3, 9, 1), (6, 5, 4)])
@pytest.mark.parametrize("cutoff", [0.0, 6.0])
def test_degenerate_grids(dev, machine_dir, shape, cutoff):
    from DoseCUDA.engine import IMPTDoseEngine
    _say("grid %s cutoff %g" % (shape, cutoff))
    g, plan = _grid_plan(shape, 3.0, _spots([0, 5, 10]))
    eng = IMPTDoseEngine(g, plan, gpu_id=0, cutoff_sigma=cutoff)
    eng.compute_wet()
    dose = eng.dose().cpu().numpy().astype(np.float64)
    ref, wets = _oracle_dose_wet(g, plan, machine_dir)
    d_wet = np.abs(eng.beams[0].wet.cpu().numpy() - wets[0]).max()
    dmax = float(ref.max())
    err = float(np.abs(dose - ref).max())
    _say("  wet %.2e cm, dose err %.2e (D_max %.3g)" % (d_wet, err, dmax))
    assert d_wet <= 2e-4
    assert err <= max(1e-3 * dmax, 1e-6)


@pytest.mark.parametrize("wet_mode", ["full", "reach"])
def test_batch_degenerate_grids(dev, wet_mode):
    """A batch of the degenerate grids (one plan each, oblique beams): WET,
    dose and gradients bit-identical to the single-beam path."""
    from DoseCUDA.batch import BatchedDose
    from DoseCUDA.engine import IMPTDoseEngine
    _say("batch of degenerate grids, %s WET" % wet_mode)
    shapes = [(1, 1, 1), (1, 6, 3), (2, 2, 2), (5, 1, 7), (3, 9, 1), (6, 5, 4), (9, 33, 37)]
    engines = []
    for n, shape in enumerate(shapes):
        g, plan = _grid_plan(shape, 3.0, _spots([0, 5, 10]), gantry=30.0 + 40.0 * n, couch=20.0, seed=n)
        engines.append(IMPTDoseEngine(g, plan, gpu_id=0, cutoff_sigma=6.0))
    rng = np.random.default_rng(2)
    grads = [torch.from_numpy(rng.normal(size=e.shape).astype(np.float32)).to(dev) for e in engines]
    want = []
    for e, gd in zip(engines, grads):
        e.compute_wet(force=True)
        want.append(([b.wet.clone() for b in e.beams], e.dose().clone(),
                     [(m.clone(), x.clone()) for m, x in e.adjoint(gd)]))
    bd = BatchedDose(engines, wet_mode=wet_mode)
    if wet_mode == "full":
        bd.compute_wet(force=True)
    doses = bd.forward()
    grads_b = bd.adjoint(grads)
    torch.cuda.synchronize()
    k = 0
    for p, e in enumerate(engines):
        wets, dose, adj = want[p]
        assert torch.equal(doses[p], dose), "dose plan %d" % p
        for j, b in enumerate(e.beams):
            keep = ~(b.wet >= 1e29) if wet_mode == "reach" else torch.ones_like(b.wet, dtype=torch.bool)
            assert torch.equal(b.wet[keep], wets[j][keep]), "WET plan %d" % p
            assert torch.equal(grads_b[k][0], adj[j][0]) and torch.equal(grads_b[k][1], adj[j][1]), "grad plan %d" % p
            k += 1


@pytest.mark.parametrize("ids", [[98, 120], [-1, 3, 7]])
def test_beam_without_usable_layers(dev, machine_dir, ids):
    from DoseCUDA.engine import IMPTDoseEngine
    _say("energy ids %s" % ids)
    g, plan = _grid_plan((12, 14, 10), 3.0, _spots(ids))
    eng = IMPTDoseEngine(g, plan, gpu_id=0)
    params = eng.parameters()
    dose = eng.differentiable_dose(params)
    (dose * dose).sum().backward()
    ref, _ = _oracle_dose_wet(g, plan, machine_dir)
    assert float(np.abs(ref).max()) == 0.0
    assert float(dose.detach().abs().max()) == 0.0
    assert float(params[0][0].grad.abs().max()) == 0.0 and float(params[0][1].grad.abs().max()) == 0.0


@pytest.mark.parametrize("cutoff", [0.0, 6.0])
def test_range_shifter_model(dev, machine_dir, cutoff):
    """Beams on the second beam model (range-shifter label '4'): negative
    sigma_ms, and energy 0 with R80 = -0.65 mm (no voxel in range).  Dose within
    1e-3 of D_max of the oracle, which picks the model by the same label."""
    from DoseCUDA import IMPTPlan
    from DoseCUDA.engine import IMPTDoseEngine
    from DoseCUDA.workloads import _beam, _grid
    _say("range-shifter model, cutoff %g" % cutoff)
    g = _grid((20, 26, 22), 2.5)
    g.HU = np.random.default_rng(6).uniform(-300.0, 400.0, size=(20, 26, 22)).astype(np.float32)
    plan = IMPTPlan()
    for ga, ca in ((60.0, 10.0), (200.0, 0.0)):
        b = _beam(ga, np.asarray(_spots([0, 3, 30, 60]), np.float32), couch=ca)
        b.dicom_rangeshifter_label = '4'
        plan.addBeam(b)
    eng = IMPTDoseEngine(g, plan, gpu_id=0, cutoff_sigma=cutoff)
    params = eng.parameters()
    dose = eng.differentiable_dose(params)
    (dose * dose).sum().backward()
    ref, _ = _oracle_dose_wet(g, plan, machine_dir)
    D = dose.detach().cpu().numpy().astype(np.float64)
    dmax = float(ref.max())
    err = float(np.abs(D - ref).max())
    _say("  dose err %.2e (D_max %.3g)" % (err, dmax))
    assert dmax > 0.0 and err <= 1e-3 * dmax
    for mu, xy in params:
        assert torch.isfinite(mu.grad).all() and torch.isfinite(xy.grad).all()


def test_spots_out_of_reach(dev, machine_dir):
    from DoseCUDA.engine import IMPTDoseEngine
    _say("spots far outside the volume")
    g, plan = _grid_plan((10, 12, 14), 3.0, _spots([5, 20], coords=(400.0, 430.0)))
    eng = IMPTDoseEngine(g, plan, gpu_id=0)
    params = eng.parameters()
    dose = eng.differentiable_dose(params)
    (dose * dose).sum().backward()
    ref, _ = _oracle_dose_wet(g, plan, machine_dir)
    D = dose.detach().cpu().numpy().astype(np.float64)
    assert float(np.abs(D - ref).max()) <= max(1e-3 * float(ref.max()), 1e-12)
    assert torch.isfinite(params[0][0].grad).all() and torch.isfinite(params[0][1].grad).all()


def test_batch_with_empty_beam(dev):
    from DoseCUDA import IMPTPlan
    from DoseCUDA.batch import BatchedDose
    from DoseCUDA.engine import IMPTDoseEngine
    from DoseCUDA.workloads import _beam, _grid
    _say("batch with a beam without spots")
    g = _grid((16, 20, 18), 2.5)
    g.HU = np.random.default_rng(8).uniform(-200.0, 300.0, size=(16, 20, 18)).astype(np.float32)
    plan = IMPTPlan()
    plan.addBeam(_beam(0.0, np.asarray(_spots([2, 8]), np.float32)))
    plan.addBeam(_beam(90.0, np.zeros((0, 4), np.float32)))
    plan.addBeam(_beam(200.0, np.asarray(_spots([4]), np.float32), couch=15.0))
    engines = [IMPTDoseEngine(g, plan, gpu_id=0) for _ in range(2)]
    single = engines[0]
    params = single.parameters()
    d_single = single.differentiable_dose(params)
    (d_single * d_single).sum().backward()
    bd = BatchedDose(engines, wet_mode="reach")
    mus = [b.mu0.clone() for e in engines for b in e.beams]
    xys = [b.xy0.clone() for e in engines for b in e.beams]
    doses = bd.forward(mus, xys)
    grads = bd.adjoint([2.0 * d for d in doses])
    torch.cuda.synchronize()
    for d in doses:
        assert torch.equal(d, d_single.detach())
    for k, (gm, gx) in enumerate(grads[:3]):
        assert torch.equal(gm, params[k][0].grad) and torch.equal(gx, params[k][1].grad)


def _layer_tensor(rows, dev, dvp):
    """float32 [L, 5] layer records (spot_start, n_spots, energy_id as int32; R80; energy)."""
    lay = np.zeros((len(rows), 5), np.float32)
    lay.view(np.int32)[:, :3] = np.asarray(rows, np.int32).reshape(-1, 3)
    ids = np.clip(lay.view(np.int32)[:, 2], 0, dvp.shape[0] - 1)
    lay[:, 3] = dvp[ids, 1]
    lay[:, 4] = dvp[ids, 0]
    return torch.from_numpy(lay).to(dev)


@pytest.mark.parametrize("cutoff", [0.0, 6.0])
def test_overlapping_layer_records(dev, cutoff):
    """Through the C ABI, a layer record whose spot range intersects an earlier
    usable layer's is skipped: the dose equals that of the table without the
    skipped records bit for bit, gradients within 1e-5.  80 one-spot layers
    make the overlap scan run over two 64-lane passes (an overlap found in the
    second pass when the bad records come last, in the first when one comes
    first and shadows three later layers)."""
    from DoseCUDA import IMPTPlan
    from DoseCUDA import _native as N
    from DoseCUDA import device as D
    from DoseCUDA.workloads import _grid
    _say("overlapping layer records, cutoff %g" % cutoff)
    model = IMPTPlan().beam_models[0]
    dvp = model.divergence_params
    g = _grid((20, 24, 22), 2.5)
    g.HU = np.random.default_rng(5).uniform(-100.0, 200.0, size=(20, 24, 22)).astype(np.float32)
    ids = np.arange(10, 90)
    rng = np.random.default_rng(6)
    sl = np.stack([rng.uniform(-8, 8, ids.size), rng.uniform(-8, 8, ids.size), rng.uniform(0.3, 1.0, ids.size),
                   ids], axis=1).astype(np.float32)
    grid = D.make_grid(g.HU.shape, 2.5)
    geo = D.make_geom(D.beam_geometry(np.zeros(3, np.float32), np.asarray(g.origin, np.float32), 30.0, 0.0,
                                      model.VSADX, model.VSADY))
    dmodel = D.DeviceBeamModel(model, dev)
    rlsp = torch.from_numpy(np.asarray(g.RLSPFromHU("HitachiProbeatJHU"), np.float32)).to(dev)
    _, wet = D.wet_from_padded(D.rlsp_pad(rlsp, grid, dev), grid, geo, dev)
    gd = torch.from_numpy(rng.normal(size=g.HU.shape).astype(np.float32)).to(dev)
    base = [(k, 1, int(e)) for k, e in enumerate(ids)]

    def run(rows):
        sp = D.DeviceSpots(sl, dvp, dev, spatial=False)
        sp.layers = _layer_tensor(rows, dev, dvp)
        sp.n_layers = len(rows)
        sp.c = N.dh_spot_set(N.ptr(sp.spots).value, N.ptr(sp.layers).value, sp.n_spots, sp.n_layers)
        ws, nb = D.alloc_workspace(sp.n_spots, len(base) + 2, grid, dev)
        dose = torch.empty_like(wet)
        D.pb_forward(wet, dose, grid, geo, dmodel, sp, ws, nb, cutoff_sigma=cutoff)
        gm, gxy = D.pb_adjoint(wet, gd, grid, geo, dmodel, sp, ws, nb, cutoff_sigma=cutoff)
        torch.cuda.synchronize()
        return dose.cpu().numpy(), gm.cpu().numpy(), gxy.cpu().numpy()

    shadow = (69, 3, 50)   # spots 69..71 as one layer of energy 50
    cases = [
        (base + [shadow, (71, 2, 51)], base),                                        # extra records skipped
        ([shadow] + base, [shadow] + [r for r in base if r[0] not in (69, 70, 71)]),  # later layers skipped
    ]
    for bad, clean in cases:
        d1, m1, x1 = run(bad)
        d0, m0, x0 = run(clean)
        assert d0.max() > 0 and np.array_equal(d1, d0)
        assert np.abs(m1 - m0).max() <= 1e-5 * np.abs(m0).max()
        assert np.abs(x1 - x0).max() <= 1e-5 * np.abs(x0).max()
