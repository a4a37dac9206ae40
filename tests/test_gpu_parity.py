"""HIP kernels vs the FP64 oracle (GPU).

Tolerances (float32 kernels vs float64 oracle):
  * WET (cm): max |dWET| <= 2e-4 cm (0.002 mm) over the volume, raw and
    smoothed.
  * dose: max |dD| / D_max <= 1e-3 end to end (the dose bar is 5e-3), and
    <= 2e-5 for the pencil-beam kernel alone on identical WET input; 2%/2 mm
    gamma pass rate 1.0.
  * gradients: relative error of dL/dmu, dL/dxy vs the analytic FP64 adjoint
    <= 1e-4 (in the max norm over spots, relative to the max magnitude).
"""

import numpy as np
import pytest
import torch

from oracle import impt_oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-parity] " + msg, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


def _phantom_small(shape=(24, 28, 20), spacing=2.5, seed=7):
    from DoseCUDA.workloads import _grid
    g = _grid(shape, spacing)
    rng = np.random.default_rng(seed)
    hu = rng.uniform(-200.0, 300.0, size=shape).astype(np.float32)
    hu[:, 5:9, :] = 1200.0
    hu[3:6, :, 4:7] = -1000.0
    g.HU = hu
    return g


def _gpu_wet(g, plan, beam, dev):
    from DoseCUDA import device as D
    from DoseCUDA.plan_impt import _model_for_beam
    model = _model_for_beam(plan, beam)
    rlsp = g.RLSPFromHU(plan.machine_name)
    grid = D.make_grid(rlsp.shape, float(np.float32(g.spacing[0])))
    geo = D.make_geom(D.beam_geometry(np.asarray(beam.iso, np.float32), np.asarray(g.origin, np.float32),
                                      beam.gantry_angle, beam.couch_angle, model.VSADX, model.VSADY))
    padded = D.rlsp_pad(torch.from_numpy(rlsp).to(dev), grid, dev)
    raw, wet = D.wet_from_padded(padded, grid, geo, dev)
    return raw.cpu().numpy(), wet.cpu().numpy(), rlsp, model


def _oracle_geom(g, beam, model):
    return O.beam_geometry(np.asarray(beam.iso, np.float32), np.asarray(g.origin, np.float32), beam.gantry_angle,
                           beam.couch_angle, model.VSADX, model.VSADY)


@pytest.mark.parametrize("gantry,couch", [(0.0, 0.0), (120.0, 0.0), (240.0, 30.0), (90.0, 0.0)])
def test_wet_small_phantom(dev, gantry, couch):
    from DoseCUDA import IMPTPlan
    from DoseCUDA.workloads import _beam, grid_spots
    _say("wet small phantom gantry=%g couch=%g" % (gantry, couch))
    g = _phantom_small()
    plan = IMPTPlan()
    plan.addBeam(_beam(gantry, grid_spots([30], [0.0, 5.0], [0.0]), couch=couch))
    beam = plan.beam_list[0]
    raw, wet, rlsp, model = _gpu_wet(g, plan, beam, dev)
    geom = _oracle_geom(g, beam, model)
    s = float(np.float32(g.spacing[0]))
    raw_o = O.raymarch(rlsp, s, geom)
    wet_o = O.smooth(raw_o, s, geom)
    d_raw = np.abs(raw - raw_o)
    d_wet = np.abs(wet - wet_o)
    _say("max |d raw| %.3e cm, max |d wet| %.3e cm" % (d_raw.max(), d_wet.max()))
    assert d_raw.max() <= 2e-4 and d_wet.max() <= 2e-4


def test_rlsp_from_hu_matches_np_interp(dev):
    from DoseCUDA import device as D
    from DoseCUDA.engine import _hu_knots
    _say("rlsp_from_hu")
    rng = np.random.default_rng(3)
    hu = rng.uniform(-1500, 3500, size=(9, 11, 13)).astype(np.float32)
    hu[0, 0, :5] = [-1000.0, -742.5, 0.0, 3096.0, 24.0]
    hk, rk = _hu_knots("HitachiProbeatJHU")
    ref = np.array(np.interp(hu, hk, rk), dtype=np.single)
    grid = D.make_grid(hu.shape, 2.0)
    out = D.rlsp_from_hu(torch.from_numpy(hu).to(dev), grid, hk, rk, dev).cpu().numpy()
    nz, ny, nx = hu.shape
    q = out[: 4 * nx * ny * (nz + 1)].reshape(nz + 1, ny, nx, 4)
    expect = _quads_np(ref)
    # first component: np.interp's value within 1 ulp
    np.testing.assert_array_max_ulp(q[..., 0], expect[..., 0], maxulp=1)
    # bilinear terms: differences of values that are each within 1 ulp
    np.testing.assert_allclose(q[..., 1:], expect[..., 1:], rtol=0, atol=4 * np.finfo(np.float32).eps *
                               float(np.abs(ref).max()))


def _quads_np(rho):
    """Padded quads of rho without flags: (a, b - a, c - a, (d - c) - (b - a)) in float32."""
    nz, ny, nx = rho.shape
    pad = np.zeros((nz + 1, ny + 1, nx + 1), np.float32)
    pad[:nz, :ny, :nx] = rho
    a, b, c, d = pad[:, :ny, :nx], pad[:, :ny, 1:], pad[:, 1:, :nx], pad[:, 1:, 1:]
    return np.stack([a, b - a, c - a, (d - c) - (b - a)], axis=-1)


def _block_radius_np(table, rmax=15):
    """Run radius per block by brute force: the largest r <= rmax such that every
    block within Chebyshev distance r (inside the block grid) is uniform with
    the block's value; 0 for non-uniform blocks."""
    bz, by, bx = table.shape
    out = np.zeros(table.shape, np.float32)
    for k in range(bz):
        for j in range(by):
            for i in range(bx):
                v = table[k, j, i]
                if np.isnan(v):
                    continue
                r = 0
                while r < rmax:
                    q = r + 1
                    cube = table[max(k - q, 0):k + q + 1, max(j - q, 0):j + q + 1, max(i - q, 0):i + q + 1]
                    if not np.all(cube == v):
                        break
                    r = q
                out[k, j, i] = r
    return out


def _block_table_np(rho):
    """Reference block table: per 4x4x4 block of cells, the common value of the
    corners its cells touch (zero past the high end), NaN if they differ."""
    nz, ny, nx = rho.shape
    pad = np.zeros((nz + 1, ny + 1, nx + 1), np.float32)
    pad[:nz, :ny, :nx] = rho
    bz, by, bx = (nz + 3) // 4, (ny + 3) // 4, (nx + 3) // 4
    out = np.full((bz, by, bx), np.nan, np.float32)
    for k in range(bz):
        for j in range(by):
            for i in range(bx):
                c = pad[4 * k: min(4 * k + 5, nz + 1), 4 * j: min(4 * j + 5, ny + 1), 4 * i: min(4 * i + 5, nx + 1)]
                if np.all(c == c.flat[0]):
                    out[k, j, i] = c.flat[0]
    return out


@pytest.mark.parametrize("gantry,couch", [(0.0, 0.0), (120.0, 30.0), (270.0, 0.0)])
def test_block_table_and_uniform_skipping(dev, gantry, couch):
    """Air box with a noisy tissue cube and a water slab: the block table equals
    a numpy restatement, and the ray-march (which adds runs of samples in
    uniform blocks at once) matches the oracle's sample-by-sample WET."""
    from DoseCUDA import IMPTPlan
    from DoseCUDA.workloads import _beam, _grid, grid_spots
    _say("block table / skipping gantry=%g couch=%g" % (gantry, couch))
    shape = (30, 36, 33)
    g = _grid(shape, 2.5)
    rng = np.random.default_rng(11)
    hu = np.full(shape, -1000.0, np.float32)
    hu[8:22, 10:26, 9:24] = rng.normal(40, 20, size=(14, 16, 15))
    hu[:, 28:31, :] = 0.0
    g.HU = hu
    plan = IMPTPlan()
    plan.addBeam(_beam(gantry, grid_spots([30], [0.0], [0.0]), couch=couch))
    beam = plan.beam_list[0]
    raw, wet, rlsp, model = _gpu_wet(g, plan, beam, dev)
    # table of the padded volume
    from DoseCUDA import device as D
    grid = D.make_grid(rlsp.shape, float(np.float32(g.spacing[0])))
    padded = D.rlsp_pad(torch.from_numpy(rlsp).to(dev), grid, dev).cpu().numpy()
    nz, ny, nx = shape
    nb = ((nz + 3) // 4) * ((ny + 3) // 4) * ((nx + 3) // 4)
    assert padded.size == 4 * nx * ny * (nz + 1) + 4 * nb
    pairs = padded[4 * nx * ny * (nz + 1): 4 * nx * ny * (nz + 1) + 2 * nb].reshape(nb, 2)
    table = pairs[:, 0].reshape((nz + 3) // 4, (ny + 3) // 4, (nx + 3) // 4)
    radius = pairs[:, 1].reshape(table.shape)
    want = _block_table_np(rlsp)
    assert np.array_equal(np.isnan(table), np.isnan(want))
    assert np.array_equal(table[~np.isnan(want)], want[~np.isnan(want)])
    want_r = _block_radius_np(want)
    assert np.array_equal(radius, want_r), "run radii differ from the brute force"
    assert want_r.max() >= 1
    # quads: the bilinear form, bit for bit
    q = padded[: 4 * nx * ny * (nz + 1)].reshape(nz + 1, ny, nx, 4)
    assert np.array_equal(q.view(np.uint32), _quads_np(rlsp).view(np.uint32))
    uniform = float(np.mean(~np.isnan(want)))
    geom = _oracle_geom(g, beam, model)
    s = float(np.float32(g.spacing[0]))
    raw_o = O.raymarch(rlsp, s, geom)
    wet_o = O.smooth(raw_o, s, geom)
    _say("uniform blocks %.2f, max |d raw| %.3e cm, max |d wet| %.3e cm"
         % (uniform, np.abs(raw - raw_o).max(), np.abs(wet - wet_o).max()))
    assert uniform > 0.3
    assert np.abs(raw - raw_o).max() <= 2e-4 and np.abs(wet - wet_o).max() <= 2e-4


def test_water_slab_end_to_end(dev, machine_dir):
    """computeIMPTPlan through the dose_kernels drop-in vs the oracle plan."""
    from DoseCUDA.workloads import water_slab_plan
    _say("water slab 100^3 end to end")
    g, plan = water_slab_plan()
    g.computeIMPTPlan(plan)
    ref, _ = O.compute_plan(g, plan, machine_dir)
    m = O.compare(g.dose, ref)
    gam = O.gamma_pass_rate(g.dose, ref, 2.5)
    _say("max|dD|/Dmax %.3e mean|dD| %.3e gamma %.4f" % (m["max_rel_to_dmax"], m["mean_abs"], gam))
    assert m["max_rel_to_dmax"] <= 1e-3
    assert gam == 1.0


def _pb_inputs(g, plan, dev, wet_f64):
    from DoseCUDA import device as D
    beam = plan.beam_list[0]
    from DoseCUDA.plan_impt import _model_for_beam
    model = _model_for_beam(plan, beam)
    grid = D.make_grid(g.HU.shape, float(np.float32(g.spacing[0])))
    geo = D.make_geom(D.beam_geometry(np.asarray(beam.iso, np.float32), np.asarray(g.origin, np.float32),
                                      beam.gantry_angle, beam.couch_angle, model.VSADX, model.VSADY))
    dmodel = D.DeviceBeamModel(model, dev)
    dspots = D.DeviceSpots(beam.spot_list, model.divergence_params, dev)
    ws, nb = D.alloc_workspace(dspots.n_spots, dspots.n_layers, grid, dev)
    wet = torch.from_numpy(wet_f64.astype(np.float32)).to(dev)
    return D, model, grid, geo, dmodel, dspots, ws, nb, wet


@pytest.mark.parametrize("cutoff", [0.0, 6.0])
def test_pencil_beam_kernel_same_wet(dev, cutoff):
    from DoseCUDA.workloads import water_slab_plan
    _say("pencil beam kernel only, cutoff=%g" % cutoff)
    g, plan = water_slab_plan()
    beam = plan.beam_list[0]
    model = plan.beam_models[0]
    rlsp = g.RLSPFromHU(plan.machine_name)
    s = float(np.float32(g.spacing[0]))
    geom = _oracle_geom(g, beam, model)
    wet_o = O.wet(rlsp, s, geom).astype(np.float32).astype(np.float64)
    D, model, grid, geo, dmodel, dspots, ws, nb, wet = _pb_inputs(g, plan, dev, wet_o)
    dose = torch.empty_like(wet)
    D.pb_forward(wet, dose, grid, geo, dmodel, dspots, ws, nb, cutoff_sigma=cutoff)
    _, spots, layers, r80 = O.spots_and_layers(beam.spot_list, model.divergence_params)
    ref = O.pb_dose(wet_o, s, geom, spots, layers, r80, O.ModelArrays(model))
    m = O.compare(dose.cpu().numpy(), ref)
    _say("max|dD|/Dmax %.3e" % m["max_rel_to_dmax"])
    assert m["max_rel_to_dmax"] <= 2e-5


def test_adjoint_vs_oracle(dev):
    from DoseCUDA.workloads import water_slab_plan
    _say("adjoint vs oracle analytic gradient")
    g, plan = water_slab_plan()
    beam = plan.beam_list[0]
    model = plan.beam_models[0]
    rlsp = g.RLSPFromHU(plan.machine_name)
    s = float(np.float32(g.spacing[0]))
    geom = _oracle_geom(g, beam, model)
    wet_o = O.wet(rlsp, s, geom).astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(11)
    gvol = np.zeros(g.HU.shape, np.float32)
    gvol[30:70, 20:60, 30:70] = rng.normal(size=(40, 40, 40)).astype(np.float32)
    D, model, grid, geo, dmodel, dspots, ws, nb, wet = _pb_inputs(g, plan, dev, wet_o)
    gmu, gxy = D.pb_adjoint(wet, torch.from_numpy(gvol).to(dev), grid, geo, dmodel, dspots, ws, nb,
                            cutoff_sigma=6.0)
    gmu = dspots.to_input_order(gmu).cpu().numpy()
    gxy = dspots.to_input_order(gxy).cpu().numpy()
    order, spots, layers, r80 = O.spots_and_layers(beam.spot_list, model.divergence_params)
    rmu, rxy = O.pb_grad(wet_o, gvol, s, geom, spots, layers, r80, O.ModelArrays(model))
    ref_mu = np.empty_like(rmu)
    ref_mu[order] = rmu
    ref_xy = np.empty_like(rxy)
    ref_xy[order] = rxy
    e_mu = np.abs(gmu - ref_mu).max() / np.abs(ref_mu).max()
    e_xy = np.abs(gxy - ref_xy).max() / np.abs(ref_xy).max()
    _say("rel err mu %.3e xy %.3e" % (e_mu, e_xy))
    assert e_mu <= 1e-4 and e_xy <= 1e-4


def test_forward_deterministic(dev):
    from DoseCUDA.workloads import water_slab_plan
    from DoseCUDA.engine import IMPTDoseEngine
    _say("determinism")
    g, plan = water_slab_plan()
    eng = IMPTDoseEngine(g, plan)
    d1 = eng.dose().clone()
    d2 = eng.dose().clone()
    assert torch.equal(d1, d2)
    gd = torch.ones_like(d1)
    a1 = eng.adjoint(gd)[0]
    a2 = eng.adjoint(gd)[0]
    assert torch.equal(a1[0], a2[0]) and torch.equal(a1[1], a2[1])


def _small_plan(gantries, couches, layers, radius=15.0, pitch=3.0, seed=5):
    from DoseCUDA import IMPTPlan
    from DoseCUDA.workloads import _beam, grid_spots
    rng = np.random.default_rng(seed)
    plan = IMPTPlan()
    c = np.arange(-radius, radius + 1e-6, pitch)
    for ga, ca in zip(gantries, couches):
        sl = grid_spots(layers, c, c, disk_radius=radius)
        sl[:, 2] = rng.uniform(0.2, 1.0, size=sl.shape[0]).astype(np.float32)
        plan.addBeam(_beam(ga, sl, couch=ca))
    return plan


def test_oblique_multibeam_end_to_end(dev, machine_dir):
    """Heterogeneous phantom, beams at gantry 120 / couch 30 and gantry 240,
    81-spot layers: drop-in computeIMPTPlan vs the oracle plan."""
    _say("oblique multi-beam end to end")
    g = _phantom_small(shape=(40, 56, 48), spacing=2.5, seed=3)
    plan = _small_plan([120.0, 240.0], [30.0, 0.0], [22, 27, 31])
    plan.n_fractions = 2
    g.computeIMPTPlan(plan)
    ref, ref_beams = O.compute_plan(g, plan, machine_dir)
    m = O.compare(g.dose, ref)
    gam = O.gamma_pass_rate(g.dose, ref, 2.5)
    _say("max|dD|/Dmax %.3e gamma %.4f" % (m["max_rel_to_dmax"], gam))
    assert m["max_rel_to_dmax"] <= 1e-3 and gam == 1.0
    for got, want in zip(g.beam_doses, ref_beams):
        assert O.compare(got, want)["max_rel_to_dmax"] <= 1e-3


def test_adjoint_81_spot_layers(dev):
    """Layers of 81 spots, oblique beam: the packed adjoint runs each brick's
    reachable spots in passes with several voxel subsets per pass."""
    from DoseCUDA import device as D
    _say("adjoint, 81-spot layers, oblique beam")
    g = _phantom_small(shape=(40, 56, 48), spacing=2.5, seed=4)
    plan = _small_plan([120.0], [30.0], [22, 27, 31, 35])
    beam = plan.beam_list[0]
    model = plan.beam_models[0]
    rlsp = g.RLSPFromHU(plan.machine_name)
    s = float(np.float32(g.spacing[0]))
    geom = _oracle_geom(g, beam, model)
    wet_o = O.wet(rlsp, s, geom).astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(2)
    gvol = rng.normal(size=g.HU.shape).astype(np.float32)
    gvol[rng.uniform(size=g.HU.shape) < 0.5] = 0.0
    D_, model, grid, geo, dmodel, dspots, ws, nb, wet = _pb_inputs(g, plan, dev, wet_o)
    assert dspots.n_spots == 4 * 81
    gmu, gxy = D.pb_adjoint(wet, torch.from_numpy(gvol).to(dev), grid, geo, dmodel, dspots, ws, nb, cutoff_sigma=6.0)
    gmu = dspots.to_input_order(gmu).cpu().numpy()
    gxy = dspots.to_input_order(gxy).cpu().numpy()
    order, spots, layers, r80 = O.spots_and_layers(beam.spot_list, model.divergence_params)
    rmu, rxy = O.pb_grad(wet_o, gvol, s, geom, spots, layers, r80, O.ModelArrays(model))
    ref_mu = np.empty_like(rmu)
    ref_mu[order] = rmu
    ref_xy = np.empty_like(rxy)
    ref_xy[order] = rxy
    e_mu = np.abs(gmu - ref_mu).max() / np.abs(ref_mu).max()
    e_xy = np.abs(gxy - ref_xy).max() / np.abs(ref_xy).max()
    _say("rel err mu %.3e xy %.3e" % (e_mu, e_xy))
    assert e_mu <= 1e-4 and e_xy <= 1e-4


class _NonUniformModel:
    """A beam model whose depth grid is not uniform (exercises the bisection
    lookup): every energy's table resampled at depths k^1.5 / 20^0.5 mm."""

    def __init__(self, base):
        k = np.arange(base.lut_depths.shape[1], dtype=np.float64)
        dep = (k ** 1.5 / np.sqrt(20.0)).astype(np.float32)
        self.VSADX, self.VSADY = base.VSADX, base.VSADY
        self.divergence_params = base.divergence_params
        self.lut_depths = np.tile(dep, (base.lut_depths.shape[0], 1))
        self.lut_sigmas = np.stack([np.interp(dep, base.lut_depths[e], base.lut_sigmas[e])
                                    for e in range(base.lut_depths.shape[0])]).astype(np.float32)
        self.lut_idds = np.stack([np.interp(dep, base.lut_depths[e], base.lut_idds[e])
                                  for e in range(base.lut_depths.shape[0])]).astype(np.float32)


def test_nonuniform_lut_depths(dev):
    from DoseCUDA import device as D
    from DoseCUDA.workloads import water_slab_plan
    _say("non-uniform LUT depth grid")
    g, plan = water_slab_plan(n=48, spacing=2.5)
    beam = plan.beam_list[0]
    model = _NonUniformModel(plan.beam_models[0])
    rlsp = g.RLSPFromHU(plan.machine_name)
    s = float(np.float32(g.spacing[0]))
    geom = _oracle_geom(g, beam, model)
    wet_o = O.wet(rlsp, s, geom).astype(np.float32).astype(np.float64)
    dmodel = D.DeviceBeamModel(model, dev)
    assert not dmodel.uniform
    grid = D.make_grid(g.HU.shape, s)
    geo = D.make_geom(D.beam_geometry(np.asarray(beam.iso, np.float32), np.asarray(g.origin, np.float32),
                                      beam.gantry_angle, beam.couch_angle, model.VSADX, model.VSADY))
    dspots = D.DeviceSpots(beam.spot_list, model.divergence_params, dev)
    ws, nb = D.alloc_workspace(dspots.n_spots, dspots.n_layers, grid, dev)
    wet = torch.from_numpy(wet_o.astype(np.float32)).to(dev)
    dose = torch.empty_like(wet)
    D.pb_forward(wet, dose, grid, geo, dmodel, dspots, ws, nb, cutoff_sigma=6.0)
    _, spots, layers, r80 = O.spots_and_layers(beam.spot_list, model.divergence_params)
    ref = O.pb_dose(wet_o, s, geom, spots, layers, r80, O.ModelArrays(model))
    m = O.compare(dose.cpu().numpy(), ref)
    _say("max|dD|/Dmax %.3e" % m["max_rel_to_dmax"])
    assert m["max_rel_to_dmax"] <= 2e-5
