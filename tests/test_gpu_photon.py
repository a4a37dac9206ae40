"""Photon collapsed-cone path on the GPU (libdosehip.so dh_photon_*) against
the float64 oracle (oracle/imrt_oracle.*), stage by stage and end to end.
Reference: photon_dose_cuda (dosemodule.cu:377-537, IMRTClasses.cu:353-405),
rayTraceKernel (CudaClasses.cu:180-228), termaKernel / cccKernel
(IMRTClasses.cu:209-350), computeIMRTPlan (plan_imrt.py:126-158).

Tolerances: depth 2e-4 g/cm^2 (the proton WET bar); TERMA / electron 1e-5 of
their maximum on the same depth; dose 1e-4 of D_max on the same TERMA, and
end to end (each side its own depth and TERMA) 1e-3 of D_max with the 2%/2 mm
gamma pass rate >= 0.995 (the north-star bar is 5e-3).  A voxel whose TERMA
lies within 1e-4 of the 0.01 threshold of cccKernel may fall on either side
in float32 and float64; such voxels are counted and left out of end-to-end
comparisons."""

import os

import numpy as np
import pytest
import torch

from oracle import imrt_oracle as OR
from oracle import impt_oracle as O

pytestmark = pytest.mark.gpu

PHOT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dosecuda-jax_amd", "DoseCUDA",
                    "lookuptables", "photons")


def _say(msg):
    print("[gpu-photon] " + msg, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


@pytest.fixture(scope="module")
def varian():
    from DoseCUDA import IMRTPlan
    return IMRTPlan("VarianTrueBeamHF")


def _hetero_grid(shape, spacing, seed):
    from DoseCUDA import IMRTDoseGrid
    g = IMRTDoseGrid()
    g.spacing = np.array([spacing] * 3, dtype=np.float64)
    g.size = np.array(shape)
    g.origin = np.array([-shape[2] * spacing / 2.0, -shape[1] * spacing / 2.0, -shape[0] * spacing / 2.0])
    rng = np.random.default_rng(seed)
    hu = np.full(shape, -1000.0, np.float32)
    m = 2
    hu[m:-m, m:-m, m:-m] = rng.normal(20.0, 30.0, size=tuple(n - 2 * m for n in shape))
    nz, ny, nx = shape
    hu[nz // 3:nz // 3 + 4, m:-m, m:-m] = rng.uniform(700, 1200, size=(4, ny - 2 * m, nx - 2 * m))  # bone slab
    hu[nz // 2:nz // 2 + 5, ny // 2:ny // 2 + 6, nx // 3:nx // 3 + 6] = -800.0  # lung-like pocket
    g.HU = hu
    return g


def _cp(plan, model_index, ga, ta, ca, half_x, half_y, mu=100.0, iso=(0.0, 0.0, 0.0), irregular=False):
    from DoseCUDA.plan_imrt import IMRTControlPoint
    bm = plan.beam_models[model_index]
    leaves = np.zeros((2, bm.n_mlc_pairs), np.float32)
    leaves[0], leaves[1] = -half_x, half_x
    out = (bm.mlc_offsets < -half_y) | (bm.mlc_offsets > half_y)
    leaves[:, out] = 0.0
    if irregular:
        rng = np.random.default_rng(int(ga) + 7)
        open_ = ~out
        leaves[0, open_] += rng.uniform(-10, 10, open_.sum()).astype(np.float32)
        leaves[1, open_] += rng.uniform(-10, 10, open_.sum()).astype(np.float32)
    mlc = plan._mlc_array(leaves)
    return IMRTControlPoint(np.array(iso, np.float32), mu, mlc, ga, ca, ta, None, None)


def _stages(g, model, cp, dev, density=None):
    """GPU depth, TERMA, electron and dose of one control point (device pipeline)."""
    from DoseCUDA import photon as PH
    rho = OR.density_from_hu(g.HU, os.path.join(PHOT, "VarianTrueBeamHF")) if density is None else density
    with torch.cuda.device(dev):
        dm = PH.DevicePhotonModel(model, dev)
        beam = PH.control_point(cp, np.asarray(g.origin, np.float32), dm, dev)
        ws = PH.PhotonWorkspace(rho.shape, float(np.float32(g.spacing[0])), dev)
        ws.set_density(torch.from_numpy(rho).to(dev))
        dose = torch.empty(rho.shape, dtype=torch.float32, device=dev)
        PH.cp_dose(ws, dm, beam, dose)
        torch.cuda.synchronize()
        return rho, ws.depth.cpu().numpy(), ws.terma.cpu().numpy(), ws.electron.cpu().numpy(), dose.cpu().numpy()


def _geom(g, model, cp):
    return OR.photon_geometry(np.asarray(cp.iso, np.float32), np.asarray(g.origin, np.float32), cp.ga, cp.ta,
                              cp.ca, np.float32(model.primary_source_distance))


def test_density_records_from_hu(dev):
    """The density half of the record volume equals np.interp of HU (rounded
    to float32), with AIR_DENSITY in the padding."""
    from DoseCUDA import photon as PH
    _say("density records from HU")
    rng = np.random.default_rng(1)
    hu = rng.uniform(-1100, 3200, size=(5, 6, 7)).astype(np.float32)
    tab = np.loadtxt(os.path.join(PHOT, "VarianTrueBeamHF", "HU_Density.csv"), delimiter=",", skiprows=1)
    want = np.interp(hu, tab[:, 0], tab[:, 1]).astype(np.float32)
    with torch.cuda.device(dev):
        ws = PH.PhotonWorkspace(hu.shape, 2.0, dev)
        ws.set_density_from_hu(torch.from_numpy(hu).to(dev), tab[:, 0], tab[:, 1])
        allr = ws.records.cpu().numpy()
        torch.cuda.synchronize()
    npad = 7 * 8 * 9
    rec = allr[8 * npad:12 * npad].reshape(7, 8, 9, 4)   # density-only copy (k', j', i', 4 floats)
    pairs = allr[:8 * npad].reshape(7, 8, 9, 8)         # (TERMA form, density form) per cell
    np.testing.assert_array_equal(pairs[..., 4:], rec)
    a = rec[..., 0]
    inner = a[1:-1, 1:-1, 1:-1]
    np.testing.assert_array_equal(inner, want)
    border = np.ones_like(a, bool)
    border[1:-1, 1:-1, 1:-1] = False
    assert np.all(a[border] == np.float32(0.0012))
    # the face form: a + (b - a) is the +x neighbour b up to the rounding of the float32 difference
    np.testing.assert_allclose(rec[1:-1, 1:-1, 1:-2, 0] + rec[1:-1, 1:-1, 1:-2, 1], want[:, :, 1:], rtol=0,
                               atol=4e-7)


def _block_table_bruteforce(a, shape):
    """Block table of a padded density corner array a (k', j', i'): per block
    of 4x4x4 sample cells (cells 0 .. n per axis) the common value of corners
    4b .. 4b + 4 (AIR_DENSITY past the padded grid) or NaN, and the largest
    r <= 15 with every block of the table within Chebyshev distance r equal."""
    nb = [(n + 4) // 4 for n in shape]                  # (z, y, x)
    ext = np.full([4 * m + 1 for m in nb], np.float32(0.0012), np.float32)
    ext[:a.shape[0], :a.shape[1], :a.shape[2]] = a
    val = np.full(nb, np.nan, np.float32)
    for bz in range(nb[0]):
        for by in range(nb[1]):
            for bx in range(nb[2]):
                c = ext[4 * bz:4 * bz + 5, 4 * by:4 * by + 5, 4 * bx:4 * bx + 5]
                if np.all(c == c.flat[0]):
                    val[bz, by, bx] = c.flat[0]
    rad = np.zeros(nb, np.float32)
    for idx in np.ndindex(*nb):
        v = val[idx]
        if v != v:
            continue
        r = 0
        while r < 15:
            lo = [max(i - r - 1, 0) for i in idx]
            hi = [min(i + r + 2, m) for i, m in zip(idx, nb)]
            if not np.all(val[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] == v):
                break
            r += 1
        rad[idx] = r
    return val, rad


def _uniform_grid(shape, spacing):
    """Air around a water ellipsoid of exactly 0 HU with a bone rod and a lung
    sphere: large uniform regions (the ray-trace's runs) beside edges."""
    from DoseCUDA import IMRTDoseGrid
    g = IMRTDoseGrid()
    g.spacing = np.array([spacing] * 3, dtype=np.float64)
    g.size = np.array(shape)
    g.origin = np.array([-shape[2] * spacing / 2.0, -shape[1] * spacing / 2.0, -shape[0] * spacing / 2.0])
    z, y, x = np.meshgrid(*[np.arange(n) - (n - 1) / 2.0 for n in shape], indexing="ij")
    hu = np.full(shape, -1000.0, np.float32)
    hu[(x / (0.3 * shape[2])) ** 2 + (y / (0.3 * shape[1])) ** 2 + (z / (0.35 * shape[0])) ** 2 <= 1.0] = 0.0
    hu[((x - 3) ** 2 + (y + 2) ** 2 <= 9) & (np.abs(z) < 0.3 * shape[0])] = 1000.0
    hu[(x + 5) ** 2 + (y - 4) ** 2 + (z - 2) ** 2 <= 20] = -700.0
    g.HU = hu
    return g


def test_density_block_table(dev):
    """The block table after the records (value or NaN per 4x4x4 block of
    sample cells, run radius) equals a brute force of its definition on a
    phantom with uniform regions and edges."""
    from DoseCUDA import photon as PH
    _say("density block table")
    g = _uniform_grid((60, 64, 72), 2.0)
    tab = np.loadtxt(os.path.join(PHOT, "VarianTrueBeamHF", "HU_Density.csv"), delimiter=",", skiprows=1)
    shape = g.HU.shape
    with torch.cuda.device(dev):
        ws = PH.PhotonWorkspace(shape, 2.0, dev)
        ws.set_density_from_hu(torch.from_numpy(g.HU).to(dev), tab[:, 0], tab[:, 1])
        allr = ws.records.cpu().numpy()
    npad = int(np.prod([n + 2 for n in shape]))
    a = allr[8 * npad:12 * npad].reshape(*[n + 2 for n in shape], 4)[..., 0]
    val, rad = _block_table_bruteforce(a, shape)
    nb = val.size
    got = allr[12 * npad:12 * npad + 2 * nb].reshape(*val.shape, 2)
    np.testing.assert_array_equal(got[..., 0], val)
    np.testing.assert_array_equal(got[..., 1], rad)
    assert np.sum(val == val) > nb // 4 and rad.max() >= 2     # the phantom has runs to take


@pytest.mark.parametrize("ga,ta,spacing,dense", [(0.0, 0.0, 2.0, False), (90.0, 0.0, 2.5, False),
                                                 (33.0, 20.0, 2.0, False), (250.0, 335.0, 3.0, False),
                                                 (33.0, 20.0, 2.0, True), (150.0, 40.0, 2.5, True)])
def test_raytrace_uniform_regions_vs_oracle(dev, varian, ga, ta, spacing, dense):
    """Depth of every voxel vs the oracle on a phantom of large uniform regions
    (runs of samples added at once), axis-aligned (gantry 0 / 90: samples on
    cell boundaries, rays along one axis) and oblique.  dense: the air inside
    a 3-voxel margin becomes density 0.5, so that one sample too many or too
    few in a run there (0.05 g/cm^2) is far outside the tolerance; the margin
    keeps air at the grid's faces, where a sample's bounds test in float32
    (the reference's arithmetic) and in the oracle's float64 can differ by one
    sample of air (1.2e-4 g/cm^2)."""
    from DoseCUDA import photon as PH
    _say("uniform-region ray-trace gantry %g couch %g @ %g mm%s" % (ga, ta, spacing, ", dense" if dense else ""))
    g = _uniform_grid((48, 56, 64), spacing)
    model = varian.beam_models[0]
    cp = _cp(varian, 0, ga, ta, 0.0, 50.0, 50.0)
    rho = OR.density_from_hu(g.HU, os.path.join(PHOT, "VarianTrueBeamHF"))
    if dense:
        inner = rho[3:-3, 3:-3, 3:-3]
        inner[g.HU[3:-3, 3:-3, 3:-3] == -1000.0] = np.float32(0.5)
    dmodel = PH.DevicePhotonModel(model, dev)
    with torch.cuda.device(dev):
        ws = PH.PhotonWorkspace(g.HU.shape, float(np.float32(spacing)), dev)
        ws.set_density(torch.from_numpy(rho.astype(np.float32)).to(dev))
        cpc = PH.control_point(cp, g.origin, dmodel, dev)
        PH.raytrace(ws, cpc)
        depth = ws.depth.cpu().numpy()
    ref = OR.raytrace(rho, float(np.float32(spacing)), _geom(g, model, cp))
    err = np.max(np.abs(depth - ref))
    _say("  depth %.2e g/cm^2 (max %.1f)" % (err, ref.max()))
    assert err < 2e-4, err


@pytest.mark.parametrize("machine,energy,ga,ta,ca,spacing", [
    ("VarianTrueBeamHF", 0, 0.0, 0.0, 0.0, 2.0), ("VarianTrueBeamHF", 1, 90.0, 0.0, 30.0, 2.5),
    ("VarianTrueBeamHF", 0, 217.0, 35.0, 80.0, 3.0), ("ElektaVersaSibAxesse", 1, 300.0, 10.0, 355.0, 2.0),
    ("ElektaVersaSibAxesse", 3, 45.0, 0.0, 0.0, 1.5)])
def test_stages_vs_oracle(dev, machine, energy, ga, ta, ca, spacing):
    """Depth vs the oracle; TERMA / electron vs the oracle on the GPU's depth;
    dose vs the oracle on the GPU's TERMA, density and electron.  Both
    machines (60 leaf pairs of 5 / 10 mm; 80 of 5 mm with the MLC 643 mm from
    the isocentre), 6 MV, 10 MV, 6 MV FFF and 15 MV."""
    from DoseCUDA import IMRTPlan
    _say("stages %s energy %d ga=%g ta=%g ca=%g s=%g" % (machine, energy, ga, ta, ca, spacing))
    plan = IMRTPlan(machine)
    g = _hetero_grid((26, 30, 28), spacing, seed=int(ga))
    model = plan.beam_models[energy]
    cp = _cp(plan, energy, ga, ta, ca, 20.0, 25.0, mu=80.0, irregular=True)
    rho = OR.density_from_hu(g.HU, os.path.join(PHOT, machine))
    rho, depth, terma, el, dose = _stages(g, model, cp, dev, density=rho)
    geom = _geom(g, model, cp)
    s = float(np.float32(g.spacing[0]))
    d_ref = OR.raytrace(rho, s, geom)
    assert np.max(np.abs(depth - d_ref)) < 2e-4, np.max(np.abs(depth - d_ref))
    t_ref, e_ref = OR.terma(depth.astype(np.float64), rho.shape, s, geom, model, OR.mlc_rows(cp.mlc))
    assert np.max(np.abs(terma - t_ref)) < 1e-5 * t_ref.max(), np.max(np.abs(terma - t_ref)) / t_ref.max()
    assert np.max(np.abs(el - e_ref)) < 1e-5 * max(e_ref.max(), 1e-30)
    dose_ref = OR.ccc(terma, rho, el.astype(np.float64), s, geom, model.kernel, np.float32(model.mu_calibration),
                      np.float32(cp.mu))
    err = np.max(np.abs(dose - dose_ref)) / dose_ref.max()
    _say("  depth %.2e, TERMA %.2e, dose %.2e of D_max" % (np.max(np.abs(depth - d_ref)),
                                                           np.max(np.abs(terma - t_ref)) / t_ref.max(), err))
    assert dose_ref.max() > 0 and err < 1e-4, err


def _mixed_grid(shape, spacing, seed):
    """The uniform-region phantom with its ellipsoid's lower half replaced by
    noisy tissue: uniform runs and non-uniform samples on the same rays."""
    g = _uniform_grid(shape, spacing)
    rng = np.random.default_rng(seed)
    hu = g.HU
    nz = shape[0]
    water = hu == 0.0
    water[nz // 2:] = False
    hu[water] = rng.normal(30.0, 40.0, size=int(water.sum())).astype(np.float32)
    return g


@pytest.mark.parametrize("seed", range(8))
def test_random_geometries_vs_oracle(dev, seed):
    """Randomised control points (machine, energy, gantry, couch, collimator,
    spacing, isocentre offset, field size, irregular leaves) on a phantom with
    uniform regions and noisy tissue: depth vs the oracle (2e-4 g/cm^2),
    TERMA / electron vs the oracle on the GPU's depth (1e-5 of max), dose vs
    the oracle on the GPU's TERMA (1e-4 of D_max)."""
    from DoseCUDA import IMRTPlan
    rng = np.random.default_rng(100 + seed)
    machine = ("VarianTrueBeamHF", "ElektaVersaSibAxesse")[seed % 2]
    plan = IMRTPlan(machine)
    energy = int(rng.integers(len(plan.beam_models)))
    ga, ta, ca = float(rng.uniform(0, 360)), float(rng.uniform(-40, 40)) % 360.0, float(rng.uniform(0, 360))
    spacing = float(rng.choice([1.5, 2.0, 2.5, 3.0]))
    shape = tuple(int(v) for v in rng.integers(24, 34, size=3))
    iso = tuple(float(v) for v in rng.uniform(-8.0, 8.0, size=3))
    hx, hy = float(rng.uniform(10, 30)), float(rng.uniform(10, 30))
    _say("random %d: %s energy %d ga=%.1f ta=%.1f ca=%.1f s=%g shape %s" % (seed, machine, energy, ga, ta, ca,
                                                                             spacing, shape))
    g = _mixed_grid(shape, spacing, seed)
    model = plan.beam_models[energy]
    cp = _cp(plan, energy, ga, ta, ca, hx, hy, mu=60.0, iso=iso, irregular=True)
    rho = OR.density_from_hu(g.HU, os.path.join(PHOT, machine))
    rho, depth, terma, el, dose = _stages(g, model, cp, dev, density=rho)
    geom = _geom(g, model, cp)
    s = float(np.float32(g.spacing[0]))
    d_ref = OR.raytrace(rho, s, geom)
    assert np.max(np.abs(depth - d_ref)) < 2e-4, np.max(np.abs(depth - d_ref))
    t_ref, e_ref = OR.terma(depth.astype(np.float64), rho.shape, s, geom, model, OR.mlc_rows(cp.mlc))
    assert np.max(np.abs(terma - t_ref)) < 1e-5 * t_ref.max()
    assert np.max(np.abs(el - e_ref)) < 1e-5 * max(e_ref.max(), 1e-30)
    dose_ref = OR.ccc(terma, rho, el.astype(np.float64), s, geom, model.kernel, np.float32(model.mu_calibration),
                      np.float32(cp.mu))
    err = np.max(np.abs(dose - dose_ref)) / dose_ref.max()
    _say("  depth %.2e, TERMA %.2e, dose %.2e of D_max" % (np.max(np.abs(depth - d_ref)),
                                                           np.max(np.abs(terma - t_ref)) / t_ref.max(), err))
    assert dose_ref.max() > 0 and err < 1e-4, err


def test_dropin_two_field_plan_vs_oracle(dev, varian):
    """computeIMRTPlan (drop-in, host arrays) with two fields on a 60^3 @ 3 mm
    cube vs the full-volume oracle: 1e-3 of D_max, gamma 2%/2 mm >= 0.995."""
    from DoseCUDA import IMRTDoseGrid, IMRTPlan
    _say("drop-in two-field plan, 60^3 @ 3 mm")
    g = IMRTDoseGrid()
    g.createCubePhantom(size=[60, 60, 60], spacing=3.0)
    plan = IMRTPlan("VarianTrueBeamHF")
    plan.n_fractions = 2
    plan.addSquareField(dimx=6, dimy=5)
    plan.addSquareField('10', 4.0, 7.0, mu=50, gantry_angle=90.0, collimator_angle=20.0)
    g.computeIMRTPlan(plan)
    ref, ref_beams = OR.compute_plan(g, plan, os.path.join(PHOT, "VarianTrueBeamHF"))
    m = O.compare(g.dose, ref)
    gam = O.gamma_pass_rate(g.dose, ref, 3.0)
    _say("  max|dD|/Dmax %.2e, gamma %.4f" % (m["max_rel_to_dmax"], gam))
    assert m["max_rel_to_dmax"] < 1e-3 and gam >= 0.995, (m, gam)
    for b, rb in zip(g.beam_doses, ref_beams):
        assert np.max(np.abs(b - rb)) < 1e-3 * rb.max()


def test_reference_example_script_sampled(dev, varian, tmp_path):
    """The reference's own IMRT example (tests/test_phantom_imrt.py): default
    cube 138^3 @ 3 mm, a 10x10 field and a 15x15 field (50 MU, gantry 90,
    collimator 20), PHYSICAL dose written as NRRD.  Oracle at 3000 voxels
    (half above 10% of D_max): 1e-3 of D_max; the file equals the array."""
    from DoseCUDA import IMRTDoseGrid, IMRTPlan
    from DoseCUDA.imageio import read_nrrd
    _say("reference IMRT example, 138^3 @ 3 mm")
    dose = IMRTDoseGrid()
    plan = IMRTPlan("VarianTrueBeamHF")
    dose.createCubePhantom()
    plan.addSquareField()
    plan.addSquareField('6', 15.0, 15.0, mu=50, gantry_angle=90.0, collimator_angle=20.0)
    dose.computeIMRTPlan(plan)
    dose.writeDoseNRRD(str(tmp_path / "cube_imrt_dose.nrrd"), individual_beams=False, dose_type="PHYSICAL")
    d, _, _ = read_nrrd(str(tmp_path / "cube_imrt_dose.nrrd"))
    assert np.array_equal(d, dose.dose)

    rng = np.random.default_rng(11)
    flat = dose.dose.ravel()
    hi = np.flatnonzero(flat > 0.1 * flat.max())
    idx = np.unique(np.concatenate([rng.choice(hi, 1500, replace=False), rng.choice(flat.size, 1500, replace=False),
                                    [int(flat.argmax())]]))
    rho = OR.density_from_hu(dose.HU, os.path.join(PHOT, "VarianTrueBeamHF"))
    s = float(np.float32(dose.spacing[0]))
    ref = np.zeros(len(idx))
    skipped = 0
    for beam in plan.beam_list:
        model = plan.beam_models[list(np.asarray(plan.dicom_energy_label).astype(str)).index(beam.dicom_energy_label)]
        for cp in beam.cp_list:
            geom = _geom(dose, model, cp)
            depth = OR.raytrace(rho, s, geom)
            te, el = OR.terma(depth.astype(np.float32).astype(np.float64), rho.shape, s, geom, model,
                              OR.mlc_rows(cp.mlc))
            t32 = te.astype(np.float32)
            near = np.abs(te.ravel()[idx] - 0.01) < 1e-4
            skipped += int(near.sum())
            ref += OR.ccc(t32, rho, el.astype(np.float32).ravel()[idx].astype(np.float64), s, geom, model.kernel,
                          np.float32(model.mu_calibration), np.float32(cp.mu), idx=idx) * OR.output_factor(model, cp)
            idx_keep = ~near
            idx, ref = idx[idx_keep], ref[idx_keep]
    got = flat[idx]
    err = np.max(np.abs(got - ref)) / ref.max()
    _say("  %d voxels, %d near the TERMA threshold left out, max|dD|/Dmax %.2e" % (len(idx), skipped, err))
    assert skipped < 20 and err < 1e-3, err


def test_engine_matches_dropin_and_is_deterministic(dev, varian):
    """PhotonDoseEngine (density from HU on the device, beams accumulated on
    the device) equals the drop-in's plan dose to float rounding; two runs
    are bit-identical, and so are the pipelined (two workspaces, side stream)
    and one-stream schedules."""
    from DoseCUDA import IMRTDoseGrid, IMRTPlan
    from DoseCUDA.photon import PhotonDoseEngine
    from DoseCUDA.plan_imrt import IMRTBeam
    _say("engine vs drop-in")
    g = _hetero_grid((30, 34, 32), 2.5, seed=5)
    plan = IMRTPlan("VarianTrueBeamHF")
    beam = IMRTBeam()
    beam.dicom_energy_label = '6'
    for k, ga in enumerate((10.0, 12.0, 14.0)):   # an arc-like beam: three segments
        beam.addControlPoint(_cp(plan, 0, ga, 5.0, 15.0, 15.0 + 5 * k, 20.0, mu=30.0 + k, irregular=True))
    plan.addBeam(beam)
    plan.addSquareField('10', 5.0, 5.0, mu=40, gantry_angle=250.0, collimator_angle=45.0, table_angle=10.0)
    g.computeIMRTPlan(plan)
    eng = PhotonDoseEngine(g, plan, gpu_id=0)
    d1, b1 = eng.compute()
    d1 = d1.cpu().numpy().copy()
    b1 = [b.cpu().numpy().copy() for b in b1]
    d2 = eng.compute()[0].cpu().numpy().copy()
    assert np.array_equal(d1, d2)
    # one stream, one workspace: the same kernels in the same order, bit for bit
    d3, b3 = eng.compute(pipeline=False)
    assert np.array_equal(d1, d3.cpu().numpy())
    assert all(np.array_equal(x, y.cpu().numpy()) for x, y in zip(b1, b3))
    seq = PhotonDoseEngine(g, plan, gpu_id=0, pipeline=False)
    assert np.array_equal(d1, seq.compute()[0].cpu().numpy())
    assert np.max(np.abs(d1 - g.dose)) < 1e-5 * g.dose.max()
    for b, hb in zip(b1, g.beam_doses):
        assert np.max(np.abs(b - hb)) < 1e-5 * hb.max()


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 3, 4), (5, 1, 7)])
def test_degenerate_grids(dev, varian, shape):
    _say("degenerate grid %s" % (shape,))
    g = _hetero_grid((8, 8, 8), 3.0, seed=2)
    g.HU = np.random.default_rng(3).uniform(-200, 300, size=shape).astype(np.float32)
    g.size = np.array(shape)
    model = varian.beam_models[0]
    cp = _cp(varian, 0, 30.0, 10.0, 5.0, 50.0, 50.0)
    rho, depth, terma, el, dose = _stages(g, model, cp, dev)
    geom = _geom(g, model, cp)
    s = float(np.float32(g.spacing[0]))
    assert np.max(np.abs(depth - OR.raytrace(rho, s, geom))) < 2e-4
    dose_ref = OR.ccc(terma, rho, el.astype(np.float64), s, geom, model.kernel, np.float32(model.mu_calibration),
                      np.float32(cp.mu))
    assert np.max(np.abs(dose - dose_ref)) <= 1e-4 * max(dose_ref.max(), 1e-30)


def test_closed_leaves_give_no_dose(dev, varian):
    """Every leaf pair closed: zero transmission, zero TERMA and zero dose
    (the 0/0 ratio of the two transmissions reads as 1, as fminf gives)."""
    _say("closed leaves")
    g = _hetero_grid((10, 12, 11), 3.0, seed=4)
    model = varian.beam_models[0]
    cp = _cp(varian, 0, 0.0, 0.0, 0.0, 0.0, 0.0)
    rho, depth, terma, el, dose = _stages(g, model, cp, dev)
    assert np.all(terma == 0) and np.all(el == 0) and np.all(dose == 0)


def test_dropin_argument_errors(dev, varian):
    """photon_dose_cuda raises the reference's exceptions (dosemodule.cu:50-66)."""
    from DoseCUDA import dose_kernels
    from DoseCUDA.plan import VolumeObject
    _say("drop-in argument errors")
    vol = VolumeObject()
    vol.voxel_data = np.zeros((4, 4, 4), np.float64)
    vol.spacing = np.array([2.0] * 3, np.float32)
    vol.origin = np.zeros(3, np.float32)
    cp = _cp(varian, 0, 0.0, 0.0, 0.0, 10.0, 10.0)
    with pytest.raises(ValueError, match="'voxel_data' must be 3-dimensional and of type float."):
        dose_kernels.photon_dose_cuda(varian.beam_models[0], vol, cp, 0)
    vol.voxel_data = np.ones((4, 4, 4), np.float32)

    class NoMu:
        pass
    bad = NoMu()
    for k, v in vars(cp).items():
        if k != "mu":
            setattr(bad, k, v)
    with pytest.raises(AttributeError):
        dose_kernels.photon_dose_cuda(varian.beam_models[0], vol, bad, 0)
    out = dose_kernels.photon_dose_cuda(varian.beam_models[0], vol, cp, 0)
    assert out.shape == (4, 4, 4) and out.dtype == np.float32


def test_body_grid_sampled(dev, varian):
    """The body-size grid (300x512x512 @ 1 mm: lungs, spine; 3.8 GB of
    records), one 10 x 12 cm 10 MV field at gantry 90, collimator 30, through
    PhotonDoseEngine; 2000 sampled voxels (half above 10% of D_max): depth and
    TERMA vs the oracle at them, the dose vs the oracle's cone sum on the
    GPU's TERMA volume."""
    from DoseCUDA import IMRTDoseGrid, IMRTPlan
    from DoseCUDA import photon as PH
    from DoseCUDA.workloads import body_plan
    _say("body grid 300x512x512 @ 1 mm")
    bg, _, _ = body_plan()
    g = IMRTDoseGrid()
    g.spacing, g.size, g.origin, g.HU = bg.spacing, bg.size, bg.origin, bg.HU
    del bg
    plan = IMRTPlan("VarianTrueBeamHF")
    plan.addSquareField('10', 10.0, 12.0, mu=100, gantry_angle=90.0, collimator_angle=30.0)
    eng = PH.PhotonDoseEngine(g, plan, gpu_id=0)
    dose = eng.compute()[0].cpu().numpy().ravel()
    dmodel, cps = eng.beams[0]
    cpc, of, _ = cps[0]
    depth = eng.ws.depth.cpu().numpy()
    terma = eng.ws.terma.cpu().numpy()
    el = eng.ws.electron.cpu().numpy().ravel()
    rng = np.random.default_rng(5)
    hi = np.flatnonzero(dose > 0.1 * dose.max())
    idx = np.unique(np.concatenate([rng.choice(hi, 1000, replace=False), rng.choice(dose.size, 1000, replace=False),
                                    [int(dose.argmax())]]))
    model = plan.beam_models[list(np.asarray(plan.dicom_energy_label).astype(str)).index('10')]
    cp = plan.beam_list[0].cp_list[0]
    geom = _geom(g, model, cp)
    rho = OR.density_from_hu(g.HU, os.path.join(PHOT, "VarianTrueBeamHF"))
    d_ref = OR.raytrace(rho, 1.0, geom, idx=idx)
    assert np.max(np.abs(depth.ravel()[idx] - d_ref)) < 2e-4, np.max(np.abs(depth.ravel()[idx] - d_ref))
    t_ref, e_ref = OR.terma(depth.ravel()[idx].astype(np.float64), rho.shape, 1.0, geom, model, OR.mlc_rows(cp.mlc),
                            idx=idx)
    assert np.max(np.abs(terma.ravel()[idx] - t_ref)) < 1e-5 * terma.max()
    ref = OR.ccc(terma, rho, el[idx].astype(np.float64), 1.0, geom, model.kernel, np.float32(model.mu_calibration),
                 np.float32(cp.mu), idx=idx) * of
    err = np.max(np.abs(dose[idx] - ref)) / dose.max()
    _say("  %d voxels: depth %.2e, TERMA %.2e, dose %.2e of D_max" % (
        len(idx), np.max(np.abs(depth.ravel()[idx] - d_ref)), np.max(np.abs(terma.ravel()[idx] - t_ref)) / terma.max(),
        err))
    assert err < 1e-4, err
