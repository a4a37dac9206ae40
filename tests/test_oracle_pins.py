"""Pins of the FP64 oracle against sources independent of its own code.

* LUT tables: values read straight from the machine CSV files.
* Geometry: hand-computed rotations at 0/90 deg, the source maps to (0, 0, SAD)
  in the beam's-eye view for every gantry/couch, head <-> image inverse.
* Ray-march: closed-form WET along an axis-aligned ray in water and through a
  slab; brute force with scipy.ndimage.map_coordinates('grid-constant').
* Smoothing: pure-Python brute force; invariants (uniform field unchanged,
  WET <= 0.1 cm takes no neighbours).
* Pencil beam: pure-Python brute force using np.interp (LUT) and a cross
  product (axis distance); planar normalisation sum(D s^2) = mu * IDD;
  linearity in the weights; mirror symmetry.
* Gradient: central finite differences of the oracle dose; torch float64
  autograd of an independent vectorised restatement.

The halo fit (IMPTClasses.cu:104-121) has no source independent of its
transcription, so its coefficients stay unpinned; what the mathematics fixes
is checked: its clamps (0 <= weight <= 0.9, sigma >= 0), that it is finite and
continuous over the whole clamp range [0.1, R80 - 0.1] of all 196 energies
(the poles at w = R80 + 5 and w^2 = (R80 + 3)^2 lie outside it, so a sign
slip that moves one inside fails), the plateaus outside the range, the
collapse for R80 < 0.2 mm onto the continuation of the range end, and the
planar normalisation (1 - u) + u = 1 at depths where the halo weight u > 0.
"""

import math

import numpy as np
import pytest

from oracle import impt_oracle as O

from bruteforce import pb_bruteforce as _pb_bruteforce
from bruteforce import raymarch_bruteforce as _raymarch_bruteforce
from bruteforce import smooth_bruteforce as _smooth_bruteforce


# ---------------------------------------------------------------------------
# lookup tables
def test_lut_values_from_csv(plan_models, machine_dir):
    rs0, rs4 = plan_models.beam_models
    assert (rs0.VSADX, rs0.VSADY) == (1340.0, 1930.0)
    np.testing.assert_array_equal(rs0.divergence_params[0],
                                  np.array([70.2, 39.27, 2.961E-06, -1.816E-03, 1.205E+00], np.float32))
    assert rs0.divergence_params.shape == (98, 5) and rs0.lut_depths.shape == (98, 400)
    np.testing.assert_array_equal(rs0.lut_depths, np.tile(np.arange(400, dtype=np.float32), (98, 1)))
    assert rs0.lut_sigmas[0, 0] == np.float32(-0.27) and rs0.lut_idds[0, 0] == np.float32(64.26)
    assert rs0.lut_idds[0, 3] == np.float32(66.88)
    # last kept row is depth 399 (row at 400 mm dropped)
    assert rs0.lut_depths[0, -1] == 399.0
    assert rs4.divergence_params[0, 1] == np.float32(-0.65)
    assert list(plan_models.dicom_rangeshifter_label.astype(str)) == ["0", "4"]


def test_rlsp_knots(plan_models):
    from DoseCUDA import IMPTDoseGrid
    g = IMPTDoseGrid()
    g.HU = np.array([-2000.0, -1000.0, 0.0, 24.0, 3096.0, 5000.0], np.float32)
    r = g.RLSPFromHU("HitachiProbeatJHU")
    np.testing.assert_array_equal(r, np.array([0.00113, 0.00113, 1.00653, 1.02058, 2.36825, 2.36825], np.float32))
    g.HU = np.array([-871.25], np.float32)  # midway between the first two knots
    assert abs(float(g.RLSPFromHU("HitachiProbeatJHU")[0]) - (0.00113 + 0.25771) / 2) < 1e-6


# ---------------------------------------------------------------------------
# geometry
def _geom(gantry, couch, iso=(0.0, 0.0, 0.0), origin=(0.0, 0.0, 0.0)):
    return O.beam_geometry(np.array(iso), np.array(origin), gantry, couch, 1340.0, 1930.0)


def test_geometry_closed_forms():
    g0 = _geom(0.0, 0.0)
    np.testing.assert_array_equal(g0[3:6], [0.0, -1635.0, 0.0])
    p = np.array([3.0, -7.0, 11.0])
    np.testing.assert_allclose(O.image_to_head(g0, p), [3.0, 11.0, 7.0])      # (x, z, -y)
    g90 = _geom(90.0, 0.0)
    np.testing.assert_array_equal(g90[3:6], [1635.0, 0.0, 0.0])
    np.testing.assert_allclose(O.image_to_head(g90, p), [-7.0, 11.0, 3.0])    # (y, z, x)


@pytest.mark.parametrize("gantry,couch", [(0, 0), (37.5, 0), (120, 0), (240, 30), (300, -45), (90, 90)])
def test_source_on_beam_axis(gantry, couch):
    g = _geom(gantry, couch)
    h = O.image_to_head(g, g[3:6])
    np.testing.assert_allclose(h, [0.0, 0.0, 1635.0], atol=1e-9)
    p = np.array([12.5, -40.0, 7.25])
    hp = O.image_to_head(g, p)
    assert abs(np.linalg.norm(hp) - np.linalg.norm(p)) < 1e-9


# ---------------------------------------------------------------------------
# ray-march
def _centered_grid(n, s):
    """iso at the exact centre voxel of an odd-sized grid."""
    return np.array([(n[2] - 1) * s / 2, (n[1] - 1) * s / 2, (n[0] - 1) * s / 2])


def test_raymarch_axis_closed_form():
    n = (9, 21, 9)
    s = 2.5
    rho = np.full(n, 1.00653, np.float32)
    iso_adj = _centered_grid(n, s)
    g = O.beam_geometry(np.zeros(3), -iso_adj, 0.0, 0.0, 1340.0, 1930.0)
    raw = O.raymarch(rho, s, g)
    for j in range(n[1]):
        # samples n = 0..floor(2.5 j) along -y, all of value rho (no x/z blend)
        count = math.floor(s * j) + 1
        assert abs(raw[4, j, 4] - (-0.05 + 0.1 * np.float64(np.float32(1.00653)) * count)) < 1e-12


def test_raymarch_slab_closed_form():
    n = (5, 40, 5)
    s = 2.0
    rho = np.full(n, 1.0, np.float32)
    rho[:, 10:15, :] = 1.75
    iso_adj = _centered_grid(n, s)
    g = O.beam_geometry(np.zeros(3), -iso_adj, 0.0, 0.0, 1340.0, 1930.0)
    raw = O.raymarch(rho, s, g)
    col = rho[2, :, 2].astype(np.float64)
    for j in range(n[1]):
        c = j - np.arange(0, int(s * j) + 1) / s           # continuous y index of each sample
        lo = np.floor(c).astype(int)
        f = c - lo
        hi = np.minimum(lo + 1, n[1] - 1)
        val = col[lo] * (1 - f) + col[hi] * f
        assert abs(raw[2, j, 2] - (-0.05 + 0.1 * val.sum())) < 1e-12


@pytest.mark.parametrize("gantry,couch", [(120.0, 30.0), (37.0, 0.0)])
def test_raymarch_bruteforce(gantry, couch):
    rng = np.random.default_rng(5)
    rho = rng.uniform(-0.2, 2.0, size=(6, 7, 8)).astype(np.float32)
    s = 3.0
    g = O.beam_geometry(np.array([1.0, -2.0, 0.5]), np.array([-10.0, -12.0, -8.0]), gantry, couch, 1340.0, 1930.0)
    raw = O.raymarch(rho, s, g)
    for vox in [(0, 0, 0), (5, 6, 7), (3, 2, 4), (1, 6, 0), (4, 0, 5)]:
        assert abs(raw[vox] - _raymarch_bruteforce(rho, s, g, vox)) < 1e-12


# ---------------------------------------------------------------------------
# smoothing
@pytest.mark.parametrize("gantry,couch", [(0.0, 0.0), (120.0, 30.0), (120.0, 0.0)])
def test_smooth_bruteforce(gantry, couch):
    rng = np.random.default_rng(9)
    raw = rng.uniform(-0.05, 2.0, size=(5, 6, 7))
    s = 2.0
    g = O.beam_geometry(np.zeros(3), np.array([-6.0, -5.0, -4.0]), gantry, couch, 1340.0, 1930.0)
    np.testing.assert_allclose(O.smooth(raw, s, g), _smooth_bruteforce(raw, s, g), rtol=0, atol=1e-13)


def test_smooth_half_voxel_ties_go_even():
    """Gantry 120 (BEV x at 300 deg), 2.5 mm voxels: direction 60 deg at
    dr = 5 mm moves exactly half a voxel along image x; the neighbour must be
    the even index whatever the float rounding of the rotation."""
    s = 2.5
    raw = np.zeros((3, 3, 8))
    raw[1, 1, :] = np.arange(8) * 1.0 + 100.0     # distinct values along x
    g = O.beam_geometry(np.zeros(3), np.array([-3.75 * 2, -2.5, -2.5]), 120.0, 0.0, 1340.0, 1930.0)
    # direct check of the exact image offset of (a=1, dr=5): x = -2.5*... voxel units
    hx, hy = 5 * 0.5, 5 * math.sqrt(3) / 2
    sg, cg, sc, cc = g[6:10]
    off_x = (-hx * cg) * cc + hy * sc
    assert abs(off_x / s - round(off_x / s * 2) / 2) < 1e-12 and abs(abs(off_x / s) % 1 - 0.5) < 1e-12
    ref = _smooth_bruteforce(raw, s, g)
    np.testing.assert_allclose(O.smooth(raw, s, g), ref, rtol=0, atol=1e-13)


def test_smooth_invariants():
    s = 2.0
    g = O.beam_geometry(np.zeros(3), np.array([-8.0, -9.0, -7.0]), 120.0, 0.0, 1340.0, 1930.0)
    flat = np.full((8, 9, 10), 3.7)
    np.testing.assert_allclose(O.smooth(flat, s, g), flat, rtol=1e-15)
    thin = np.full((8, 9, 10), 0.09)
    thin[4, 4, 4] = 0.1
    out = O.smooth(thin * np.random.default_rng(1).uniform(0.5, 1.0, thin.shape), s, g)
    np.testing.assert_array_equal(out, thin * np.random.default_rng(1).uniform(0.5, 1.0, thin.shape))


# ---------------------------------------------------------------------------
# pencil beam
def _small_pb_case(plan_models, gantry=120.0, couch=30.0, shape=(4, 5, 6), s=3.0):
    model = plan_models.beam_models[0]
    marr = O.ModelArrays(model)
    rng = np.random.default_rng(2)
    wet = rng.uniform(0.0, 9.0, size=shape)
    g = O.beam_geometry(np.array([0.5, 0.0, -1.0]), np.array([-8.0, -7.0, -6.0]), gantry, couch,
                        model.VSADX, model.VSADY)
    sl = np.array([[1.0, -2.0, 0.7, 30], [-3.0, 4.0, 1.3, 30], [0.0, 0.0, 0.5, 12], [6.0, 1.0, 0.9, 12],
                   [2.0, 2.0, 1.1, 45]], np.float32)
    order, spots, layers, r80 = O.spots_and_layers(sl, model.divergence_params)
    return marr, wet, g, sl, spots, layers, r80, s


def test_pb_bruteforce(plan_models):
    marr, wet, g, sl, spots, layers, r80, s = _small_pb_case(plan_models)
    ref = _pb_bruteforce(wet, s, g, spots, layers, r80, marr)
    out = O.pb_dose(wet, s, g, spots, layers, r80, marr)
    assert ref.max() > 0
    np.testing.assert_allclose(out, ref, rtol=1e-11, atol=1e-12 * ref.max())


def test_pb_planar_normalisation(plan_models):
    """One spot on the central axis, WET constant per depth plane: the dose
    integrated over a plane is mu * IDD(wet) (the two Gaussians carry weights
    (1 - u) and u and are normalised)."""
    model = plan_models.beam_models[0]
    marr = O.ModelArrays(model)
    n = (121, 3, 121)
    s = 1.0
    wet = np.empty(n)
    depths_cm = np.array([2.0, 5.5, 8.0])
    wet[:] = depths_cm[None, :, None]
    iso_adj = _centered_grid(n, s)
    g = O.beam_geometry(np.zeros(3), -iso_adj, 0.0, 0.0, model.VSADX, model.VSADY)
    eid = 40
    sl = np.array([[0.0, 0.0, 1.7, eid], [0.0, 0.0, 0.0, eid]], np.float32)
    _, spots, layers, r80 = O.spots_and_layers(sl, model.divergence_params)
    dose = O.pb_dose(wet, s, g, spots, layers, r80, marr)
    for j, dcm in enumerate(depths_cm):
        idd = np.interp(dcm * 10.0, marr.depths[eid], marr.idds[eid])
        assert abs(dose[:, j, :].sum() * s * s / (1.7 * idd) - 1.0) < 2e-4


def test_pb_linearity_and_mirror(plan_models):
    model = plan_models.beam_models[0]
    marr = O.ModelArrays(model)
    n = (9, 7, 11)
    s = 2.0
    rng = np.random.default_rng(4)
    half = rng.uniform(0.0, 8.0, size=(9, 7, 6))
    wet = np.concatenate([half, half[:, :, -2::-1]], axis=2)  # mirror-symmetric in x
    assert wet.shape == n
    iso_adj = _centered_grid(n, s)
    g = O.beam_geometry(np.zeros(3), -iso_adj, 0.0, 0.0, model.VSADX, model.VSADY)

    def dose(rows):
        _, spots, layers, r80 = O.spots_and_layers(np.array(rows, np.float32), model.divergence_params)
        return O.pb_dose(wet, s, g, spots, layers, r80, marr)

    a = [[4.0, 1.0, 0.8, 33]]
    b = [[-2.0, -3.0, 1.4, 50]]
    np.testing.assert_allclose(dose(a + b), dose(a) + dose(b), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(dose([[4.0, 1.0, 1.6, 33]]), 2.0 * dose(a), rtol=1e-13)
    # gantry 0: head x = image x, so a spot at -x gives the x-mirrored dose
    np.testing.assert_allclose(dose([[-4.0, 1.0, 0.8, 33]]), dose(a)[:, :, ::-1], rtol=1e-11, atol=1e-14)


def test_pb_halo_clamps(plan_models):
    """Halo weight in [0, 0.9] and sigma >= 0 for every (wet, R80) of both
    range shifters: the dose of a spot is bounded by mu * IDD / (2 pi min
    sigma^2) and never negative."""
    for model in plan_models.beam_models:
        marr = O.ModelArrays(model)
        wet = np.linspace(0, 40.0, 64).reshape(4, 4, 4)
        g = O.beam_geometry(np.zeros(3), np.array([-3.0, -3.0, -3.0]), 0.0, 0.0, model.VSADX, model.VSADY)
        for eid in (0, 30, 97):
            sl = np.array([[0.0, 0.0, 1.0, eid], [1.0, 0.0, 1.0, eid]], np.float32)
            _, spots, layers, r80 = O.spots_and_layers(sl, model.divergence_params)
            d = O.pb_dose(wet, 2.0, g, spots, layers, r80, marr)
            assert np.all(np.isfinite(d)) and np.all(d >= 0)


def _energy_ranges(plan_models):
    for model in plan_models.beam_models:
        for e, r80 in enumerate(model.divergence_params[:, 1].astype(np.float64)):
            yield model, e, float(r80)


def test_halo_fit_finite_and_continuous_on_clamp_range(plan_models):
    """On [0.1, R80 - 0.1] both fits are finite and bounded, and their largest
    step between neighbouring depths shrinks in proportion to the grid step
    (Lipschitz): a pole inside the range (e.g. (w - R80 + 5)^2 for
    (w - R80 - 5)^2) blows the values up or stops the steps from shrinking."""
    n_checked = 0
    for _, e, r80 in _energy_ranges(plan_models):
        if r80 < 0.2:
            continue
        n_checked += 1
        coarse = np.linspace(0.1, r80 - 0.1, 401)
        fine = np.linspace(0.1, r80 - 0.1, 4001)
        sn_c, un_c = O.halo(r80, coarse)
        sn_f, un_f = O.halo(r80, fine)
        for v in (sn_c, un_c, sn_f, un_f):
            assert np.all(np.isfinite(v))
        assert sn_f.min() >= 0.0 and sn_f.max() < 40.0          # halo widths of millimetres to centimetres
        assert un_f.min() >= 0.0 and un_f.max() <= 0.9
        for c, f in ((sn_c, sn_f), (un_c, un_f)):
            jc = np.abs(np.diff(c)).max()
            jf = np.abs(np.diff(f)).max()
            assert jf <= 0.15 * jc + 1e-12, (e, r80, jc, jf)   # step ratio 10: jumps shrink ~10x
    assert n_checked == 195                                    # every energy but RS4's R80 = -0.65 mm


def test_halo_fit_clamp_plateaus(plan_models):
    """Depths below 0.1 mm give the value at 0.1 mm and depths beyond R80 - 0.1
    the value at R80 - 0.1, exactly (the depth is clamped before the fit)."""
    for _, e, r80 in _energy_ranges(plan_models):
        if r80 < 0.2:
            continue
        lo = np.array([-5.0, 0.0, 0.05, 0.1])
        hi = np.array([r80 - 0.1, r80, 1.1 * r80, 2.0 * r80 + 50.0])
        for w in (lo, hi):
            sn, un = O.halo(r80, w)
            assert np.all(sn == sn[-1 if w is lo else 0]) and np.all(un == un[-1 if w is lo else 0])


def test_halo_fit_collapsed_range(plan_models):
    """RS4 energy 0 has R80 = -0.65 mm: the clamp [0.1, R80 - 0.1] is empty and
    the depth collapses to R80 - 0.1 (fmin(fmax(w, 0.1), R80 - 0.1), as
    jnp.clip): the fit is constant in the depth and equals the continuation of
    its value at the end of the range, h(R) = fit(R - 0.1; R), extrapolated
    smoothly from R in [0.2, 3] mm (the other clamp order would give
    fit(0.1; R), a halo width about 0.05 mm larger)."""
    rs4 = plan_models.beam_models[1]
    r80 = float(rs4.divergence_params[0, 1])
    assert r80 < 0.0
    sn, un = O.halo(r80, np.linspace(-5.0, 400.0, 50))
    assert np.all(sn == sn[0]) and np.all(un == un[0]) and np.isfinite(sn[0]) and np.isfinite(un[0])
    assert 0.0 <= un[0] <= 0.9
    # sigma_n stays positive near the range end, so its continuation is smooth
    # (u_n sits on its clamp at 0 there, a kink, and is only range-checked)
    rr = np.linspace(0.2, 3.0, 40)
    end = np.array([O.halo(R, [R - 0.1])[0][0] for R in rr])      # sigma_n at the range end
    start = np.array([O.halo(R, [0.1])[0][0] for R in rr])        # the other order's continuation
    assert end.min() > 0.5
    want = np.polyval(np.polyfit(rr, end, 5), r80)
    other = np.polyval(np.polyfit(rr, start, 5), r80)
    assert abs(want - sn[0]) < 2e-5, (want, sn[0])
    assert abs(other - sn[0]) > 1e-2, (other, sn[0])               # the check tells the two orders apart


@pytest.mark.parametrize("mi,eid,depths_cm", [(0, 10, (2.0, 4.5)), (0, 40, (5.0, 9.0)), (0, 80, (6.0, 15.0)),
                                              (1, 30, (1.5, 3.0))])
def test_pb_planar_normalisation_with_halo(plan_models, mi, eid, depths_cm):
    """At depths where the halo weight u > 0 the primary and halo Gaussians
    carry (1 - u) and u of the integral depth dose: one spot, WET constant
    per beam's-eye-view plane, plane sum of D s^2 = mu * IDD(wet).  The
    energy's sigma_air is made constant (A = B = 0, C = 4 mm above the row's
    most negative sigma_ms): with the real coefficients sigma_air follows each
    voxel's distance to the source, which grows off axis and widens the
    Gaussians slightly (+3.9e-4 of the plane sum with the range shifter)."""
    model = plan_models.beam_models[mi]
    marr = O.ModelArrays(model)
    marr.dvp = marr.dvp.copy()
    marr.dvp[eid, 2:5] = (0.0, 0.0, 4.0 - min(0.0, float(marr.sigmas[eid].min())))
    r80 = float(model.divergence_params[eid, 1])
    _, un = O.halo(r80, np.array(depths_cm) * 10.0)
    assert np.all(un > 0.02), un
    n = (181, len(depths_cm), 181)
    s = 1.0
    wet = np.empty(n)
    wet[:] = np.array(depths_cm)[None, :, None]
    iso_adj = _centered_grid(n, s)
    g = O.beam_geometry(np.zeros(3), -iso_adj, 0.0, 0.0, model.VSADX, model.VSADY)
    sl = np.array([[0.0, 0.0, 1.3, eid], [0.0, 0.0, 0.0, eid]], np.float32)
    _, spots, layers, r80s = O.spots_and_layers(sl, model.divergence_params)
    dose = O.pb_dose(wet, s, g, spots, layers, r80s, marr)
    for j, dcm in enumerate(depths_cm):
        idd = np.interp(dcm * 10.0, marr.depths[eid], marr.idds[eid])
        assert abs(dose[:, j, :].sum() * s * s / (1.3 * idd) - 1.0) < 2e-5


# ---------------------------------------------------------------------------
# layer table (importLayers semantics)
def test_layer_table_quirks(plan_models):
    dvp = plan_models.beam_models[0].divergence_params
    sl = np.array([[0, 0, 1, 5.9], [1, 1, 1, 3], [2, 2, 1, 5], [3, 3, 1, 120]], np.float32)
    order, spots, layers, r80 = O.spots_and_layers(sl, dvp)
    assert layers.tolist() == [[0, 1, 3], [1, 2, 5]]           # 5.9 truncates to 5; 120 >= n_E dropped
    np.testing.assert_array_equal(r80, dvp[[3, 5], 1].astype(np.float64))
    sl2 = np.vstack([sl, [[9, 9, 1, -1]]]).astype(np.float32)
    assert O.spots_and_layers(sl2, dvp)[2].shape == (0, 3)     # a negative id blocks every layer


# ---------------------------------------------------------------------------
# gradient
def test_grad_central_fd(plan_models):
    marr, wet, g, sl, spots, layers, r80, s = _small_pb_case(plan_models, shape=(5, 6, 7), s=2.0)
    rng = np.random.default_rng(8)
    gvol = rng.normal(size=wet.shape)
    gmu, gxy = O.pb_grad(wet, gvol, s, g, spots, layers, r80, marr)

    def loss(sp):
        return float((gvol * O.pb_dose(wet, s, g, sp, layers, r80, marr)).sum())

    for q in range(spots.shape[0]):
        hw = 1e-3
        up, dn = spots.copy(), spots.copy()
        up[q, 2] += hw
        dn[q, 2] -= hw
        fd = (loss(up) - loss(dn)) / (2 * hw)
        assert abs(fd - gmu[q]) <= 1e-7 * max(1.0, abs(gmu[q]))
        for a in (0, 1):
            hx = 0.05
            up, dn = spots.copy(), spots.copy()
            up[q, a] += hx
            dn[q, a] -= hx
            fd = (loss(up) - loss(dn)) / (2 * hx)
            scale = max(np.abs(gxy).max(), 1e-12)
            assert abs(fd - gxy[q, a]) <= 2e-4 * scale


def test_grad_torch_autograd(plan_models):
    torch = pytest.importorskip("torch")
    marr, wet, g, sl, spots, layers, r80, s = _small_pb_case(plan_models, shape=(5, 6, 7), s=2.0)
    rng = np.random.default_rng(12)
    gvol = rng.normal(size=wet.shape)
    gmu, gxy = O.pb_grad(wet, gvol, s, g, spots, layers, r80, marr)

    # independent vectorised float64 restatement, differentiated by torch
    nz, ny, nx = wet.shape
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    P = np.stack([I * s, J * s, K * s], -1).reshape(-1, 3) - g[0:3]
    H = np.stack([O.image_to_head(g, p) for p in P])
    dsrc = np.linalg.norm(g[3:6] - P, axis=1)
    w_mm = wet.reshape(-1) * 10.0
    xy = torch.tensor(spots[:, :2], dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(spots[:, 2], dtype=torch.float64, requires_grad=True)
    Ht = torch.tensor(H)
    dose = torch.zeros(len(P), dtype=torch.float64)
    for (st, ns, eid), R in zip(layers, r80):
        dep = marr.depths[eid].astype(np.float64)
        idd = torch.tensor(np.interp(w_mm, dep, marr.idds[eid]))
        sms = np.interp(w_mm, dep, marr.sigmas[eid])
        A, B, C = (float(c) for c in marr.dvp[eid, 2:5])
        d = dsrc - w_mm + 0.7 * R
        sig_t = torch.tensor(A * d * d + B * d + C + sms)
        w = np.minimum(np.maximum(w_mm, 0.1), R - 0.1)
        sn = np.maximum(2.85 + 0.0014 * R * np.log(w + 3) + 0.06 * w - 7.4e-5 * w * w - 0.22 * R / (w - R - 5) ** 2,
                        0)
        un = np.clip(0.052 * np.log(1.13 + w / (11.2 - 0.023 * R))
                     + 0.35 * (0.0017 * R * R - R) / ((R + 3) ** 2 - w * w) - 1.61e-9 * w * (R + 3) ** 2, 0, 0.9)
        un = torch.tensor(un)
        sig_h = torch.sqrt(sig_t ** 2 + torch.tensor(sn) ** 2)
        mask = torch.tensor(w_mm <= 1.1 * R, dtype=torch.float64)
        for q in range(st, st + ns):
            t = torch.stack([xy[q, 0] / g[10], xy[q, 1] / g[11], torch.tensor(-1.0, dtype=torch.float64)])
            r = Ht - torch.stack([xy[q, 0], xy[q, 1], torch.tensor(0.0, dtype=torch.float64)])
            cr = torch.linalg.cross(r, t.expand_as(r))
            d2 = (cr * cr).sum(1) / (t * t).sum()
            pd = (1 - un) * idd / (2 * math.pi * sig_t ** 2 + 1e-8) * torch.exp(-d2 / (2 * sig_t ** 2))
            hd = un * idd / (2 * math.pi * sig_h ** 2 + 1e-8) * torch.exp(-d2 / (2 * sig_h ** 2))
            dose = dose + mask * mu[q] * (pd + hd)
    loss = (torch.tensor(gvol.reshape(-1)) * dose).sum()
    loss.backward()
    np.testing.assert_allclose(mu.grad.numpy(), gmu, rtol=1e-9, atol=1e-12 * np.abs(gmu).max())
    np.testing.assert_allclose(xy.grad.numpy(), gxy, rtol=1e-9, atol=1e-12 * np.abs(gxy).max())
