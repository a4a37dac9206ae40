"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the FP64 oracle) and the checks that pin them without the oracle's code.

* rlsp: np.interp over the HU_RLSP.csv knots read with the csv module.
* wet_raw: scipy map_coordinates(order=1, mode='grid-constant') brute-force
  march at sampled voxels; wet: the pure-Python smoothing brute force.
* dose: the pure-Python pencil-beam brute force (np.interp LUT, cross-product
  axis distance) over the whole grid.
* grad_mu: dose linearity, dL/dmu_s = sum(grad_field * dose of spot s at
  mu = 1) from the oracle's forward code (not its adjoint); grad_xy: central
  finite differences (h = 0.05 mm) of the same.
* geometry_kat: hand-derived source positions for whole-quadrant angles, and
  the product's host-side geometry equal to the fixture.
The oracle must also reproduce every fixture (a regression pin).
"""

import csv
import os

import numpy as np
import pytest

from bruteforce import pb_bruteforce, raymarch_bruteforce, smooth_bruteforce

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["water_slab_g0", "hetero_g120_c30"]


def _load(name):
    return dict(np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False))


@pytest.fixture(scope="module")
def model(plan_models):
    return plan_models.beam_models[0]


def _geom(f, model):
    from oracle import impt_oracle as O
    return O.beam_geometry(f["iso"], np.asarray(f["origin"], np.float32), float(f["gantry"]), float(f["couch"]),
                           model.VSADX, model.VSADY)


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_fixture(name):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    f = _load(name)
    out = mg.compute(f["hu"], float(f["spacing"]), tuple(f["iso"]), float(f["gantry"]), float(f["couch"]),
                     f["spot_list"])
    np.testing.assert_array_equal(out["rlsp"], f["rlsp"])
    for k in ("wet_raw", "wet", "dose", "grad_field", "grad_mu", "grad_xy"):
        np.testing.assert_allclose(out[k], f[k], rtol=1e-12, atol=1e-14 * max(1.0, float(np.abs(f[k]).max())))


@pytest.mark.parametrize("name", CASES)
def test_rlsp_from_csv_knots(name, machine_dir):
    with open(os.path.join(machine_dir, "HU_RLSP.csv")) as fh:
        rows = [r for r in csv.reader(fh) if r]
    hu_k = np.array([float(r[0]) for r in rows[1:]])
    rl_k = np.array([float(r[1]) for r in rows[1:]])
    f = _load(name)
    np.testing.assert_array_equal(f["rlsp"], np.interp(f["hu"], hu_k, rl_k).astype(np.float32))


@pytest.mark.parametrize("name", CASES)
def test_wet_bruteforce(name, model):
    f = _load(name)
    g = _geom(f, model)
    s = float(np.float32(f["spacing"]))
    nz, ny, nx = f["hu"].shape
    rng = np.random.default_rng(3)
    vox = [(0, 0, 0), (nz - 1, ny - 1, nx - 1), (nz // 2, ny // 2, nx // 2)]
    vox += [tuple(int(rng.integers(0, n)) for n in (nz, ny, nx)) for _ in range(9)]
    for v in vox:
        assert abs(f["wet_raw"][v] - raymarch_bruteforce(f["rlsp"], s, g, v)) < 1e-12
    np.testing.assert_allclose(f["wet"], smooth_bruteforce(f["wet_raw"], s, g), rtol=0, atol=1e-12)


@pytest.mark.parametrize("name", CASES)
def test_dose_bruteforce(name, model):
    from oracle import impt_oracle as O
    f = _load(name)
    g = _geom(f, model)
    s = float(np.float32(f["spacing"]))
    _, spots, layers, r80 = O.spots_and_layers(f["spot_list"], model.divergence_params)
    ref = pb_bruteforce(f["wet"], s, g, spots, layers, r80, O.ModelArrays(model))
    assert (ref > 0.05 * ref.max()).sum() > 100, "the fixture must put dose in the volume"
    np.testing.assert_allclose(f["dose"], ref, rtol=1e-11, atol=1e-12 * ref.max())


@pytest.mark.parametrize("name", CASES)
def test_gradients_from_forward(name, model):
    from oracle import impt_oracle as O
    f = _load(name)
    g = _geom(f, model)
    s = float(np.float32(f["spacing"]))
    marr = O.ModelArrays(model)
    sl = f["spot_list"]

    def dose_of(rows):
        _, spots, layers, r80 = O.spots_and_layers(rows, model.divergence_params)
        return O.pb_dose(f["wet"], s, g, spots, layers, r80, marr)

    gf = f["grad_field"]
    for k in range(sl.shape[0]):
        one = sl[k:k + 1].copy()
        one[0, 2] = 1.0
        assert abs(f["grad_mu"][k] - np.sum(gf * dose_of(one))) <= 1e-10 * max(1.0, abs(f["grad_mu"][k]))
    h = 0.05
    scale = np.abs(f["grad_xy"]).max()
    for k in (0, 4, 13, 26):
        for a in range(2):
            p, m = sl[k:k + 1].astype(np.float64), sl[k:k + 1].astype(np.float64)
            p[0, a] += h
            m[0, a] -= h
            fd = (np.sum(gf * dose_of(p)) - np.sum(gf * dose_of(m))) / (2 * h)
            assert abs(f["grad_xy"][k, a] - fd) <= 2e-3 * scale


def test_geometry_kat():
    from DoseCUDA.geometry import beam_geometry
    f = _load("geometry_kat")
    iso, origin, (vx, vy) = f["iso"], f["origin"], f["vsad"]
    sad = (vx + vy) / 2.0
    # whole-quadrant gantry angles at couch 0: source at (-sad sin(g + 180), sad cos(g + 180), 0)
    hand = {0.0: (0.0, -sad, 0.0), 90.0: (sad, 0.0, 0.0), 180.0: (0.0, sad, 0.0), 270.0: (-sad, 0.0, 0.0)}
    for (gantry, couch), row in zip(f["angles"], f["geom"]):
        np.testing.assert_array_equal(row[:3], iso - origin)
        if couch == 0.0 and gantry in hand:
            np.testing.assert_array_equal(row[3:6], hand[gantry])
        assert abs(np.linalg.norm(row[3:6]) - sad) < 1e-9
        assert abs(row[6] ** 2 + row[7] ** 2 - 1) < 1e-15 and abs(row[8] ** 2 + row[9] ** 2 - 1) < 1e-15
        d = beam_geometry(iso, origin, gantry, couch, vx, vy)
        got = np.concatenate([d["iso"], d["src"], [d["sin_g"], d["cos_g"], d["sin_c"], d["cos_c"], d["vsadx"],
                                                   d["vsady"]]])
        np.testing.assert_array_equal(got, row)
