#!/usr/bin/env python3
"""Write the golden fixtures of this directory from the FP64 oracle.

    python tests/golden/make_golden.py

Plans (HitachiProbeatJHU, range shifter 0, one beam, 3 layers x 9 spots):

* water_slab_g0.npz: 12 x 14 x 12 (z, y, x) water grid at 5 mm with an
  HU = 1000 slab at y indices 5..7, gantry 0, couch 0, isocentre at the grid
  centre; energy ids 0, 3, 6 (R80 39-45 mm); spots on {-8, 0, 8}^2 mm.
* hetero_g120_c30.npz: 10 x 12 x 14 grid at 4 mm, HU ~ U(-200, 300) with a
  bone bar (HU 1200) and an air pocket (-1000), gantry 120, couch 30,
  isocentre off the centre; energy ids 1, 4, 8; spots on {-6, 0, 6}^2 mm.

Each file holds the inputs (hu, spacing, origin, iso, gantry, couch,
spot_list [x, y, mu, energy id], n_fractions) and the oracle's outputs:
rlsp (np.interp, float32), wet_raw and wet (cm), dose (plan total times
n_fractions), grad_field (a seeded dL/dD) and the gradients of
L = sum(grad_field * dose / n_fractions) with respect to the spot weights
(grad_mu) and positions (grad_xy), in spot-list order.

geometry_kat.npz: beam geometry vectors (iso - origin, source, sin/cos of
gantry + 180, sin/cos couch, VSADX, VSADY) for a set of gantry/couch angles.

The checks that pin these values independently of the oracle's code live in
tests/test_golden.py.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "dosecuda-jax_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

MACHINE = "HitachiProbeatJHU"
MACHINE_DIR = os.path.join(ROOT, "dosecuda-jax_amd", "DoseCUDA", "lookuptables", "protons", MACHINE)
GEOMETRY_ANGLES = [(0.0, 0.0), (90.0, 0.0), (180.0, 0.0), (270.0, 0.0), (120.0, 30.0), (240.0, -45.0),
                   (37.0, 10.0), (315.0, 90.0)]


def _spots(energy_ids, coords, seed):
    rng = np.random.default_rng(seed)
    rows = [(x, y, 0.0, e) for e in energy_ids for y in coords for x in coords]
    sl = np.array(rows, np.float64)
    sl[:, 2] = rng.uniform(0.5, 1.5, size=len(rows))
    return sl.astype(np.float32)


def cases():
    """(name, hu, spacing, iso, gantry, couch, spot_list)."""
    hu1 = np.zeros((12, 14, 12), np.float32)
    hu1[:, 5:8, :] = 1000.0
    rng = np.random.default_rng(21)
    hu2 = rng.uniform(-200.0, 300.0, size=(10, 12, 14)).astype(np.float32)
    hu2[:, 4:6, 3:11] = 1200.0
    hu2[2:5, 7:10, 5:9] = -1000.0
    return [
        ("water_slab_g0", hu1, 5.0, (0.0, 0.0, 0.0), 0.0, 0.0, _spots([0, 3, 6], [-8.0, 0.0, 8.0], 1)),
        ("hetero_g120_c30", hu2, 4.0, (2.0, -3.0, 1.5), 120.0, 30.0, _spots([1, 4, 8], [-6.0, 0.0, 6.0], 2)),
    ]


def origin_of(shape, spacing):
    nz, ny, nx = shape
    return np.array([-nx * spacing / 2.0, -ny * spacing / 2.0, -nz * spacing / 2.0])


def grad_field(shape, seed):
    return np.random.default_rng(seed).normal(size=shape)


def compute(hu, spacing, iso, gantry, couch, spot_list, field_seed=7):
    from DoseCUDA import IMPTPlan
    from oracle import impt_oracle as O
    model = IMPTPlan().beam_models[0]
    origin = origin_of(hu.shape, spacing)
    rlsp = O.rlsp_from_hu(hu, MACHINE_DIR)
    s = float(np.float32(spacing))
    geom = O.beam_geometry(np.asarray(iso, np.float32), np.asarray(origin, np.float32), gantry, couch,
                           model.VSADX, model.VSADY)
    raw = O.raymarch(rlsp, s, geom)
    wet = O.smooth(raw, s, geom)
    order, spots, layers, r80 = O.spots_and_layers(spot_list, model.divergence_params)
    marr = O.ModelArrays(model)
    dose = O.pb_dose(wet, s, geom, spots, layers, r80, marr)
    gf = grad_field(hu.shape, field_seed) * (dose > 0)
    gmu_s, gxy_s = O.pb_grad(wet, gf, s, geom, spots, layers, r80, marr)
    gmu = np.empty_like(gmu_s)
    gxy = np.empty_like(gxy_s)
    gmu[order] = gmu_s
    gxy[order] = gxy_s
    return {"rlsp": rlsp, "wet_raw": raw, "wet": wet, "dose": dose, "grad_field": gf, "grad_mu": gmu,
            "grad_xy": gxy, "origin": origin, "geom": geom}


def main():
    from DoseCUDA import IMPTPlan
    from oracle import impt_oracle as O
    for name, hu, spacing, iso, gantry, couch, sl in cases():
        out = compute(hu, spacing, iso, gantry, couch, sl)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), hu=hu, spacing=np.float64(spacing),
                            origin=out["origin"], iso=np.asarray(iso, np.float32), gantry=np.float64(gantry),
                            couch=np.float64(couch), spot_list=sl, n_fractions=np.int64(1), rlsp=out["rlsp"],
                            wet_raw=out["wet_raw"], wet=out["wet"], dose=out["dose"], grad_field=out["grad_field"],
                            grad_mu=out["grad_mu"], grad_xy=out["grad_xy"])
        print(name, "dose max %.4g, voxels with dose %d" % (out["dose"].max(), int((out["dose"] > 0).sum())))
    model = IMPTPlan().beam_models[0]
    iso = np.array([1.5, -2.0, 4.0])
    origin = np.array([-30.0, -35.0, -25.0])
    geoms = np.array([O.beam_geometry(iso, origin, g, c, model.VSADX, model.VSADY) for g, c in GEOMETRY_ANGLES])
    np.savez_compressed(os.path.join(HERE, "geometry_kat.npz"), angles=np.array(GEOMETRY_ANGLES), iso=iso,
                        origin=origin, vsad=np.array([model.VSADX, model.VSADY], np.float64), geom=geoms)


if __name__ == "__main__":
    main()
