"""Host-side beam geometry and spot/layer tables shared by the dose entry points.

* ``beam_geometry``  CudaBeam constructor (CudaClasses.cu:17-51) with the
  isocentre / gantry adjustments of dosemodule.cu:148-164, evaluated in float64.
  Sines and cosines of angles that are whole multiples of 90 degrees are exact
  (0, +-1): the geometry of axis-aligned beams is then exact in float32 too,
  which keeps exact ray/plane and half-voxel ties identical between float32
  kernels and a float64 evaluation.
* ``sort_spots`` / ``layer_table``  make_spot_array + importLayers
  (dosemodule.cu:211-225, IMPTClasses.cu:37-67).
"""

import math

import numpy as np


def sincos_deg(angle):
    """(sin, cos) of an angle in degrees, exact at multiples of 90 degrees."""
    a = math.fmod(float(angle), 360.0)
    if a < 0.0:
        a += 360.0
    q = a / 90.0
    if q == math.floor(q):
        return [(0.0, 1.0), (1.0, 0.0), (0.0, -1.0), (-1.0, 0.0)][int(q) % 4]
    r = math.radians(a)
    return math.sin(r), math.cos(r)


def beam_geometry(iso, origin, gantry_angle, couch_angle, vsadx, vsady):
    """Beam geometry in float64.

    Returns a dict with ``iso`` (isocentre minus grid origin, mm), ``src``
    (source relative to the isocentre), ``sin_g``/``cos_g`` of gantry + 180 deg,
    ``sin_c``/``cos_c`` of the couch angle and the two VSADs.  The source sits
    at distance (VSADX + VSADY) / 2 (IMPTClasses.cuh:36), rotated by the
    gantry about z, then by the couch about y.
    """
    iso = np.asarray(iso, dtype=np.float64)
    origin = np.asarray(origin, dtype=np.float64)
    iso_adj = iso[:3] - origin[:3]
    sg, cg = sincos_deg(math.fmod(float(gantry_angle) + 180.0, 360.0))
    sc, cc = sincos_deg(float(couch_angle))
    sad = (float(vsadx) + float(vsady)) / 2.0
    xg = -sad * sg
    yg = sad * cg
    src = np.array([xg * cc, yg, -xg * sc], dtype=np.float64)
    return {"iso": iso_adj, "src": src, "sin_g": sg, "cos_g": cg, "sin_c": sc, "cos_c": cc,
            "vsadx": float(vsadx), "vsady": float(vsady)}


def energy_ids(spot_list):
    """Integer energy ids of a float [S, 4] spot list: truncation toward zero,
    as static_cast<int>(float) (dosemodule.cu:220)."""
    with np.errstate(invalid="ignore"):
        return np.asarray(spot_list, dtype=np.float32)[:, 3].astype(np.int32)


def morton_key(x, y, quantum=1.0):
    """Z-order key of spot positions on a ``quantum`` mm grid (16 bits per axis).
    Used to order spots inside a layer so that consecutive spots are close
    together; it only changes the order of summation."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    if x.size == 0:
        return np.zeros(0, dtype=np.int64)
    qx = np.nan_to_num((x - np.nanmin(x)) / quantum)
    qy = np.nan_to_num((y - np.nanmin(y)) / quantum)
    qx = np.clip(qx, 0, 65535).astype(np.int64)
    qy = np.clip(qy, 0, 65535).astype(np.int64)

    def spread(v):
        v = (v | (v << 8)) & 0x00FF00FF
        v = (v | (v << 4)) & 0x0F0F0F0F
        v = (v | (v << 2)) & 0x33333333
        v = (v | (v << 1)) & 0x55555555
        return v

    return spread(qx) | (spread(qy) << 1)


def sort_spots(spot_list, spatial=True):
    """Permutation sorting spots by energy id (then by Z-order position inside
    a layer when ``spatial``).  Returns (order, sorted integer energy ids)."""
    sl = np.asarray(spot_list, dtype=np.float32)
    ids = energy_ids(sl)
    if spatial:
        order = np.lexsort((morton_key(sl[:, 0], sl[:, 1]), ids))
    else:
        order = np.argsort(ids, kind="stable")
    return order, ids[order]


def layer_table(sorted_ids, divergence_params):
    """Energy layers of a sorted id array, as importLayers builds them
    (IMPTClasses.cu:37-67): energies 0 .. n_E-1 are matched in turn against
    the run that starts where the previous layer ended.  Consequences kept on
    purpose: ids >= n_E are dropped, and a negative id (which sorts first and
    never matches) leaves the beam with no layers at all.

    Returns an int32 array [L, 3] (spot_start, n_spots, energy_id) and float32
    arrays r80 [L] and energy [L].
    """
    dvp = np.asarray(divergence_params, dtype=np.float32)
    n_e = dvp.shape[0]
    ids = np.asarray(sorted_ids, dtype=np.int64)
    rows = []
    if ids.size and ids[0] >= 0:
        vals, starts, counts = np.unique(ids, return_index=True, return_counts=True)
        for v, s, c in zip(vals, starts, counts):
            if v >= n_e:
                break
            rows.append((int(s), int(c), int(v)))
    tab = np.array(rows, dtype=np.int32).reshape(-1, 3)
    r80 = dvp[tab[:, 2], 1] if len(rows) else np.zeros(0, np.float32)
    energy = dvp[tab[:, 2], 0] if len(rows) else np.zeros(0, np.float32)
    return tab, r80.astype(np.float32), energy.astype(np.float32)
