"""Synthetic dose-grid / plan builders for the benchmark workloads.

Each builder returns ``(IMPTDoseGrid, IMPTPlan)`` with HU in (z, y, x), an
isotropic grid centred on the isocentre (origin = -size * spacing / 2 per axis,
as createCubePhantom, plan.py:226-232) and spots in ``spot_list`` rows
(x mm, y mm, weight, energy id) on range shifter '0' (RS0).

* water_slab_plan    100^3 @ 2.5 mm water, 20 mm HU=1000 slab at 60 mm depth,
                     gantry 0, 5 layers (R80 80-147 mm), 8x8 spots at 5 mm
                     (320 spots)
* head_plan          160x200x200 @ 2 mm head-size phantom (ellipsoidal soft
                     tissue, 8 mm bone shell, two air spheres), gantries
                     0/120/240, 40 layers (RS0 ids 20..59), 81 spots per layer
                     on a 3 mm grid inside a 15 mm disk (9720 spots)
* target_plan        water_slab geometry, 50 mm spherical target at the
                     isocentre, 25 layers x 400 spots (10000 spots), plus the
                     target / shell masks of its loss
* uneven_plans       plans of random cube size, beam count and spot count
* body_plan          300x512x512 @ 1 mm body (elliptic cylinder N(20, 30) HU,
                     two lungs U(-850, -700), spine U(300, 1100)), gantries
                     0/90/180/270, 50 layers (RS0 ids 25..74) x spots on a
                     3 mm grid in a 30 mm disk (~63k spots), 30 mm target
* imrt_head_plan     photons: the head phantom on an IMRTDoseGrid, a 7-field
                     step-and-shoot plan (gantries 0, 51.4, ..., 308.6 deg),
                     5 MLC segments per beam shrinking from 80x80 mm with
                     random leaf offsets, 10-40 MU each (35 control points)
"""

import numpy as np

from .plan_impt import IMPTDoseGrid, IMPTPlan, IMPTBeam


def _grid(shape_zyx, spacing):
    g = IMPTDoseGrid()
    g.spacing = np.array([spacing] * 3, dtype=np.float64)
    g.size = np.array(shape_zyx)
    # origin (x, y, z) = -n * s / 2 along each axis
    g.origin = np.array([-shape_zyx[2] * spacing / 2.0, -shape_zyx[1] * spacing / 2.0,
                         -shape_zyx[0] * spacing / 2.0])
    return g


def voxel_centres(g):
    """Voxel-centre coordinates (z, y, x) relative to the grid centre (mm)."""
    nz, ny, nx = (int(v) for v in g.size)
    s = float(g.spacing[0])
    z = g.origin[2] + s * np.arange(nz)
    y = g.origin[1] + s * np.arange(ny)
    x = g.origin[0] + s * np.arange(nx)
    return z, y, x


def _beam(gantry, spots, couch=0.0):
    b = IMPTBeam()
    b.dicom_rangeshifter_label = '0'
    b.gantry_angle = float(gantry)
    b.couch_angle = float(couch)
    b.spot_list = np.asarray(spots, dtype=np.single)
    b.n_spots = b.spot_list.shape[0]
    return b


def grid_spots(energy_ids, coords_x, coords_y, weights=1.0, disk_radius=None):
    rows = []
    for e in energy_ids:
        for y in coords_y:
            for x in coords_x:
                if disk_radius is not None and x * x + y * y > disk_radius * disk_radius + 1e-9:
                    continue
                rows.append((x, y, 1.0, e))
    sl = np.array(rows, dtype=np.float64)
    if np.isscalar(weights):
        sl[:, 2] = weights
    else:
        sl[:, 2] = np.asarray(weights)[: sl.shape[0]]
    return sl.astype(np.float32)


def water_slab_plan(n=100, spacing=2.5, slab_start_mm=60.0, slab_mm=20.0):
    """Water cube, HU = 1000 slab 60-80 mm deep from the first voxel centre on
    the beam-entrance (-y) side, gantry 0, 5 layers x 64 spots."""
    g = _grid((n, n, n), spacing)
    hu = np.zeros((n, n, n), dtype=np.float32)
    j0 = int(round(slab_start_mm / spacing))
    j1 = int(round((slab_start_mm + slab_mm) / spacing))
    hu[:, j0:j1, :] = 1000.0
    g.HU = hu
    plan = IMPTPlan()
    xs = np.arange(-17.5, 17.5 + 1e-6, 5.0)
    plan.addBeam(_beam(0.0, grid_spots([35, 42, 49, 56, 62], xs, xs)))
    return g, plan


def head_phantom(shape=(160, 200, 200), spacing=2.0, seed=2):
    g = _grid(shape, spacing)
    rng = np.random.default_rng(seed)
    z, y, x = voxel_centres(g)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    ax, ay, az = 75.0, 95.0, 110.0
    r_out = (X / ax) ** 2 + (Y / ay) ** 2 + (Z / az) ** 2
    sh = 8.0
    r_in = (X / (ax - sh)) ** 2 + (Y / (ay - sh)) ** 2 + (Z / (az - sh)) ** 2
    hu = np.full(shape, -1000.0, dtype=np.float32)
    body = r_out <= 1.0
    hu[body] = np.clip(rng.normal(40.0, 20.0, size=int(body.sum())), -100.0, 100.0)
    shell = body & (r_in > 1.0)
    hu[shell] = rng.uniform(700.0, 1200.0, size=int(shell.sum()))
    for cx, cy, cz in ((-30.0, 20.0, 40.0), (30.0, 20.0, 40.0)):
        hu[(X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= 15.0 ** 2] = -1000.0
    g.HU = hu.astype(np.float32)
    return g, rng


def head_plan(seed=2, gantries=(0.0, 120.0, 240.0), layers=range(20, 60)):
    g, rng = head_phantom(seed=seed)
    plan = IMPTPlan()
    c = np.arange(-15.0, 15.0 + 1e-6, 3.0)
    for ga in gantries:
        sl = grid_spots(list(layers), c, c, disk_radius=15.0)
        sl[:, 2] = rng.uniform(0.0, 1.0, size=sl.shape[0]).astype(np.float32)
        plan.addBeam(_beam(ga, sl))
    return g, plan


def target_plan(n=100, spacing=2.5, n_layers=25, grid=20, pitch=2.5, first_energy=None):
    """Water-slab geometry with ~10k spots covering a 50 mm target at the
    isocentre.  Returns (dose_grid, plan, target_mask, shell_mask)."""
    g, _ = water_slab_plan(n, spacing)
    plan = IMPTPlan()
    model = plan.beam_models[0]
    r80 = model.divergence_params[:, 1]
    if first_energy is None:
        # layers whose R80 spans the target's water-equivalent depths
        first_energy = int(np.searchsorted(r80, 105.0))
    ids = list(range(first_energy, first_energy + n_layers))
    half = (grid - 1) * pitch / 2.0
    c = np.linspace(-half, half, grid)
    plan.addBeam(_beam(0.0, grid_spots(ids, c, c)))
    z, y, x = voxel_centres(g)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    target = r <= 25.0
    shell = (r > 25.0) & (r <= 35.0)
    return g, plan, target, shell


def uneven_plans(n_plans=64, seed=4, sides=(64, 256), spacing=2.0, spots=(500, 40000)):
    """Plans of uneven size: cube side ~ U{sides}, 1-4 beams at random gantry,
    spot count log-uniform in ``spots``; water with a 20 mm HU = 1000 slab at
    60 mm depth (the water_slab pattern)."""
    out = []
    for p in range(n_plans):
        rng = np.random.default_rng(seed * 1000 + p)
        n = int(rng.integers(sides[0], sides[1] + 1))
        g = _grid((n, n, n), spacing)
        hu = np.zeros((n, n, n), dtype=np.float32)
        j0, j1 = int(round(60.0 / spacing)), int(round(80.0 / spacing))
        hu[:, j0:j1, :] = 1000.0
        g.HU = hu
        plan = IMPTPlan()
        nb = int(rng.integers(1, 5))
        total = int(round(np.exp(rng.uniform(np.log(spots[0]), np.log(spots[1])))))
        per_beam = max(1, total // nb)
        for _ in range(nb):
            ga = float(rng.uniform(0.0, 360.0))
            n_layers = int(min(40, max(1, per_beam // 100)))
            per_layer = max(1, per_beam // n_layers)
            side = int(np.ceil(np.sqrt(per_layer)))
            c = (np.arange(side) - (side - 1) / 2.0) * 3.0
            ids = list(range(20, 20 + n_layers))
            sl = grid_spots(ids, c, c)[: n_layers * per_layer]
            sl[:, 2] = rng.uniform(0.0, 1.0, size=sl.shape[0]).astype(np.float32)
            plan.addBeam(_beam(ga, sl))
        out.append((g, plan))
    return out


def body_plan(shape=(300, 512, 512), spacing=1.0, seed=5, gantries=(0.0, 90.0, 180.0, 270.0),
              layers=range(25, 75), spot_radius=30.0, pitch=3.0):
    """Large body phantom (seed 5): elliptic cylinder along z (semi-axes 180 mm
    in x, 120 mm in y) of soft tissue N(20, 30) HU, two lungs (ellipsoids at
    x = +-80 mm, semi-axes 55 x 65 x 120 mm) U(-850, -700), a spine (cylinder of
    radius 15 mm at y = +85 mm) U(300, 1100), air outside.  Returns (dose_grid,
    plan, target_mask) with a 30 mm spherical target at the isocentre."""
    g = _grid(shape, spacing)
    rng = np.random.default_rng(seed)
    z, y, x = voxel_centres(g)
    Z = z[:, None, None]
    Y = y[None, :, None]
    X = x[None, None, :]
    hu = np.full(shape, -1000.0, dtype=np.float32)
    body = np.broadcast_to((X / 180.0) ** 2 + (Y / 120.0) ** 2 <= 1.0, shape)
    hu[body] = rng.normal(20.0, 30.0, size=int(body.sum())).astype(np.float32)
    for cx in (-80.0, 80.0):
        lung = np.broadcast_to(((X - cx) / 55.0) ** 2 + ((Y + 5.0) / 65.0) ** 2 + (Z / 120.0) ** 2 <= 1.0, shape)
        hu[lung] = rng.uniform(-850.0, -700.0, size=int(lung.sum())).astype(np.float32)
    spine = np.broadcast_to(X ** 2 + (Y - 85.0) ** 2 <= 15.0 ** 2, shape)
    hu[spine] = rng.uniform(300.0, 1100.0, size=int(spine.sum())).astype(np.float32)
    g.HU = hu
    plan = IMPTPlan()
    c = np.arange(-spot_radius, spot_radius + 1e-6, pitch)
    for ga in gantries:
        sl = grid_spots(list(layers), c, c, disk_radius=spot_radius)
        sl[:, 2] = rng.uniform(0.0, 1.0, size=sl.shape[0]).astype(np.float32)
        plan.addBeam(_beam(ga, sl))
    target = np.broadcast_to(X ** 2 + Y ** 2 + Z ** 2 <= 30.0 ** 2, shape)
    return g, plan, np.ascontiguousarray(target)


def imrt_head_plan(seed=2, n_beams=7, segments=5, machine="VarianTrueBeamHF"):
    """Photon step-and-shoot plan on the head phantom.  Returns
    (IMRTDoseGrid, IMRTPlan)."""
    from .plan_imrt import IMRTDoseGrid, IMRTPlan, IMRTBeam, IMRTControlPoint
    hg, _ = head_phantom(seed=seed)
    g = IMRTDoseGrid()
    g.spacing, g.size, g.origin, g.HU = hg.spacing, hg.size, hg.origin, hg.HU
    rng = np.random.default_rng(seed + 100)
    plan = IMRTPlan(machine)
    bm = plan.beam_models[0]
    for b in range(n_beams):
        beam = IMRTBeam()
        beam.dicom_energy_label = str(plan.dicom_energy_label[0])
        ga = 360.0 * b / n_beams
        ca = float(rng.choice([0.0, 10.0, 350.0]))
        for s in range(segments):
            half = 40.0 - 6.0 * s
            leaves = np.zeros((2, bm.n_mlc_pairs), np.float32)
            open_ = (bm.mlc_offsets >= -half) & (bm.mlc_offsets <= half)
            leaves[0, open_] = (-half + rng.uniform(-8, 8, open_.sum())).astype(np.float32)
            leaves[1, open_] = (half + rng.uniform(-8, 8, open_.sum())).astype(np.float32)
            leaves[1] = np.maximum(leaves[1], leaves[0])
            mlc = plan._mlc_array(leaves)
            beam.addControlPoint(IMRTControlPoint(np.zeros(3, np.float32), float(rng.uniform(10, 40)), mlc, ga, ca,
                                                  0.0, None, None))
        plan.addBeam(beam)
    return g, plan
