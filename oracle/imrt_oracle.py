"""FP64 oracle of the photon collapsed-cone (CCC) dose path -- TEST INFRASTRUCTURE ONLY.

Imported by tests/ and by bench.py's photon CPU-baseline leg, and only to
compare against: the product (DoseCUDA package, libdosehip.so) never imports
it.

What it restates (reference files under DoseCUDA/):
  * beam geometry          CudaClasses.cu:17-51 with the primary source
                           distance (IMRTClasses.cu:30-39), dosemodule.cu:497-509
  * HU -> density          plan_imrt.py:114-124 (np.interp, cast to float32)
  * radiological depth     rayTraceKernel CudaClasses.cu:180-228 on the density
                           texture (border AIR_DENSITY)
  * TERMA and electron     termaKernel IMRTClasses.cu:209-250
  * dose                   cccKernel IMRTClasses.cu:252-350
  * output factor, plan    IMRTPhotonEnergy.outputFactor plan_imrt.py:61-92,
    loop                   computeIMRTPlan plan_imrt.py:126-158
in float64 (imrt_oracle.c), with the texture semantics of TextureClasses.cu
(border addressing, unnormalised coordinates, real-number trilinear weights).

Pinning (tests/test_photon_oracle_pins.py): closed forms (uniform-density
depth counts, the collapsed-cone recursion in a uniform medium, a contiguous
open field's transmission), independent brute force (scipy map_coordinates
'grid-constant' with the border value, a numerical integral of the Gaussian
source over the projected aperture, scipy's Rotation for the kernel tilt),
np.interp for the tables, and the tables straight from the CSV text.
"""

import ctypes
import math
import os

import numpy as np

from . import impt_oracle as _IO

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}
AIR_DENSITY = 0.0012


def lib(precision="f64"):
    name = "libimrt_oracle.so" if precision == "f64" else "libimrt_oracle_f32.so"
    if name not in _LIBS:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            _IO.build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i, d, i64, f = ctypes.c_int, ctypes.c_double, ctypes.c_int64, ctypes.c_float
        L.orc_ph_raytrace.argtypes = [P, i, i, i, d, P, P, i64, P]
        L.orc_ph_head_transmission.argtypes = [P, i64, d, d, d, d, P, i, P]
        L.orc_ph_terma.argtypes = [P, i, i, i, d, P, P, P, i, P, i, P, i, P, i64, P, P]
        L.orc_ph_tilt.argtypes = [P, P, P, i64, P]
        L.orc_ph_ccc.argtypes = [P, P, P, i, i, i, f, P, P, d, d, P, i64, P]
        L.orc_ph_frames.argtypes = [P, P, i64, i, P]
        for fn in (L.orc_ph_raytrace, L.orc_ph_head_transmission, L.orc_ph_terma, L.orc_ph_tilt, L.orc_ph_ccc,
                   L.orc_ph_frames):
            fn.restype = ctypes.c_int
        _LIBS[name] = L
    return _LIBS[name]


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ---------------------------------------------------------------------------
# geometry
def photon_geometry(iso, origin, gantry, couch, collimator, sad):
    """float64 [12]: iso - origin, source (iso-relative, at distance sad),
    sin/cos of gantry + 180, couch and collimator (IMRTClasses.cu:30-39,
    CudaClasses.cu:17-51, dosemodule.cu:497-509)."""
    iso = np.asarray(iso, np.float64)[:3] - np.asarray(origin, np.float64)[:3]
    sg, cg = _IO.sincos_deg(math.fmod(float(gantry) + 180.0, 360.0))
    sc, cc = _IO.sincos_deg(couch)
    sca, cca = _IO.sincos_deg(collimator)
    xg, yg = -float(sad) * sg, float(sad) * cg
    return np.array([iso[0], iso[1], iso[2], xg * cc, yg, -xg * sc, sg, cg, sc, cc, sca, cca], np.float64)


def image_to_head(geom, p):
    """numpy twin of img_to_head (IMRTClasses.cu:41-79), for tests."""
    sg, cg, sc, cc, sca, cca = geom[6:12]
    x, y, z = (np.asarray(c, np.float64) for c in p)
    xt = x * cc - z * sc
    zt = x * sc + z * cc
    xg = xt * cg + y * sg
    yg = -xt * sg + y * cg
    xc = xg * cca - zt * sca
    zc = xg * sca + zt * cca
    return np.array([-xc, zc, yg])


# ---------------------------------------------------------------------------
# model tables
def model_arrays(model):
    """(scalars [10], profile [3, n], spectrum [3, n], kernel float32 [6, 6])
    of an IMRTPhotonEnergy-like object; kernel rows are cones (theta, Am, am,
    Bm, bm, ray_length), as the reference's cccKernel reads them."""
    m = np.array([model.primary_source_distance, model.scatter_source_distance, model.primary_source_size,
                  model.scatter_source_size, model.mlc_distance, model.scatter_source_weight,
                  model.electron_attenuation, model.electron_source_weight, model.electron_fitted_dmax,
                  model.mu_calibration], np.float64)
    prof = _f64(np.stack([np.asarray(model.profile_radius, np.float32), np.asarray(model.profile_intensities, np.float32),
                          np.asarray(model.profile_softening, np.float32)]).astype(np.float64))
    spec = _f64(np.stack([np.asarray(model.spectrum_attenuation_coefficients, np.float32),
                          np.asarray(model.spectrum_primary_weights, np.float32),
                          np.asarray(model.spectrum_scatter_weights, np.float32)]).astype(np.float64))
    kern = _f32(np.asarray(model.kernel, np.float32).reshape(6, 6))
    return m, prof, spec, kern


def density_from_hu(hu, machine_dir):
    """np.interp of HU over the HU_Density.csv knots, cast to float32
    (IMRTDoseGrid.DensityFromHU, plan_imrt.py:114-124)."""
    tab = np.loadtxt(os.path.join(machine_dir, "HU_Density.csv"), delimiter=",", skiprows=1)
    return np.array(np.interp(hu, tab[:, 0], tab[:, 1]), dtype=np.single)


# ---------------------------------------------------------------------------
# kernels
def raytrace(density, spacing, geom, idx=None, precision="f64"):
    rho = _f32(density)
    nz, ny, nx = rho.shape
    idx_a = None if idx is None else np.ascontiguousarray(idx, np.int64)
    out = np.zeros(rho.size if idx is None else len(idx_a), np.float64)
    lib(precision).orc_ph_raytrace(_p(rho), nx, ny, nz, float(spacing), _p(_f64(geom)), _p(idx_a),
                                   0 if idx_a is None else len(idx_a), _p(out))
    return out.reshape(rho.shape) if idx is None else out


def head_transmission(points_head, sad_p, mlc_distance, iso_to_source, sigma, mlc, precision="f64"):
    pts = _f64(np.asarray(points_head, np.float64).reshape(-1, 3))
    mlc_a = _f64(np.asarray(mlc, np.float64).reshape(-1, 4))
    out = np.zeros(len(pts), np.float64)
    lib(precision).orc_ph_head_transmission(_p(pts), len(pts), float(sad_p), float(mlc_distance),
                                            float(iso_to_source), float(sigma), _p(mlc_a), len(mlc_a), _p(out))
    return out


def terma(wet, shape, spacing, geom, model, mlc, idx=None, precision="f64"):
    """(TERMA, electron) at every voxel (wet: full volume) or at the listed
    voxels (wet: one value per listed voxel)."""
    nz, ny, nx = shape
    m, prof, spec, _ = model_arrays(model)
    w = _f64(np.asarray(wet, np.float64).ravel())
    mlc_a = _f64(np.asarray(mlc, np.float64).reshape(-1, 4))
    idx_a = None if idx is None else np.ascontiguousarray(idx, np.int64)
    n = nx * ny * nz if idx is None else len(idx_a)
    te = np.zeros(n, np.float64)
    el = np.zeros(n, np.float64)
    lib(precision).orc_ph_terma(_p(w), nx, ny, nz, float(spacing), _p(_f64(geom)), _p(m), _p(prof), prof.shape[1],
                                _p(spec), spec.shape[1], _p(mlc_a), len(mlc_a), _p(idx_a),
                                0 if idx_a is None else len(idx_a), _p(te), _p(el))
    if idx is None:
        return te.reshape(shape), el.reshape(shape)
    return te, el


def frames(geom, pts, to_head=True):
    """Image -> head (to_head) or head -> image of points [n, 3] (C oracle)."""
    pts = _f64(np.asarray(pts, np.float64).reshape(-1, 3))
    out = np.zeros_like(pts)
    lib().orc_ph_frames(_p(_f64(geom)), _p(pts), len(pts), 1 if to_head else 0, _p(out))
    return out


def tilt(geom, vox, vec):
    vox = _f64(np.asarray(vox, np.float64).reshape(-1, 3))
    vec = _f64(np.asarray(vec, np.float64).reshape(-1, 3))
    out = np.zeros_like(vec)
    lib().orc_ph_tilt(_p(_f64(geom)), _p(vox), _p(vec), len(vec), _p(out))
    return out


def ccc(terma32, density, electron, spacing, geom, kernel, mu_cal, mu, idx=None, precision="f64"):
    """Dose (Gy-like units of the model) at every voxel or the listed ones;
    terma32 and density are the float32 volumes the reference's textures hold,
    electron one value per voxel (full volume) or per listed voxel."""
    t = _f32(terma32)
    rho = _f32(density)
    nz, ny, nx = rho.shape
    el = _f64(np.asarray(electron, np.float64).ravel())
    idx_a = None if idx is None else np.ascontiguousarray(idx, np.int64)
    out = np.zeros(rho.size if idx is None else len(idx_a), np.float64)
    lib(precision).orc_ph_ccc(_p(t), _p(rho), _p(el), nx, ny, nz, float(np.float32(spacing)), _p(_f64(geom)),
                              _p(_f32(kernel)), float(mu_cal), float(mu), _p(idx_a),
                              0 if idx_a is None else len(idx_a), _p(out))
    return out.reshape(rho.shape) if idx is None else out


def mlc_rows(mlc):
    """MLC leaf pairs as rows (x1, x2, y_offset, y_width): the control point's
    [n, 4] array read by index (the reference's make_mlc_array reads the same
    values out of the transposed view IMRTPlan builds, dosemodule.cu:228-240)."""
    return np.asarray(mlc, np.float32).reshape(-1, 4).astype(np.float64)


def cp_dose(density, spacing, origin, model, cp, return_parts=False):
    """One control point's dose (photon_dose_cuda semantics, IMRTClasses.cu:353-405):
    depth, TERMA / electron, then the collapsed-cone sum, with TERMA rounded to
    float32 as the reference's TERMA texture holds it."""
    s = float(np.float32(spacing))
    geom = photon_geometry(np.asarray(cp.iso, np.float32), np.asarray(origin, np.float32), cp.ga, cp.ta, cp.ca,
                           np.float32(model.primary_source_distance))
    rho = _f32(density)
    wet = raytrace(rho, s, geom)
    mlc = mlc_rows(cp.mlc)
    te, el = terma(wet.astype(np.float32).astype(np.float64), rho.shape, s, geom, model, mlc)
    t32 = te.astype(np.float32)
    dose = ccc(t32, rho, el.astype(np.float32).astype(np.float64), s, geom, model_arrays(model)[3],
               np.float32(model.mu_calibration), np.float32(cp.mu))
    if return_parts:
        return dose, {"wet": wet, "terma": te, "electron": el, "geom": geom}
    return dose


def output_factor(model, cp):
    """Equivalent-square output factor (IMRTPhotonEnergy.outputFactor, plan_imrt.py:61-92)."""
    mlc = np.asarray(cp.mlc, np.float32).reshape(-1, 4)
    min_y, max_y, max_x, area = 10000.0, -10000.0, 0.0, 0.0
    for x1, x2, yo, yw in mlc.astype(np.float64):
        area += (x2 - x1) * yw
        if (x2 - x1) > 3.0 and (x2 - x1) > max_x:
            max_x = x2 - x1
        if (x2 - x1) > 3.0:
            min_y = min(min_y, yo - yw / 2.0)
            max_y = max(max_y, yo + yw / 2.0)
    eq = 4 * area / (2 * (max_y - min_y) + 2 * max_x)
    return float(np.interp(eq, model.output_factor_equivalent_squares, model.output_factor_values))


def compute_plan(dose_grid, plan, machine_dir):
    """Total dose (float64, (z, y, x)) times n_fractions, plus per-beam doses
    (computeIMRTPlan, plan_imrt.py:126-158)."""
    if not (dose_grid.spacing[0] == dose_grid.spacing[1] == dose_grid.spacing[2]):
        raise ValueError("Spacing must be isotropic")
    rho = density_from_hu(dose_grid.HU, machine_dir)
    labels = [str(x) for x in plan.dicom_energy_label]
    total = np.zeros(rho.shape)
    beams = []
    for beam in plan.beam_list:
        model = plan.beam_models[labels.index(str(beam.dicom_energy_label))]
        bd = np.zeros(rho.shape)
        for cp in beam.cp_list:
            bd += cp_dose(rho, dose_grid.spacing[0], dose_grid.origin, model, cp) * output_factor(model, cp)
        beams.append(bd)
        total += bd
    return total * plan.n_fractions, beams
