"""FP64 oracle of the proton pencil-beam dose path -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline
leg, and only to compare against: the product (DoseCUDA package, libdosehip.so)
never imports it.

What it restates (reference files under DoseCUDA/):
  * beam geometry           CudaClasses.cu:17-51, dosemodule.cu:148-164,
                            impt_jax.py:736-782
  * HU -> RLSP              plan_impt.py:158-168 (np.interp, cast to float32)
  * spot sort / layers      impt_jax.py:797-859 (= importLayers IMPTClasses.cu:37-67)
  * ray-march, smoothing,   impt_jax.py:235-541 in float64, CUDA axis convention
    pencil-beam dose        (numpy axis 2 = x), see impt_oracle.c
  * plan loop               impt_jax.py:866-929 / plan_impt.py:170-206
  * gradient                analytic adjoint of the dose in float64

Pinning (tests/test_oracle_*.py): closed forms (uniform-density WET counts,
Gaussian normalisation, axis distance by cross product), independent brute
force (pure-Python loops, scipy map_coordinates 'grid-constant'), LUT values
straight from the CSV files, invariants (symmetry, linearity) and central
finite differences for the gradient.  The halo-width / halo-weight fit
(IMPTClasses.cu:104-121) has no independent source for its coefficients
(parity unpinned for them); what the mathematics fixes is pinned: finite and
continuous over the whole clamp range of all 196 energies (its poles at
w = R80 + 5 and w^2 = (R80 + 3)^2 outside it), the clamp plateaus, the
collapse of the range for R80 < 0.2 mm, and the two-Gaussian normalisation
with a non-zero halo weight.
"""

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(precision="f64"):
    name = "libimpt_oracle.so" if precision == "f64" else "libimpt_oracle_f32.so"
    if name not in _LIBS:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i, d, i64 = ctypes.c_int, ctypes.c_double, ctypes.c_int64
        L.orc_raymarch.argtypes = [P, i, i, i, d, P, P, i64, P]
        L.orc_smooth.argtypes = [P, i, i, i, d, P, P, i64, P]
        L.orc_wet_at.argtypes = [P, i, i, i, d, P, P, i64, P]
        L.orc_pb_dose.argtypes = [P, i, i, i, d, P, P, P, P, i, P, i, P, P, P, i, P, i64, P]
        L.orc_pb_grad.argtypes = [P, P, i, i, i, d, P, P, i, P, P, i, P, i, P, P, P, i, P, P]
        L.orc_halo.argtypes = [d, P, i64, P, P]
        for fn in (L.orc_raymarch, L.orc_smooth, L.orc_wet_at, L.orc_pb_dose, L.orc_pb_grad, L.orc_halo):
            fn.restype = ctypes.c_int
        _LIBS[name] = L
    return _LIBS[name]


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


# ---------------------------------------------------------------------------
# geometry
def sincos_deg(a):
    """sin/cos of degrees; whole multiples of 90 degrees give exact 0 / +-1 (the
    real-number values; float evaluation of sin(pi) would leave ~1e-16)."""
    a = math.fmod(float(a), 360.0)
    if a < 0:
        a += 360.0
    if a % 90.0 == 0.0:
        return {0: (0.0, 1.0), 1: (1.0, 0.0), 2: (0.0, -1.0), 3: (-1.0, 0.0)}[int(a // 90.0) % 4]
    return math.sin(math.radians(a)), math.cos(math.radians(a))


def beam_geometry(iso, origin, gantry, couch, vsadx, vsady):
    """float64 array [12]: iso - origin, source position, sin/cos(gantry + 180),
    sin/cos(couch), VSADX, VSADY (impt_jax.py:736-782, CudaClasses.cu:17-51)."""
    iso = np.asarray(iso, np.float64)[:3] - np.asarray(origin, np.float64)[:3]
    sg, cg = sincos_deg((float(gantry) + 180.0) % 360.0)
    sc, cc = sincos_deg(couch)
    sad = (float(vsadx) + float(vsady)) / 2.0
    xg, yg = -sad * sg, sad * cg
    return np.array([iso[0], iso[1], iso[2], xg * cc, yg, -xg * sc, sg, cg, sc, cc, float(vsadx), float(vsady)],
                    dtype=np.float64)


def image_to_head(geom, p):
    """numpy twin of the C transform, for tests."""
    sg, cg, sc, cc = geom[6:10]
    x, y, z = p
    xt = x * cc - z * sc
    zt = x * sc + z * cc
    xg = xt * cg + y * sg
    yg = -xt * sg + y * cg
    return np.array([-xg, zt, yg])


# ---------------------------------------------------------------------------
# spots and layers (impt_jax.py:797-859)
def spots_and_layers(spot_list, divergence_params):
    """Sort spots by integer energy id and build layers by the run scan of
    _extract_layer_data: energies 0..n_E-1 matched in turn from the current
    start (a negative id, sorting first, blocks every layer; ids >= n_E drop).

    Returns (order, spots [S,3] float64 sorted, layers int32 [L,3], r80 [L])."""
    sl = np.asarray(spot_list, dtype=np.float32)
    if sl.ndim == 1:
        sl = sl.reshape(1, -1)
    with np.errstate(invalid="ignore"):
        ids = sl[:, 3].astype(np.int32)
    order = np.argsort(ids, kind="stable")
    ids_s = ids[order]
    dvp = np.asarray(divergence_params, np.float32)
    n_e = dvp.shape[0]
    layers = []
    start = 0
    n = len(ids_s)
    for e in range(n_e):
        end = start
        while end < n and ids_s[end] == e:
            end += 1
        if end > start:
            layers.append((start, end - start, e))
            start = end
    lay = np.array(layers, dtype=np.int32).reshape(-1, 3)
    r80 = dvp[lay[:, 2], 1].astype(np.float64) if len(layers) else np.zeros(0)
    spots = sl[order, :3].astype(np.float64)
    return order, np.ascontiguousarray(spots), np.ascontiguousarray(lay), np.ascontiguousarray(r80)


# ---------------------------------------------------------------------------
# kernels
def raymarch(rlsp, spacing, geom, idx=None, precision="f64"):
    rho = np.ascontiguousarray(rlsp, dtype=np.float32)
    nz, ny, nx = rho.shape
    ix = None if idx is None else np.ascontiguousarray(idx, dtype=np.int64)
    n = rho.size if ix is None else ix.size
    out = np.empty(n, dtype=np.float64)
    g = np.ascontiguousarray(geom, np.float64)
    lib(precision).orc_raymarch(_p(rho), nx, ny, nz, float(spacing), _p(g), _p(ix), n, _p(out))
    return out.reshape(rho.shape) if ix is None else out


def smooth(raw, spacing, geom, idx=None, precision="f64"):
    r = np.ascontiguousarray(raw, dtype=np.float64)
    nz, ny, nx = r.shape
    ix = None if idx is None else np.ascontiguousarray(idx, dtype=np.int64)
    n = r.size if ix is None else ix.size
    out = np.empty(n, dtype=np.float64)
    g = np.ascontiguousarray(geom, np.float64)
    lib(precision).orc_smooth(_p(r), nx, ny, nz, float(spacing), _p(g), _p(ix), n, _p(out))
    return out.reshape(r.shape) if ix is None else out


def wet(rlsp, spacing, geom, precision="f64"):
    return smooth(raymarch(rlsp, spacing, geom, precision=precision), spacing, geom, precision=precision)


def wet_at(rlsp, spacing, geom, idx, precision="f64"):
    rho = np.ascontiguousarray(rlsp, dtype=np.float32)
    nz, ny, nx = rho.shape
    ix = np.ascontiguousarray(idx, dtype=np.int64)
    out = np.empty(ix.size, dtype=np.float64)
    g = np.ascontiguousarray(geom, np.float64)
    lib(precision).orc_wet_at(_p(rho), nx, ny, nz, float(spacing), _p(g), _p(ix), ix.size, _p(out))
    return out


class ModelArrays:
    """float32 contiguous copies of an IMPTBeamModel's tables."""

    def __init__(self, model):
        self.dvp = np.ascontiguousarray(model.divergence_params, np.float32)
        self.depths = np.ascontiguousarray(model.lut_depths, np.float32)
        self.sigmas = np.ascontiguousarray(model.lut_sigmas, np.float32)
        self.idds = np.ascontiguousarray(model.lut_idds, np.float32)
        self.vsadx = float(model.VSADX)
        self.vsady = float(model.VSADY)


def pb_dose(wet_cm, spacing, geom, spots, layers, r80, marr, idx=None, precision="f64"):
    w = np.ascontiguousarray(wet_cm, dtype=np.float64)
    nz, ny, nx = w.shape
    ix = None if idx is None else np.ascontiguousarray(idx, dtype=np.int64)
    n = w.size if ix is None else ix.size
    out = np.empty(n, dtype=np.float64)
    g = np.ascontiguousarray(geom, np.float64)
    sp = np.ascontiguousarray(spots, np.float64)
    la = np.ascontiguousarray(layers, np.int32)
    r8 = np.ascontiguousarray(r80, np.float64)
    lib(precision).orc_pb_dose(_p(w), nx, ny, nz, float(spacing), _p(g), _p(sp), _p(la), _p(r8), la.shape[0],
                               _p(marr.dvp), marr.dvp.shape[1], _p(marr.depths), _p(marr.sigmas), _p(marr.idds),
                               marr.depths.shape[1], _p(ix), n, _p(out))
    return out.reshape(w.shape) if ix is None else out


def pb_grad(wet_cm, grad_dose, spacing, geom, spots, layers, r80, marr, precision="f64"):
    """(dL/dmu [S], dL/dxy [S, 2]) in sorted spot order."""
    w = np.ascontiguousarray(wet_cm, dtype=np.float64)
    gv = np.ascontiguousarray(grad_dose, dtype=np.float64)
    nz, ny, nx = w.shape
    g = np.ascontiguousarray(geom, np.float64)
    sp = np.ascontiguousarray(spots, np.float64)
    la = np.ascontiguousarray(layers, np.int32)
    r8 = np.ascontiguousarray(r80, np.float64)
    S = sp.shape[0]
    gmu = np.zeros(S)
    gxy = np.zeros((S, 2))
    lib(precision).orc_pb_grad(_p(w), _p(gv), nx, ny, nz, float(spacing), _p(g), _p(sp), S, _p(la), _p(r8), la.shape[0],
                      _p(marr.dvp), marr.dvp.shape[1], _p(marr.depths), _p(marr.sigmas), _p(marr.idds),
                      marr.depths.shape[1], _p(gmu), _p(gxy))
    return gmu, gxy


def halo(r80, wet_mm, precision="f64"):
    """(sigma_n [mm], u_n) of the nuclear-halo fit at the depths wet_mm (mm) for
    an energy of range r80 (mm): the term voxel_terms uses (test hook)."""
    w = np.ascontiguousarray(np.atleast_1d(wet_mm), dtype=np.float64)
    sn = np.empty(w.size)
    un = np.empty(w.size)
    lib(precision).orc_halo(float(r80), _p(w), w.size, _p(sn), _p(un))
    return sn, un


# ---------------------------------------------------------------------------
# plan level (computeIMPTPlanJax, impt_jax.py:866-929, in CUDA axis order)
def rlsp_from_hu(hu, machine_dir):
    tab = np.genfromtxt(os.path.join(machine_dir, "HU_RLSP.csv"), delimiter=",", names=True)
    return np.array(np.interp(hu, tab["HU"], tab["RLSP"]), dtype=np.single)


def beam_model_for(plan, beam):
    labels = [str(v) for v in np.asarray(plan.dicom_rangeshifter_label)]
    return plan.beam_models[labels.index(beam.dicom_rangeshifter_label)]


def compute_plan(dose_grid, plan, machine_dir, return_wet=False):
    """Total dose (float64, (z, y, x)) times n_fractions, plus per-beam doses."""
    if not (dose_grid.spacing[0] == dose_grid.spacing[1] == dose_grid.spacing[2]):
        raise ValueError("Spacing must be isotropic for IMPT dose calculation")
    rlsp = rlsp_from_hu(dose_grid.HU, machine_dir)
    s = float(np.float32(dose_grid.spacing[0]))
    origin = np.asarray(dose_grid.origin, np.float32)
    total = np.zeros(rlsp.shape)
    beams, wets = [], []
    for beam in plan.beam_list:
        model = beam_model_for(plan, beam)
        geom = beam_geometry(np.asarray(beam.iso, np.float32), origin, beam.gantry_angle, beam.couch_angle,
                             model.VSADX, model.VSADY)
        w = wet(rlsp, s, geom)
        _, spots, layers, r80 = spots_and_layers(beam.spot_list, model.divergence_params)
        d = pb_dose(w, s, geom, spots, layers, r80, ModelArrays(model))
        beams.append(d * plan.n_fractions)
        wets.append(w)
        total += d
    total *= plan.n_fractions
    return (total, beams, wets) if return_wet else (total, beams)


# ---------------------------------------------------------------------------
# comparison metrics (tests/compare_cuda_jax.py:56-109 of the reference, plus
# max |dD| / D_max and a gamma index)
def compare(test, ref):
    test = np.asarray(test, np.float64)
    ref = np.asarray(ref, np.float64)
    dmax = float(ref.max()) if ref.size else 0.0
    diff = np.abs(test - ref)
    return {
        "max_abs": float(diff.max()) if diff.size else 0.0,
        "mean_abs": float(diff.mean()) if diff.size else 0.0,
        "max_rel_to_dmax": float(diff.max() / dmax) if dmax > 0 else 0.0,
        "within_1pct": float(np.mean(diff <= 0.01 * dmax)) if dmax > 0 else 1.0,
        "within_5pct": float(np.mean(diff <= 0.05 * dmax)) if dmax > 0 else 1.0,
    }


def gamma_pass_rate(test, ref, spacing, dose_pct=2.0, dta_mm=2.0, threshold_pct=10.0):
    """Global gamma (dose_pct of D_max, dta_mm) pass rate over reference voxels
    above threshold_pct of D_max; exhaustive search over the DTA ball on the
    voxel grid (no sub-voxel interpolation, which can only lower gamma)."""
    test = np.asarray(test, np.float64)
    ref = np.asarray(ref, np.float64)
    dmax = ref.max()
    if dmax <= 0:
        return 1.0
    dd = dose_pct / 100.0 * dmax
    r = int(math.ceil(dta_mm / spacing))
    mask = ref >= threshold_pct / 100.0 * dmax
    best = np.full(ref.shape, np.inf)
    pad = np.pad(test, r, mode="edge")
    nz, ny, nx = ref.shape
    for dz in range(-r, r + 1):
        for dy in range(-r, r + 1):
            for dx in range(-r, r + 1):
                dist2 = (dx * dx + dy * dy + dz * dz) * spacing * spacing
                if dist2 > dta_mm * dta_mm * (1 + 1e-9) and (dx or dy or dz):
                    continue
                shifted = pad[r + dz:r + dz + nz, r + dy:r + dy + ny, r + dx:r + dx + nx]
                g2 = dist2 / (dta_mm * dta_mm) + ((shifted - ref) / dd) ** 2
                np.minimum(best, g2, out=best)
    return float(np.mean(best[mask] <= 1.0)) if mask.any() else 1.0
