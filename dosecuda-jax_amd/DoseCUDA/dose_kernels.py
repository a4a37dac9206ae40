"""Drop-in for the reference's ``dose_kernels`` extension module (proton part).

``proton_raytrace_cuda(model, vol, beam, gpu_id)`` and
``proton_spot_cuda(model, vol, wet, beam, gpu_id)`` keep the names, argument
meaning, attribute protocol, return values and exceptions of
DoseCUDA/dose_kernels/dosemodule.cu:104-374:

* attributes are read by name; a missing one raises AttributeError, a
  non-numeric float TypeError, and an array that is not a numpy float32 array
  of the required rank raises
  ``ValueError("'<attr>' must be <n>-dimensional and of type float.")``
  (dosemodule.cu:50-66);
* outputs are new C-contiguous float32 numpy arrays shaped like the input
  volume; inputs are not modified;
* device failures raise RuntimeError("CUDA error: ...") (dosemodule.cu:193-195).

The work runs on the HIP device ``gpu_id`` through libdosehip.so, on torch's
current stream of that device; the call returns once the result is on the host,
as the reference's blocking copy does.

``photon_dose_cuda(model, vol, cp, gpu_id)`` does the same for the photon
collapsed-cone dose of one control point (photon_dose, dosemodule.cu:377-537).
"""

import numpy as np
import torch

from . import _native as N
from . import device as D
from . import photon as PH

_native = N.load()  # fail at import if the HIP library is missing


def _getfloat(obj, attr):
    # PyFloat_AsDouble semantics: numbers (and objects with __float__) only
    val = getattr(obj, attr)
    if isinstance(val, (str, bytes, bytearray)):
        raise TypeError("must be real number, not %s" % type(val).__name__)
    return float(val)


def _getarray(obj, attr, ndim):
    arr = getattr(obj, attr)
    if not (isinstance(arr, np.ndarray) and arr.ndim == ndim and arr.dtype == np.float32):
        raise ValueError("'%s' must be %d-dimensional and of type float." % (attr, ndim))
    return arr


def _device(gpu_id):
    if not torch.cuda.is_available():
        raise RuntimeError("CUDA error: no HIP device available")
    gpu_id = int(gpu_id)
    if gpu_id < 0 or gpu_id >= torch.cuda.device_count():
        raise RuntimeError("CUDA error: invalid device ordinal %d" % gpu_id)
    return torch.device("cuda", gpu_id)


def proton_raytrace_cuda(model, vol, beam, gpu_id):
    """Smoothed water-equivalent thickness (cm) of every voxel for one beam
    (proton_raytrace, dosemodule.cu:104-201)."""
    vsadx = _getfloat(model, "VSADX")
    vsady = _getfloat(model, "VSADY")
    rlsp = _getarray(vol, "voxel_data", 3)
    spacing = _getarray(vol, "spacing", 1)
    origin = _getarray(vol, "origin", 1)
    iso = _getarray(beam, "iso", 1)
    ga = _getfloat(beam, "gantry_angle")
    ta = _getfloat(beam, "couch_angle")

    dev = _device(gpu_id)
    geo = D.make_geom(D.beam_geometry(iso, origin, ga, ta, vsadx, vsady))
    grid = D.make_grid(rlsp.shape, spacing[0])
    with torch.cuda.device(dev):
        d_rlsp = torch.from_numpy(np.ascontiguousarray(rlsp)).to(dev)
        padded = D.rlsp_pad(d_rlsp, grid, dev)
        _, wet = D.wet_from_padded(padded, grid, geo, dev)
        out = wet.cpu().numpy()
    return np.ascontiguousarray(out)


def proton_spot_cuda(model, vol, wet, beam, gpu_id, cutoff_sigma=0.0):
    """Pencil-beam dose of one beam from its WET volume (proton_spot,
    dosemodule.cu:243-374).  Not multiplied by the number of fractions.

    By default every spot is evaluated for every voxel with WET <= 1.1 R80, as
    the reference's pencilBeamKernel does (IMPTClasses.cu:255-267).
    ``cutoff_sigma`` (keyword, not in the reference) > 0 skips spot/voxel
    pairs farther than that many halo sigmas apart (each skipped term is below
    exp(-cutoff_sigma^2 / 2) of its spot's peak)."""
    vsadx = _getfloat(model, "VSADX")
    vsady = _getfloat(model, "VSADY")
    dvp = _getarray(model, "divergence_params", 2)
    _getarray(model, "lut_depths", 2)
    _getarray(model, "lut_sigmas", 2)
    _getarray(model, "lut_idds", 2)
    _getarray(vol, "voxel_data", 3)
    spacing = _getarray(vol, "spacing", 1)
    origin = _getarray(vol, "origin", 1)
    wet_arr = _getarray(wet, "voxel_data", 3)
    iso = _getarray(beam, "iso", 1)
    spots = _getarray(beam, "spot_list", 2)
    ga = _getfloat(beam, "gantry_angle")
    ta = _getfloat(beam, "couch_angle")
    if spots.shape[1] != 4:
        raise ValueError("'spot_list' must be 2-dimensional and of type float.")

    dev = _device(gpu_id)
    geo = D.make_geom(D.beam_geometry(iso, origin, ga, ta, vsadx, vsady))
    grid = D.make_grid(wet_arr.shape, spacing[0])
    with torch.cuda.device(dev):
        dmodel = D.DeviceBeamModel(model, dev)
        dspots = D.DeviceSpots(spots, dvp, dev)
        d_wet = torch.from_numpy(np.ascontiguousarray(wet_arr)).to(dev)
        dose = torch.empty_like(d_wet)
        ws, nbytes = D.alloc_workspace(dspots.n_spots, dspots.n_layers, grid, dev)
        D.pb_forward(d_wet, dose, grid, geo, dmodel, dspots, ws, nbytes, cutoff_sigma=cutoff_sigma)
        out = dose.cpu().numpy()
    return np.ascontiguousarray(out)


def _getbool(obj, attr):
    # PyObject_IsTrue semantics (dosemodule.cu:91-101)
    return bool(getattr(obj, attr))


def photon_dose_cuda(model, vol, cp, gpu_id):
    """Collapsed-cone dose of one control point (photon_dose,
    dosemodule.cu:377-537): radiological depth through the density volume,
    TERMA and electron contamination through the leaf pairs, then the sum over
    6 cones x 12 azimuths.  Not multiplied by the output factor.

    ``model.kernel`` ([6, 6], one row per cone) and ``cp.mlc`` ([n, 4], one row
    (x1, x2, y_offset, y_width) per leaf pair) are read by index.  The
    reference reads their raw buffers column-wise, which gives the same values
    for the Fortran-ordered arrays IMRTPlan builds (the pandas table and the
    transposed leaf stack, dosemodule.cu:228-240, IMRTClasses.cu:292-297)."""
    for attr in ("mu_calibration", "primary_source_distance", "scatter_source_distance", "primary_source_size",
                 "scatter_source_size", "mlc_distance", "scatter_source_weight", "electron_attenuation",
                 "electron_source_weight", "electron_fitted_dmax", "jaw_transmission", "mlc_transmission"):
        _getfloat(model, attr)
    _getbool(model, "has_xjaws")
    _getbool(model, "has_yjaws")
    for attr in ("profile_radius", "profile_intensities", "profile_softening", "spectrum_attenuation_coefficients",
                 "spectrum_primary_weights", "spectrum_scatter_weights"):
        _getarray(model, attr, 1)
    kern = _getarray(model, "kernel", 2)
    density = _getarray(vol, "voxel_data", 3)
    spacing = _getarray(vol, "spacing", 1)
    origin = _getarray(vol, "origin", 1)
    _getarray(cp, "iso", 1)
    mlc = _getarray(cp, "mlc", 2)
    mu = _getfloat(cp, "mu")
    for attr in ("ga", "ca", "ta"):
        _getfloat(cp, attr)
    if kern.shape != (6, 6):
        raise ValueError("'kernel' must be 2-dimensional and of type float.")
    if mlc.shape[1] != 4:
        raise ValueError("'mlc' must be 2-dimensional and of type float.")

    dev = _device(gpu_id)
    with torch.cuda.device(dev):
        dmodel = PH.DevicePhotonModel(model, dev)
        beam = PH.control_point(cp, origin, dmodel, dev, mu=mu)
        ws = PH.PhotonWorkspace(density.shape, float(spacing[0]), dev)
        ws.set_density(torch.from_numpy(np.ascontiguousarray(density)).to(dev))
        dose = torch.empty(density.shape, dtype=torch.float32, device=dev)
        PH.cp_dose(ws, dmodel, beam, dose)
        out = dose.cpu().numpy()
    return np.ascontiguousarray(out)
