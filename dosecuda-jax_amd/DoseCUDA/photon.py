"""Device-resident photon collapsed-cone dose over the C ABI of libdosehip.so.

* ``DevicePhotonModel``  an IMRTPhotonEnergy's tables on the device
  (IMRTBeam::Model, IMRTClasses.cuh:22-58; marshalled like dosemodule.cu:377-532)
* ``control_point``      a control point's dh_photon_beam (geometry with the
  collimator, leaf pairs on the device, meterset)
* ``PhotonWorkspace``    record volume (density + TERMA) and the depth, TERMA
  and electron volumes of one grid
* ``cp_dose``            depth -> TERMA -> collapsed-cone dose of one control
  point into a dose volume (photon_dose_cuda, IMRTClasses.cu:353-405)
* ``PhotonDoseEngine``   a whole plan on the device: density from HU once,
  every control point's dose times its output factor accumulated per beam
  (computeIMRTPlan, plan_imrt.py:126-158)

Everything stays in device memory and is enqueued on torch's current stream
(the engine's pipelined plan also on a side stream, ordered with the current
stream by events); nothing here synchronises.
"""

import numpy as np
import torch

from . import _native as N
from .geometry import beam_geometry, sincos_deg
from .device import make_grid, make_geom


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class DevicePhotonModel:
    """Profile and spectrum tables of one photon energy uploaded once; the six
    cones go into the dh_photon_model by value.  ``kernel`` is read by index
    as [cone, (theta, Am, am, Bm, bm, ray_length)]."""

    def __init__(self, model, device):
        prof = np.stack([_f32(model.profile_radius), _f32(model.profile_intensities), _f32(model.profile_softening)])
        spec = np.stack([_f32(model.spectrum_attenuation_coefficients), _f32(model.spectrum_primary_weights),
                         _f32(model.spectrum_scatter_weights)])
        kern = np.asarray(model.kernel, dtype=np.float32)
        if kern.shape != (6, 6):
            raise ValueError("'kernel' must be 2-dimensional and of type float.")
        if prof.shape[1] < 1:
            raise ValueError("'profile_radius' must be 1-dimensional and of type float.")
        self.device = device
        self.profile = torch.from_numpy(np.ascontiguousarray(prof)).to(device)
        self.spectrum = torch.from_numpy(np.ascontiguousarray(spec)).to(device)
        c = N.dh_photon_model()
        c.profile = N.ptr(self.profile).value
        c.spectrum = N.ptr(self.spectrum).value
        c.n_profile = prof.shape[1]
        c.n_spectrum = spec.shape[1]
        for i in range(6):
            for j in range(6):
                c.kernel[i][j] = float(kern[i, j])
        c.mu_cal = float(model.mu_calibration)
        c.primary_src_dist = float(model.primary_source_distance)
        c.scatter_src_dist = float(model.scatter_source_distance)
        c.primary_src_size = float(model.primary_source_size)
        c.scatter_src_size = float(model.scatter_source_size)
        c.mlc_distance = float(model.mlc_distance)
        c.scatter_src_weight = float(model.scatter_source_weight)
        c.electron_attenuation = float(model.electron_attenuation)
        c.electron_src_weight = float(model.electron_source_weight)
        c.electron_fitted_dmax = float(model.electron_fitted_dmax)
        self.c = c
        self.primary_source_distance = float(model.primary_source_distance)


def photon_geometry(iso, origin, gantry, couch, collimator, sad):
    """dh_beam_geom plus (sin, cos) of the collimator: the CudaBeam of an
    IMRTBeam (source at the primary source distance, gantry + 180 deg,
    IMRTClasses.cu:30-39, dosemodule.cu:497-509)."""
    geo = make_geom(beam_geometry(iso, origin, gantry, couch, sad, sad))
    return geo, sincos_deg(float(collimator))


class control_point:
    """dh_photon_beam of one control point; keeps its leaf pairs on the device."""

    def __init__(self, cp, origin, dmodel, device, mu=None):
        mlc = np.asarray(cp.mlc, dtype=np.float32)
        if mlc.ndim != 2 or mlc.shape[1] != 4:
            raise ValueError("'mlc' must be 2-dimensional and of type float.")
        self.mlc = torch.from_numpy(np.ascontiguousarray(mlc)).to(device)
        geo, (s, c) = photon_geometry(np.asarray(cp.iso, np.float32), np.asarray(origin, np.float32), float(cp.ga),
                                      float(cp.ta), float(cp.ca), dmodel.primary_source_distance)
        b = N.dh_photon_beam()
        b.geom = geo
        b.sin_col, b.cos_col = s, c
        b.mlc = N.ptr(self.mlc).value if mlc.shape[0] else None
        b.n_pairs = mlc.shape[0]
        b.mu = float(cp.mu if mu is None else mu)
        self.c = b


class PhotonWorkspace:
    """Record volume and per-control-point volumes of one grid (``depth``: a
    depth volume to share with another workspace of the same grid)."""

    def __init__(self, shape_zyx, spacing, device, depth=None):
        self.grid = make_grid(shape_zyx, spacing)
        n = N.load().dh_photon_record_len(self.grid)
        N.check(0 if n > 0 else N.DH_ERR_ARG, "dh_photon_record_len")
        self.device = device
        # 32-byte aligned: torch's caching allocator hands out 512-byte aligned blocks
        self.records = torch.empty(int(n), dtype=torch.float32, device=device)
        self.depth = torch.empty(tuple(shape_zyx), dtype=torch.float32, device=device) if depth is None else depth
        self.terma = torch.empty(tuple(shape_zyx), dtype=torch.float32, device=device)
        self.electron = torch.empty(tuple(shape_zyx), dtype=torch.float32, device=device)
        self.shape = tuple(shape_zyx)

    def set_density(self, density):
        """Density half of the records from a device density volume."""
        L = N.load()
        N.check(L.dh_photon_density(N.ptr(density), N.ptr(self.records), self.grid, N.stream_handle(self.device)),
                "dh_photon_density")

    def set_density_from_hu(self, hu, hu_knots, density_knots):
        L = N.load()
        hk = np.ascontiguousarray(hu_knots, np.float64)
        dk = np.ascontiguousarray(density_knots, np.float64)
        N.check(L.dh_photon_density_from_hu(N.ptr(hu), N.ptr(self.records), self.grid,
                                            hk.ctypes.data_as(N.ctypes.POINTER(N.ctypes.c_double)),
                                            dk.ctypes.data_as(N.ctypes.POINTER(N.ctypes.c_double)), len(hk),
                                            N.stream_handle(self.device)), "dh_photon_density_from_hu")


def raytrace(ws, beam):
    """Radiological depth of every voxel (rayTraceKernel on the density)."""
    N.check(N.load().dh_photon_raytrace(N.ptr(ws.records), N.ptr(ws.depth), ws.grid, N.ctypes.byref(beam.c),
                                        N.stream_handle(ws.device)), "dh_photon_raytrace")


def terma(ws, dmodel, beam):
    """TERMA and electron contamination of every voxel, then the TERMA half of the records."""
    L = N.load()
    st = N.stream_handle(ws.device)
    N.check(L.dh_photon_terma(N.ptr(ws.depth), N.ptr(ws.terma), N.ptr(ws.electron), ws.grid,
                              N.ctypes.byref(dmodel.c), N.ctypes.byref(beam.c), st), "dh_photon_terma")
    N.check(L.dh_photon_terma_records(N.ptr(ws.terma), N.ptr(ws.records), ws.grid, st), "dh_photon_terma_records")


def ccc(ws, dmodel, beam, dose, scale=1.0, accumulate=False):
    """Collapsed-cone sum into ``dose`` (dose = scale * D, or += with accumulate)."""
    N.check(N.load().dh_photon_ccc(N.ptr(ws.records), N.ptr(ws.terma), N.ptr(ws.electron), N.ptr(dose), ws.grid,
                                   N.ctypes.byref(dmodel.c), N.ctypes.byref(beam.c), float(scale),
                                   1 if accumulate else 0, N.stream_handle(ws.device)), "dh_photon_ccc")


def cp_dose(ws, dmodel, beam, dose, scale=1.0, accumulate=False):
    """Depth, TERMA / electron, TERMA records, then the collapsed-cone sum of
    one control point into ``dose`` (dose = scale * D, or += with accumulate)."""
    raytrace(ws, beam)
    terma(ws, dmodel, beam)
    ccc(ws, dmodel, beam, dose, scale, accumulate)


class PhotonDoseEngine:
    """A photon plan on one device: density from HU on the device (np.interp
    of HU_Density.csv in float64, rounded to float32), then per beam the sum
    over its control points of output factor x control-point dose, and the
    plan dose = sum of beams x n_fractions (computeIMRTPlan, plan_imrt.py:126-158).

    pipeline=True keeps two workspaces (two record volumes, one depth volume)
    and computes the depth and TERMA of control point i + 1 on a side stream
    while the cone sum of control point i runs on the caller's stream: the
    TERMA kernel (VALU-bound) and the ray-trace fill issue slots the cone sum
    (bound by its gathers) leaves idle.  The cone sums still run one after the
    other in plan order on the caller's stream, so the doses are the same bit
    for bit as with pipeline=False (one workspace, everything on the caller's
    stream), which costs one record volume less (12 floats per padded cell)."""

    def __init__(self, dose_grid, plan, gpu_id=0, pipeline=True):
        from .plan_imrt import _photon_dir
        import pandas as pd
        if not (dose_grid.spacing[0] == dose_grid.spacing[1] == dose_grid.spacing[2]):
            raise Exception("Spacing must be isotropic for IMPT dose calculation - consider resampling CT")
        self.device = torch.device("cuda", int(gpu_id))
        self.plan = plan
        self.origin = np.asarray(dose_grid.origin, np.float32)
        self.shape = tuple(np.asarray(dose_grid.HU).shape)
        spacing = float(np.float32(dose_grid.spacing[0]))
        df = pd.read_csv(_photon_dir(plan.machine_name, "HU_Density.csv"))
        with torch.cuda.device(self.device):
            self.hu = torch.from_numpy(_f32(dose_grid.HU)).to(self.device)
            self.ws = PhotonWorkspace(self.shape, spacing, self.device)
            self.wss = [self.ws]
            if pipeline:
                self.wss.append(PhotonWorkspace(self.shape, spacing, self.device, depth=self.ws.depth))
                self.side = torch.cuda.Stream(self.device)
            for ws in self.wss:
                ws.set_density_from_hu(self.hu, df["HU"].to_numpy(), df["Density"].to_numpy())
            labels = list(np.asarray(plan.dicom_energy_label).astype(str))
            self.models = {}
            self.beams = []
            for beam in plan.beam_list:
                li = labels.index(beam.dicom_energy_label)
                if li not in self.models:
                    self.models[li] = DevicePhotonModel(plan.beam_models[li], self.device)
                model = plan.beam_models[li]
                cps = [(control_point(cp, self.origin, self.models[li], self.device), float(model.outputFactor(cp)),
                        _depth_key(cp)) for cp in beam.cp_list]
                self.beams.append((self.models[li], cps))
            self.beam_doses = [torch.zeros(self.shape, dtype=torch.float32, device=self.device)
                               for _ in self.beams]
            self.dose = torch.zeros(self.shape, dtype=torch.float32, device=self.device)

    def compute(self, pipeline=None):
        """Beam doses and the plan dose (device tensors), ordered on the
        current stream (pipeline: None = as constructed, False = one workspace
        on the current stream).  The radiological depth depends on the
        isocentre, gantry and couch angles only (not on the leaves, collimator
        or meterset), so consecutive control points with the same three reuse
        it: a step-and-shoot beam traces its rays once."""
        piped = len(self.wss) > 1 if pipeline is None else (bool(pipeline) and len(self.wss) > 1)
        with torch.cuda.device(self.device):
            cur = torch.cuda.current_stream(self.device)
            if piped:
                # the side stream starts after the caller's earlier work (which
                # may still read the workspaces)
                self.side.wait_stream(cur)
            done = [None, None]   # cone sum that last read workspace b
            self.dose.zero_()
            last = None
            i = 0
            for (dmodel, cps), bd in zip(self.beams, self.beam_doses):
                bd.zero_()
                for cpc, of, key in cps:
                    b = i & 1 if piped else 0
                    ws = self.wss[b]
                    if piped:
                        with torch.cuda.stream(self.side):
                            if done[b] is not None:
                                self.side.wait_event(done[b])
                            if key != last:
                                raytrace(ws, cpc)
                            terma(ws, dmodel, cpc)
                            ready = torch.cuda.Event()
                            ready.record(self.side)
                        cur.wait_event(ready)
                    else:
                        if key != last:
                            raytrace(ws, cpc)
                        terma(ws, dmodel, cpc)
                    last = key
                    ccc(ws, dmodel, cpc, bd, scale=of, accumulate=True)
                    if piped:
                        done[b] = torch.cuda.Event()
                        done[b].record(cur)
                    i += 1
                self.dose.add_(bd)
            self.dose.mul_(float(self.plan.n_fractions))
            if piped:
                cur.wait_stream(self.side)
        return self.dose, self.beam_doses

    def depth_traces(self):
        """Ray-traces one compute() runs (changes of the depth key)."""
        keys = [key for _, cps in self.beams for _, _, key in cps]
        return sum(1 for i, k in enumerate(keys) if i == 0 or k != keys[i - 1])

    def control_points(self):
        return sum(len(cps) for _, cps in self.beams)


def _depth_key(cp):
    """What the radiological depth of a control point depends on: isocentre,
    gantry and couch angle (CudaBeam, CudaClasses.cu:17-51)."""
    return (tuple(float(v) for v in np.asarray(cp.iso, np.float32)[:3]), float(cp.ga), float(cp.ta))


def cone_steps(model, spacing):
    """Samples per ray of each cone: the float32 counter L, L - sp, ... >= 0
    with sp = spacing / 10 (cccKernel, IMRTClasses.cu:313-335)."""
    sp = np.float32(np.float32(spacing) / np.float32(10.0))
    out = []
    for L in np.asarray(model.kernel, np.float32)[:, 5]:
        n, rl = 0, np.float32(L)
        while rl >= 0:
            n += 1
            rl = np.float32(rl - sp)
        out.append(n)
    return out


def samples_per_voxel(model, spacing):
    """Trilinear samples (TERMA and density pairs) of one in-field voxel: 12
    azimuths x the steps of each cone."""
    return 12 * int(sum(cone_steps(model, spacing)))

