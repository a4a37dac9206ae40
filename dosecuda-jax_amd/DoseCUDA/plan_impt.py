"""IMPT (proton pencil-beam scanning) plan API.

Re-creates the proton part of DoseCUDA/plan_impt.py of the reference with the
same class, method and attribute names:

* ``IMPTBeamModel``  (plan_impt.py:63-149)  beam-model lookup tables
* ``IMPTDoseGrid``   (plan_impt.py:152-287) HU->RLSP and the per-beam dose loop
* ``IMPTBeam``       (plan_impt.py:290-327) spot lists
* ``IMPTPlan``       (plan_impt.py:330-351) range-shifter models of a machine

The dose calls go through :mod:`DoseCUDA.dose_kernels`, whose functions keep
the reference extension's names and signatures but run hand-written gfx950
HIP kernels.  The RT-Ion-Plan DICOM reader/writer of the reference
(plan_impt.py:353-674) is not part of the dose path and is not provided.
"""

import os
import sys

import numpy as np
import pandas as pd

from .plan import Plan, Beam, DoseGrid, VolumeObject  # noqa: F401  (re-exported names)
from . import imageio
from . import dose_kernels

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))

# Rows of depth-dose / sigma table kept per energy.  The reference keeps the
# first 400 rows (depth 0..399 mm) and drops the row at 400 mm
# (plan_impt.py:118-121, LUT_LENGTH in IMPTClasses.cuh:6).
LUT_ROWS = 400


def _data_path(*parts):
    return os.path.join(_PKG_DIR, *parts)


def _read_machine_geometry(path):
    """Return (VSADX, VSADY) in mm from machine_geometry.csv (plan_impt.py:72-86)."""
    vsad = {}
    with open(path, "r") as fh:
        for line in fh:
            key, _, val = line.partition(",")
            key = key.strip()
            if key in ("VSADX", "VSADY"):
                vsad[key] = float(val)
    for key in ("VSADX", "VSADY"):
        if key not in vsad:
            raise Exception("%s not found in machine_geometry.csv" % key)
    return vsad["VSADX"], vsad["VSADY"]


def _read_energy_table(path):
    """Parse one energy_NNN.csv.

    Layout: header line, one line ``E,R80,A,B,C``, a blank line, a header, then
    rows ``depth,sigma,idd``; only the first ``LUT_ROWS`` rows are kept.
    Returns (divergence_params[5], depths, sigmas, idds).
    """
    with open(path, "r") as fh:
        lines = fh.read().splitlines()
    params = [float(v) for v in lines[1].split(",")]
    rows = []
    for line in lines[4:]:
        if len(rows) == LUT_ROWS:
            break
        fields = line.split(",")
        rows.append((float(fields[0]), float(fields[1]), float(fields[2])))
    tab = np.asarray(rows, dtype=np.float64)
    return params, tab[:, 0], tab[:, 1], tab[:, 2]


class IMPTBeamModel():
    """Lookup tables of one range-shifter configuration (plan_impt.py:63-149).

    Attributes read by the dose kernels (dosemodule.cu:116-118, 261-265):
    ``VSADX``, ``VSADY`` (float, mm); ``divergence_params`` float32 [n_E, 5]
    = (energy MeV, R80 mm, A, B, C) with sigma_air(d) = A d^2 + B d + C;
    ``lut_depths``, ``lut_sigmas``, ``lut_idds`` float32 [n_E, 400].
    """

    def __init__(self, dicom_rangeshifter_label, path_to_model, folder_rangeshifter_label):
        self.dicom_rangeshifter_label = dicom_rangeshifter_label

        model_dir = path_to_model if os.path.isabs(path_to_model) else _data_path(path_to_model)
        self.VSADX, self.VSADY = _read_machine_geometry(os.path.join(model_dir, "machine_geometry.csv"))

        rs_dir = os.path.join(model_dir, folder_rangeshifter_label)
        self.energy_table = pd.read_csv(os.path.join(rs_dir, "energies.csv"))
        self.energy_labels = self.energy_table["energy_label"].to_numpy()
        energy_ids = self.energy_table["index"].to_numpy()
        self.reference_sigmas = self.energy_table["sigma"].to_numpy()

        dvp, dep, sig, idd = [], [], [], []
        for eid in energy_ids:
            p, d, s, i = _read_energy_table(os.path.join(rs_dir, "energy_%03d.csv" % eid))
            dvp.append(p)
            dep.append(d)
            sig.append(s)
            idd.append(i)
        self.divergence_params = np.array(dvp, dtype=np.single)
        self.lut_depths = np.array(dep, dtype=np.single)
        self.lut_sigmas = np.array(sig, dtype=np.single)
        self.lut_idds = np.array(idd, dtype=np.single)

    def energyIDFromLabel(self, energy_label):
        """Row index of the energy whose label equals ``energy_label``
        (plan_impt.py:142-149)."""
        hit = self.energy_table.index[self.energy_table["energy_label"] == energy_label]
        if len(hit) != 1:
            raise Exception("Energy ID not found for %.2f" % energy_label)
        return hit[0]


def _model_for_beam(plan, beam):
    """Beam model whose DICOM range-shifter label matches the beam's.

    As in the reference (plan_impt.py:187-191) an unknown label prints a message
    and exits the process.
    """
    labels = list(np.asarray(plan.dicom_rangeshifter_label).astype(str))
    try:
        idx = labels.index(beam.dicom_rangeshifter_label)
    except ValueError:
        print("Beam model not found for rangeshifter ID %s" % beam.dicom_rangeshifter_label)
        sys.exit(1)
    return plan.beam_models[idx]


class IMPTDoseGrid(DoseGrid):
    """Dose grid with the proton dose entry points (plan_impt.py:152-287)."""

    def __init__(self):
        super().__init__()
        self.RLSP = []

    def RLSPFromHU(self, machine_name):
        """Relative linear stopping power from HU: piecewise-linear through the
        machine's HU_RLSP.csv knots, clamped at both ends (np.interp), returned as
        float32 (plan_impt.py:158-168)."""
        table = pd.read_csv(_data_path("lookuptables", "protons", machine_name, "HU_RLSP.csv"))
        hu_knots = table["HU"].to_numpy()
        rlsp_knots = table["RLSP"].to_numpy()
        return np.array(np.interp(self.HU, hu_knots, rlsp_knots), dtype=np.single)

    def _volume(self, data):
        vol = VolumeObject()
        vol.voxel_data = np.array(data, dtype=np.single)
        vol.origin = np.array(self.origin, dtype=np.single)
        vol.spacing = np.array(self.spacing, dtype=np.single)
        return vol

    def _check_isotropic(self):
        if self.spacing[0] != self.spacing[1] or self.spacing[0] != self.spacing[2]:
            raise Exception("Spacing must be isotropic for IMPT dose calculation - consider resampling CT")

    def computeIMPTPlan(self, plan, gpu_id=0):
        """Per beam: WET ray-march, then spot dose; ``beam_doses`` keeps each beam
        times ``n_fractions`` and ``dose`` their sum times ``n_fractions``
        (plan_impt.py:170-206)."""
        self.beam_doses = []
        self.dose = np.zeros(self.size, dtype=np.single)
        self.RLSP = self.RLSPFromHU(plan.machine_name)
        self._check_isotropic()
        rlsp_object = self._volume(self.RLSP)

        for beam in plan.beam_list:
            beam_model = _model_for_beam(plan, beam)
            beam_wet = dose_kernels.proton_raytrace_cuda(beam_model, rlsp_object, beam, gpu_id)
            wet_object = self._volume(beam_wet)
            beam_dose = dose_kernels.proton_spot_cuda(beam_model, rlsp_object, wet_object, beam, gpu_id)
            self.beam_doses.append(beam_dose * plan.n_fractions)
            self.dose += beam_dose

        self.dose *= plan.n_fractions

    def computeIMPTWET(self, plan, gpu_id=0):
        """WET volume of every beam, as VolumeObjects (plan_impt.py:208-236)."""
        self.RLSP = self.RLSPFromHU(plan.machine_name)
        rlsp_object = self._volume(self.RLSP)
        beam_wets = []
        for beam in plan.beam_list:
            beam_model = _model_for_beam(plan, beam)
            beam_wet = dose_kernels.proton_raytrace_cuda(beam_model, rlsp_object, beam, gpu_id)
            beam_wets.append(self._volume(beam_wet))
        return beam_wets

    def computeIMPTPlanFromWET(self, plan, wet_object, beam, gpu_id=0):
        """Spot dose of one beam from a cached WET volume; uses ``self.RLSP`` from
        an earlier call (plan_impt.py:238-257)."""
        self._check_isotropic()
        rlsp_object = self._volume(self.RLSP)
        beam_model = _model_for_beam(plan, beam)
        return dose_kernels.proton_spot_cuda(beam_model, rlsp_object, wet_object, beam, gpu_id)

    def writeWETNRRD(self, plan, wet_path, gpu_id=0):
        """Write each beam's WET volume (cm) to ``<wet_path>_beamNN.nrrd``
        (plan_impt.py:259-287)."""
        if not wet_path.endswith(".nrrd"):
            raise Exception("WET path must have .nrrd extension")
        wets = self.computeIMPTWET(plan, gpu_id)
        for i, wet in enumerate(wets):
            imageio.write_nrrd(wet_path.replace(".nrrd", "_beam%02i.nrrd" % (i + 1)),
                               np.asarray(wet.voxel_data, dtype=np.single), self.origin, self.spacing)


class IMPTBeam(Beam):
    """A scanned proton beam: ``spot_list`` is float32 [S, 4] rows of
    (x mm, y mm, MU, energy id) at the isocentre plane (plan_impt.py:290-327)."""

    def __init__(self, beam=None):
        super().__init__()
        self.spot_list = []
        self.n_spots = 0
        self.dicom_rangeshifter_label = None
        if beam is not None:
            for name in ("gantry_angle", "couch_angle", "iso", "spot_list", "n_spots",
                         "dicom_rangeshifter_label"):
                setattr(self, name, getattr(beam, name))

    def _append(self, rows):
        self.spot_list = np.vstack((self.spot_list, rows)) if self.n_spots > 0 else rows

    def addSpotData(self, cp, energy_id):
        """Append the spots of an ion control point (ScanSpotPositionMap,
        ScanSpotMetersetWeights) at one energy (plan_impt.py:307-315)."""
        xy = np.reshape(np.array(cp.ScanSpotPositionMap), (-1, 2))
        mu = np.array(cp.ScanSpotMetersetWeights)
        eid = np.full(mu.size, energy_id)
        self._append(np.array(np.column_stack((xy, mu, eid)), dtype=np.single))
        self.n_spots += mu.size

    def changeSpotEnergy(self, energy_id):
        self.spot_list[:, 3] = energy_id

    def addSingleSpot(self, x, y, mu, energy_id):
        """Append one spot.  As in the reference (plan_impt.py:320-323) the first
        spot is stored as a 1-D row, which the dose call rejects until a second
        spot makes ``spot_list`` 2-D."""
        self._append(np.array([x, y, mu, energy_id], dtype=np.single))
        self.n_spots += 1

    def resetSpots(self):
        self.spot_list = np.empty((0, 4), dtype=np.single)
        self.n_spots = 0


class IMPTPlan(Plan):
    """A proton plan on one machine: one IMPTBeamModel per range shifter listed
    in rangeshifter_labels.csv (plan_impt.py:330-342)."""

    def __init__(self, machine_name="HitachiProbeatJHU"):
        super().__init__()
        self.machine_name = machine_name
        model_dir = os.path.join("lookuptables", "protons", machine_name)
        rs = pd.read_csv(_data_path(model_dir, "rangeshifter_labels.csv"))
        self.dicom_rangeshifter_label = rs["dicom_rangeshifter_label"]
        self.folder_rangeshifter_label = rs["folder_rangeshifter_label"]
        self.beam_models = [IMPTBeamModel(d, model_dir, f)
                            for d, f in zip(self.dicom_rangeshifter_label, self.folder_rangeshifter_label)]
