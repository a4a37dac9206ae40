"""IMPT (proton pencil-beam scanning) plan API.

Re-creates the proton part of DoseCUDA/plan_impt.py of the reference with the
same class, method and attribute names:

* ``IMPTBeamModel``  (plan_impt.py:63-149)  beam-model lookup tables
* ``IMPTDoseGrid``   (plan_impt.py:152-287) HU->RLSP and the per-beam dose loop
* ``IMPTBeam``       (plan_impt.py:290-327) spot lists
* ``IMPTPlan``       (plan_impt.py:330-351) range-shifter models of a machine

The dose calls go through :mod:`DoseCUDA.dose_kernels`, whose functions keep
the reference extension's names and signatures but run hand-written gfx950
HIP kernels.  The RT Ion Plan reader / writer (``readPlanDicom``,
``writePlanDicom``, plan_impt.py:17-60, 344-674) go through
:mod:`DoseCUDA.dicomio` instead of pydicom.
"""

import datetime
import os
import sys

import numpy as np
import pandas as pd

from .plan import Plan, Beam, DoseGrid, VolumeObject, Prescription  # noqa: F401  (re-exported names)
from . import dicomio
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

    # -- RT Ion Plan DICOM (plan_impt.py:353-674), through DoseCUDA.dicomio --

    def load_ct_series(self, ct_dir):
        """Every ``.dcm`` file of ct_dir, sorted by InstanceNumber (plan_impt.py:344-351)."""
        files = [os.path.join(ct_dir, f) for f in os.listdir(ct_dir) if f.lower().endswith(".dcm")]
        series = [dicomio.dcmread(f, force=True) for f in files]
        series.sort(key=lambda ds: ds.InstanceNumber)
        return series

    def _beam_model_for_label(self, label):
        labels = list(np.asarray(self.dicom_rangeshifter_label).astype(str))
        if label not in labels:
            raise ValueError(f"Beam model not found for rangeshifter ID {label}")
        return self.beam_models[labels.index(label)]

    def writePlanDicom(self, ct_dir, output_file, rt_struct_file=None):
        """Write the plan as an RT Ion Plan (plan_impt.py:353-633).

        Patient and study identifiers come from the first CT slice of ct_dir.
        Each energy layer of a beam (descending energy id) becomes a pair of
        ion control points: the first carries the spot weights, the second
        the same positions with zero weights; cumulative meterset weights run
        over the layers.  Dose references come from the prescriptions (one
        default 2 Gy x n_fractions PTV prescription when there are none),
        with ROI numbers from rt_struct_file when it is given.
        """
        meta = dicomio.Dataset()
        meta.MediaStorageSOPClassUID = dicomio.RT_ION_PLAN_STORAGE
        meta.MediaStorageSOPInstanceUID = dicomio.generate_uid()
        meta.TransferSyntaxUID = dicomio.EXPLICIT_VR_LE

        ds = dicomio.Dataset()
        first_ct = self.load_ct_series(ct_dir)[0]
        for kw in ("PatientName", "PatientID", "StudyInstanceUID", "FrameOfReferenceUID"):
            setattr(ds, kw, getattr(first_ct, kw))
        now = datetime.datetime.now()
        day, clock = now.strftime("%Y%m%d"), now.strftime("%H%M%S.%f")
        _fill(ds, [("SeriesInstanceUID", dicomio.generate_uid()),
                   ("SOPClassUID", dicomio.RT_ION_PLAN_STORAGE),
                   ("SOPInstanceUID", meta.MediaStorageSOPInstanceUID),
                   ("Modality", "RTPLAN"),
                   ("RTPlanLabel", self.RTPlanLabel if self.RTPlanLabel else "IMPT Plan"),
                   ("RTPlanDate", day), ("RTPlanTime", clock),
                   ("InstanceCreationDate", day), ("InstanceCreationTime", clock),
                   ("NumberOfBeams", len(self.beam_list))] + _PLAN_CONSTANTS)
        rss = dicomio.Dataset()
        rss.ReferencedSOPClassUID = dicomio.RT_STRUCTURE_SET_STORAGE
        rss.ReferencedSOPInstanceUID = dicomio.generate_uid()
        ds.ReferencedStructureSetSequence = [rss]

        prescriptions = list(self.Prescriptions)
        if not prescriptions:
            rx = Prescription()
            rx.TargetPrescriptionDose = self.n_fractions * 2.0
            rx.ROIName = "PTV"
            rx.TargetUnderdoseVolumeFraction = 2.0
            prescriptions = [rx]
        ds = create_dose_reference(ds, prescriptions, rt_struct_file)

        fg = dicomio.Dataset()
        _fill(fg, [("FractionGroupNumber", 1), ("NumberOfFractionsPlanned", self.n_fractions),
                   ("NumberOfBeams", len(self.beam_list)), ("NumberOfBrachyApplicationSetups", 0),
                   ("ReferencedBeamSequence", [])])
        ds.FractionGroupSequence = [fg]
        setup = dicomio.Dataset()
        setup.PatientPosition = "HFS"
        setup.PatientSetupNumber = 1
        ds.PatientSetupSequence = [setup]

        ds.IonBeamSequence = []
        for number, beam in enumerate(self.beam_list, start=1):
            model = self._beam_model_for_label(beam.dicom_rangeshifter_label)
            ds.IonBeamSequence.append(self._ion_beam(number, beam, model, fg))
        dicomio.dcmwrite(output_file, ds, meta)

    def _ion_beam(self, number, beam, model, fg):
        """One Ion Beam Sequence item; appends the beam's meterset to fg."""
        spots = np.asarray(beam.spot_list)
        rs_label = beam.dicom_rangeshifter_label
        item = dicomio.Dataset()
        _fill(item, [("BeamNumber", number), ("BeamName", beam.BeamName),
                     ("BeamDescription", beam.BeamDescription)] + _BEAM_CONSTANTS +
              [("VirtualSourceAxisDistances", [float(model.VSADX), float(model.VSADY)]),
               ("SnoutSequence", [_dataset([("SnoutID", "0")])]),
               ("NumberOfRangeShifters", 1 if rs_label else 0)])
        if rs_label:
            item.RangeShifterSequence = [_dataset([("RangeShifterNumber", 1), ("RangeShifterID", rs_label),
                                                   ("RangeShifterType", "BINARY"),
                                                   ("AccessoryCode", rs_label)])]
        item.NumberOfLateralSpreadingDevices = 1
        item.LateralSpreadingDeviceSequence = [_dataset([("LateralSpreadingDeviceNumber", 1),
                                                         ("LateralSpreadingDeviceID", "0"),
                                                         ("LateralSpreadingDeviceType", "SCATTERER")])]
        item.NumberOfRangeModulators = 1
        item.RangeModulatorSequence = [_dataset([("RangeModulatorNumber", 0), ("RangeModulatorID", "0"),
                                                 ("RangeModulatorType", "FIXED"),
                                                 ("RangeModulatorDescription", "")])]
        total_mu = float(np.sum(spots[:, 2]))
        fg.ReferencedBeamSequence.append(_dataset([("ReferencedBeamNumber", number), ("BeamMeterset", total_mu)]))
        item.FinalCumulativeMetersetWeight = total_mu
        for tag, vr, value in _BEAM_PRIVATE:
            item.add_new(tag, vr, value)

        cps = []
        cum = 0.0
        fwhm_per_sigma_cm = 2.0 * np.sqrt(2.0 * np.log(2.0)) * 10.0
        for eid in np.unique(spots[:, 3]).astype(int)[::-1]:
            layer = spots[spots[:, 3] == eid]
            layer_mu = np.sum(layer[:, 2])
            width = model.reference_sigmas[eid] * fwhm_per_sigma_cm
            xy = layer[:, 0:2].flatten().tolist()
            lsd = [_dataset([("LateralSpreadingDeviceSetting", "IN"),
                             ("IsocenterToLateralSpreadingDeviceDistance", 2275.0),
                             ("LateralSpreadingDeviceWaterEquivalentThickness", 0.0),
                             ("ReferencedLateralSpreadingDeviceNumber", 1)])]
            rms = [_dataset([("ReferencedRangeModulatorNumber", 0)])]
            rdr = [_dataset([("ReferencedDoseReferenceNumber", k),
                             ("CumulativeDoseReferenceCoefficient", None)]) for k in (1, 2)]
            # the layer's two control points: weights, then zero weights
            for cum_w, weights in ((cum, layer[:, 2].tolist()), (cum + layer_mu, [0.0] * len(layer))):
                cp = dicomio.Dataset()
                _fill(cp, [("ControlPointIndex", len(cps)), ("NominalBeamEnergyUnit", "MEV"),
                           ("NominalBeamEnergy", float(model.energy_labels[eid])),
                           ("GantryAngle", float(beam.gantry_angle)),
                           ("PatientSupportAngle", float(beam.couch_angle)),
                           ("TableTopEccentricAngle", float(beam.couch_angle)),
                           ("IsocenterPosition", [float(c) for c in beam.iso]),
                           ("CumulativeMetersetWeight", float(cum_w)),
                           ("NumberOfScanSpotPositions", len(layer)),
                           ("ScanSpotPositionMap", xy), ("ScanSpotMetersetWeights", weights),
                           ("ScanningSpotSize", [width, width])] + _CP_CONSTANTS)
                if rs_label:
                    cp.RangeShifterSetting = "ON"
                cp.LateralSpreadingDeviceSettingsSequence = lsd
                cp.RangeModulatorSettingsSequence = rms
                cp.ReferencedDoseReferenceSequence = rdr
                cps.append(cp)
            cum += layer_mu
        item.IonControlPointSequence = cps
        item.NumberOfControlPoints = len(cps)
        return item

    def readPlanDicom(self, plan_path):
        """Replace the beams with those of an RT Ion Plan (plan_impt.py:635-674).

        Per beam: range-shifter label from the first Range Shifter Sequence
        item ('0' without one), gantry / couch angle and isocentre from the
        first control point, and the spots of every even-numbered control
        point at the energy id of its nominal energy in the first beam model.
        """
        ds = dicomio.dcmread(plan_path, force=True)
        self.n_fractions = float(ds.FractionGroupSequence[0].NumberOfFractionsPlanned)
        self.beam_list = []
        self.n_beams = 0
        for ibs in ds.IonBeamSequence:
            beam = IMPTBeam()
            if ibs.NumberOfRangeShifters and ibs.RangeShifterSequence[0].RangeShifterID is not None:
                beam.dicom_rangeshifter_label = ibs.RangeShifterSequence[0].RangeShifterID
            else:
                beam.dicom_rangeshifter_label = "0"
            cp0 = ibs.IonControlPointSequence[0]
            beam.gantry_angle = float(cp0.GantryAngle)
            beam.couch_angle = float(cp0.PatientSupportAngle)
            beam.iso = np.array(cp0.IsocenterPosition, dtype=np.single)
            if "BeamName" in ibs:
                beam.BeamName = ibs.BeamName
            if "BeamDescription" in ibs:
                beam.BeamDescription = ibs.BeamDescription
            for cp in ibs.IonControlPointSequence[::2]:
                beam.addSpotData(cp, self.beam_models[0].energyIDFromLabel(float(cp.NominalBeamEnergy)))
            self.addBeam(beam)


# Fixed attributes of the RT Ion Plan the writer emits (plan_impt.py:386-395,
# 459-475, 536-555).
_PLAN_CONSTANTS = [("RTPlanGeometry", "PATIENT"), ("ApprovalStatus", "UNAPPROVED"),
                   ("Manufacturer", "RaySearch Laboratories"), ("PlanIntent", "CURATIVE"),
                   ("TreatmentProtocols", ["PencilBeamScanning"]), ("PatientIdentityRemoved", "YES"),
                   ("DeidentificationMethod", "RayStation Custom Export"),
                   ("NumberOfBrachyApplicationSetups", 0)]
_BEAM_CONSTANTS = [("BeamType", "STATIC"), ("RadiationType", "PROTON"), ("ScanMode", "MODULATED_SPEC"),
                   ("ModulatedScanModeType", "STATIONARY"), ("TreatmentMachineName", "ProBeat_noMRF_"),
                   ("Manufacturer", "Generic"), ("PrimaryDosimeterUnit", "MU"),
                   ("TreatmentDeliveryType", "TREATMENT"), ("NumberOfWedges", 0), ("NumberOfCompensators", 0),
                   ("NumberOfBoli", 0), ("NumberOfBlocks", 0), ("PatientSupportType", "TABLE"),
                   ("PatientSupportID", "TABLE")]
_BEAM_PRIVATE = [((0x4001, 0x0010), "LO", "RAYSEARCHLABS 2.0"),
                 ((0x4001, 0x1002), "UN", b"Constant 1.1"),
                 ((0x4001, 0x1012), "UN", b"ProBeat_noMRF_"),
                 ((0x4001, 0x1033), "UN", b"RBE_NOT_INCLUDED"),
                 ((0x4001, 0x1003), "UN", b"20181030034922.000000 ")]
_CP_CONSTANTS = [("GantryRotationDirection", "NONE"), ("BeamLimitingDeviceAngle", 0.0),
                 ("BeamLimitingDeviceRotationDirection", "NONE"), ("PatientSupportRotationDirection", "NONE"),
                 ("TableTopEccentricRotationDirection", "NONE"), ("TableTopPitchAngle", 0.0),
                 ("TableTopPitchRotationDirection", "NONE"), ("TableTopRollAngle", 0.0),
                 ("TableTopRollRotationDirection", "NONE"), ("GantryPitchAngle", 0.0),
                 ("GantryPitchRotationDirection", "NONE"), ("TableTopVerticalPosition", ""),
                 ("TableTopLongitudinalPosition", ""), ("TableTopLateralPosition", ""),
                 ("SnoutPosition", 391.0), ("MetersetRate", 480.0), ("ScanSpotTuneID", "Standard"),
                 ("NumberOfPaintings", 1)]


def _fill(ds, pairs):
    for kw, value in pairs:
        setattr(ds, kw, value)
    return ds


def _dataset(pairs):
    return _fill(dicomio.Dataset(), pairs)


def get_roi_num(rt_ds, ROIName):
    """ROINumber of the structure named ROIName in an RT Structure Set, or
    None (plan_impt.py:17-21)."""
    for roi in rt_ds.StructureSetROISequence:
        if roi.ROIName == ROIName:
            return roi.ROINumber
    return None


def create_dose_reference(ds, prescriptions, rt_struct_file=None):
    """Append a target Dose Reference item per prescription (plan_impt.py:24-60).

    ROI numbers are the prescriptions' positions (1, 2, ...) or, with an RT
    Structure Set file, the numbers of the structures named by the
    prescriptions (ValueError for a missing name), whose SOP instance UID then
    becomes the plan's referenced structure set.
    """
    if "DoseReferenceSequence" not in ds:
        ds.DoseReferenceSequence = []
    rt_ds = dicomio.dcmread(rt_struct_file, force=True) if rt_struct_file else None
    for pos, rx in enumerate(prescriptions, start=1):
        roi = pos
        if rt_ds is not None:
            roi = get_roi_num(rt_ds, rx.ROIName)
            if roi is None:
                raise ValueError(f"No structure labelled '{rx.ROIName}' found in RT structure set.")
        dr = _dataset([("DoseReferenceNumber", len(ds.DoseReferenceSequence) + 1), ("ReferencedROINumber", roi),
                       ("DoseReferenceStructureType", "VOLUME"),
                       ("DoseReferenceDescription", f"{rx.ROIName} Prescription"),
                       ("DoseReferenceType", "TARGET"), ("TargetPrescriptionDose", rx.TargetPrescriptionDose),
                       ("TargetUnderdoseVolumeFraction", rx.TargetUnderdoseVolumeFraction)])
        dr.add_new((0x4001, 0x0010), "LO", "RAYSEARCHLABS 2.0")
        dr.add_new((0x4001, 0x1011), "UN", str(rx.TargetPrescriptionDose).encode())
        ds.DoseReferenceSequence.append(dr)
    if rt_ds is not None:
        ds.ReferencedStructureSetSequence[0].ReferencedSOPInstanceUID = rt_ds.SOPInstanceUID
    return ds
