"""IMRT (photon) plan API.

Re-creates DoseCUDA/plan_imrt.py of the reference with the same class, method
and attribute names:

* ``IMRTPhotonEnergy``  (plan_imrt.py:12-92)   beam model of one photon energy
                        and its equivalent-square output factor
* ``IMRTControlPoint``  (plan_imrt.py:95-105)  one MLC segment
* ``IMRTDoseGrid``      (plan_imrt.py:108-158) HU->density and the control-point
                        dose loop
* ``IMRTBeam``          (plan_imrt.py:161-171) control points of a beam
* ``IMRTPlan``          (plan_imrt.py:174-382) machine models, RT Plan reader,
                        square fields

The dose call goes through :func:`DoseCUDA.dose_kernels.photon_dose_cuda`,
which keeps the reference extension's name and attribute protocol and runs
hand-written gfx950 HIP kernels.  ``readPlanDicom`` goes through
:mod:`DoseCUDA.dicomio` instead of pydicom.
"""

import os
import sys

import numpy as np
import pandas as pd

from .plan import Plan, Beam, DoseGrid, VolumeObject
from . import dicomio
from . import dose_kernels

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))


def _photon_dir(machine_name, *parts):
    return os.path.join(_PKG_DIR, "lookuptables", "photons", machine_name, *parts)


class IMRTPhotonEnergy():
    """Beam model of one photon energy (plan_imrt.py:12-43).

    Attributes read by the dose kernel (dosemodule.cu:377-432): the floats
    ``mu_calibration``, ``primary_source_distance``, ``scatter_source_distance``,
    ``primary_source_size``, ``scatter_source_size``, ``mlc_distance``,
    ``scatter_source_weight``, ``electron_attenuation``,
    ``electron_source_weight``, ``electron_fitted_dmax``, ``jaw_transmission``,
    ``mlc_transmission``; the flags ``has_xjaws``, ``has_yjaws``; the float32
    1-D arrays ``profile_radius``, ``profile_intensities``, ``profile_softening``,
    ``spectrum_attenuation_coefficients``, ``spectrum_primary_weights``,
    ``spectrum_scatter_weights``; and the float32 [6, 6] ``kernel`` whose rows
    are the six collapsed cones (theta deg, Am, am, Bm, bm, ray length cm).
    """

    def __init__(self, dicom_energy_label):
        self.dicom_energy_label = dicom_energy_label
        self.output_factor_equivalent_squares = None
        self.output_factor_values = None
        self.mu_calibration = None
        self.primary_source_distance = None
        self.scatter_source_distance = None
        self.mlc_distance = None
        self.scatter_source_weight = None
        self.electron_attenuation = None
        self.primary_source_size = None
        self.scatter_source_size = None
        self.profile_radius = None
        self.profile_intensities = None
        self.profile_softening = None
        self.spectrum_attenuation_coefficients = None
        self.spectrum_primary_weights = None
        self.spectrum_scatter_weights = None
        self.electron_source_weight = None
        self.has_xjaws = None
        self.has_yjaws = None
        self.electron_fitted_dmax = None
        self.jaw_transmission = None
        self.mlc_transmission = None
        self.mlc_index = None
        self.mlc_widths = None
        self.mlc_offsets = None
        self.n_mlc_pairs = None
        self.kernel = None

    _REQUIRED = ('output_factor_equivalent_squares', 'output_factor_values', 'mu_calibration',
                 'primary_source_distance', 'scatter_source_distance', 'mlc_distance', 'scatter_source_weight',
                 'electron_attenuation', 'primary_source_size', 'scatter_source_size', 'profile_radius',
                 'profile_intensities', 'profile_softening', 'spectrum_attenuation_coefficients',
                 'spectrum_primary_weights', 'spectrum_scatter_weights', 'electron_source_weight', 'has_xjaws',
                 'has_yjaws', 'electron_fitted_dmax', 'jaw_transmission', 'mlc_transmission')

    def validate_parameters(self):
        """Raise if a table entry the dose needs is missing (plan_imrt.py:45-59)."""
        for name in self._REQUIRED:
            if getattr(self, name) is None:
                raise Exception(f"{name} not set in beam model")

    def outputFactor(self, cp):
        """Output factor of a control point from its MLC opening
        (plan_imrt.py:61-92): equivalent square 4 A / P of the open leaf
        pairs (area of every pair; perimeter from the widest opening and the
        y extent of pairs open by more than 3 mm), interpolated in the
        machine's output-factor table."""
        min_y, max_y = 10000.0, -10000.0
        max_x_diff, area = 0.0, 0.0
        for x1, x2, y_offset, y_width in np.asarray(cp.mlc).reshape(-1, 4):
            opening = x2 - x1
            area += opening * y_width
            if opening > 3.0:
                max_x_diff = max(max_x_diff, opening)
                min_y = min(min_y, y_offset - y_width / 2.0)
                max_y = max(max_y, y_offset + y_width / 2.0)
        perimeter = 2 * (max_y - min_y) + 2 * max_x_diff
        equivalent_square = 4 * area / perimeter
        return np.interp(equivalent_square, self.output_factor_equivalent_squares, self.output_factor_values)


class IMRTControlPoint:
    """One MLC segment (plan_imrt.py:95-105): isocentre (mm), meterset ``mu``,
    ``mlc`` float32 [n_pairs, 4] rows (x1, x2, y_offset, y_width) in mm at the
    isocentre plane, gantry / collimator / table angles (deg), jaw positions."""

    def __init__(self, iso, mu, mlc, ga, ca, ta, xjaws, yjaws):
        self.iso = iso
        self.mu = mu
        self.mlc = mlc
        self.ga = ga
        self.ca = ca
        self.ta = ta
        self.xjaws = xjaws
        self.yjaws = yjaws


class IMRTDoseGrid(DoseGrid):
    """Dose grid with the photon dose entry point (plan_imrt.py:108-158)."""

    def __init__(self):
        super().__init__()
        self.Density = []

    def DensityFromHU(self, machine_name):
        """Mass density of every voxel: np.interp over the machine's
        HU_Density.csv knots, cast to float32 (plan_imrt.py:114-124)."""
        df = pd.read_csv(_photon_dir(machine_name, "HU_Density.csv"))
        return np.array(np.interp(self.HU, df["HU"].to_numpy(), df["Density"].to_numpy()), dtype=np.single)

    def computeIMRTPlan(self, plan, gpu_id=0):
        """Dose of every beam (sum over its control points of the control
        point's dose times its output factor) and their sum times the number
        of fractions (plan_imrt.py:126-158)."""
        self.beam_doses = []
        self.dose = np.zeros(self.size, dtype=np.single)
        self.Density = self.DensityFromHU(plan.machine_name)

        if self.spacing[0] != self.spacing[1] or self.spacing[0] != self.spacing[2]:
            raise Exception("Spacing must be isotropic for IMPT dose calculation - consider resampling CT")

        density_object = VolumeObject()
        density_object.voxel_data = np.array(self.Density, dtype=np.single)
        density_object.origin = np.array(self.origin, dtype=np.single)
        density_object.spacing = np.array(self.spacing, dtype=np.single)

        labels = list(np.asarray(plan.dicom_energy_label).astype(str))
        for beam in plan.beam_list:
            beam_dose = np.zeros(self.Density.shape, dtype=np.single)
            try:
                model_index = labels.index(beam.dicom_energy_label)
            except ValueError:
                print("Beam model not found for beam energy %s" % beam.dicom_energy_label)
                sys.exit(1)
            beam_model = plan.beam_models[model_index]
            for cp in beam.cp_list:
                output_factor = beam_model.outputFactor(cp)
                cp_dose = dose_kernels.photon_dose_cuda(beam_model, density_object, cp, gpu_id)
                beam_dose += cp_dose * output_factor
            self.beam_doses.append(beam_dose)
            self.dose += beam_dose

        self.dose *= plan.n_fractions


class IMRTBeam(Beam):
    """A photon beam: its control points and DICOM energy label (plan_imrt.py:161-171)."""

    def __init__(self):
        super().__init__()
        self.cp_list = []
        self.n_cps = 0
        self.dicom_energy_label = None

    def addControlPoint(self, cp):
        self.cp_list.append(cp)
        self.n_cps += 1


def _param_values(line):
    return line.split(',')[1:]


class IMRTPlan(Plan):
    """Photon plan of one machine (plan_imrt.py:174-382): a beam model per
    energy of ``lookuptables/photons/<machine_name>``."""

    def __init__(self, machine_name="VarianTrueBeamHF"):
        super().__init__()
        self.machine_name = machine_name
        energy_list = pd.read_csv(_photon_dir(machine_name, "energy_labels.csv"))
        self.dicom_energy_label = energy_list["dicom_energy_label"]
        self.folder_energy_label = energy_list["folder_energy_label"]
        self.beam_models = []
        for d, f in zip(self.dicom_energy_label, self.folder_energy_label):
            beam_model = IMRTPhotonEnergy(d)
            self._load_beam_model_parameters(beam_model, machine_name, f)
            self.beam_models.append(beam_model)

    def _load_beam_model_parameters(self, beam_model, machine_name, folder_energy_label):
        """MLC geometry, kernel, machine geometry and beam parameters of one
        energy (plan_imrt.py:192-219)."""
        mlc_geometry = pd.read_csv(_photon_dir(machine_name, "mlc_geometry.csv"))
        beam_model.mlc_index = mlc_geometry["mlc_pair_index"].to_numpy()
        beam_model.mlc_widths = mlc_geometry["width"].to_numpy()
        beam_model.mlc_offsets = mlc_geometry["center_offset"].to_numpy()
        beam_model.n_mlc_pairs = len(beam_model.mlc_index)

        kernel = pd.read_csv(_photon_dir(machine_name, folder_energy_label, "kernel.csv"))
        beam_model.kernel = np.array(kernel.to_numpy(), dtype=np.single)

        self._load_machine_geometry(beam_model, _photon_dir(machine_name, "machine_geometry.csv"))
        self._load_beam_parameters(beam_model, _photon_dir(machine_name, folder_energy_label, "beam_parameters.csv"))
        beam_model.validate_parameters()

    def _load_machine_geometry(self, beam_model, machine_geometry_path):
        """Source and MLC distances and jaw flags (plan_imrt.py:221-233).  As
        in the reference a flag is bool() of its text, so any non-empty value
        (also ' 0') reads as True."""
        with open(machine_geometry_path) as fh:
            for line in fh:
                for key in ('primary_source_distance', 'scatter_source_distance', 'mlc_distance'):
                    if line.startswith(key):
                        setattr(beam_model, key, float(line.split(',')[1]))
                for key in ('has_xjaws', 'has_yjaws'):
                    if line.startswith(key):
                        setattr(beam_model, key, bool(line.split(',')[1]))

    _ARRAY_PARAMS = ('output_factor_equivalent_squares', 'output_factor_values', 'profile_radius',
                     'profile_intensities', 'profile_softening', 'spectrum_attenuation_coefficients',
                     'spectrum_primary_weights', 'spectrum_scatter_weights')
    _FLOAT_PARAMS = {'mu_calibration': 'mu_calibration', 'scatter_source_weight': 'scatter_source_weight',
                     'electron_attenuation': 'electron_attenuation', 'primary_source_size': 'primary_source_size',
                     'scatter_source_size': 'scatter_source_size', 'electron_source_weight': 'electron_source_weight',
                     'electron_fitted_dmax': 'electron_fitted_dmax', 'jaw_transmission': 'jaw_transmission',
                     'mlc_transmission': 'mlc_transmission'}

    def _load_beam_parameters(self, beam_model, beam_parameter_path):
        """Per-energy parameters, one ``name, v1, v2, ...`` line each
        (plan_imrt.py:235-271); arrays as float32, scalars as float."""
        with open(beam_parameter_path) as fh:
            for line in fh:
                for key in self._ARRAY_PARAMS:
                    if line.startswith(key):
                        setattr(beam_model, key, np.array(_param_values(line), dtype=np.single))
                for key, attr in self._FLOAT_PARAMS.items():
                    if line.startswith(key):
                        setattr(beam_model, attr, float(line.split(',')[1]))

    def _mlc_array(self, mlc):
        """[2, n] leaf positions -> float32 [n, 4] rows (x1, x2, y_offset,
        y_width) with the machine's leaf offsets and widths
        (plan_imrt.py:311-313, 329-330, 373-374)."""
        bm = self.beam_models[0]
        full = np.array(np.vstack((mlc, bm.mlc_offsets.reshape(1, -1), bm.mlc_widths.reshape(1, -1))), dtype=np.single)
        return np.transpose(full)

    def readPlanDicom(self, plan_path):
        """Replace the beams with the TREATMENT beams of an RT Plan
        (plan_imrt.py:273-351).

        Per beam: energy label, isocentre and collimator / couch / gantry
        angles of control point 0; one IMRTControlPoint per later control
        point with meterset (cumulative weight difference) x the first
        referenced beam's meterset, and the leaf (or, without an MLC, jaw)
        aperture in force before it.  As in the reference the X and Y jaw
        positions are read from the second and the MLC leaves from the third
        item of the device position sequence, whatever its order.
        """
        ds = dicomio.dcmread(plan_path, force=True)
        total_mu = float(ds.FractionGroupSequence[0].ReferencedBeamSequence[0].BeamMeterset)
        self.n_fractions = float(ds.FractionGroupSequence[0].NumberOfFractionsPlanned)
        self.beam_list = []
        self.n_beams = 0
        n_pairs = self.beam_models[0].n_mlc_pairs
        offsets = self.beam_models[0].mlc_offsets

        def apertures(cp):
            xjaws = yjaws = mlc = None
            seq = cp.BeamLimitingDevicePositionSequence
            for ps in seq:
                if ps.RTBeamLimitingDeviceType == 'X':
                    xjaws = np.array(seq[1].LeafJawPositions, dtype=np.single)
                if ps.RTBeamLimitingDeviceType == 'Y':
                    yjaws = np.array(seq[1].LeafJawPositions, dtype=np.single)
                if ps.RTBeamLimitingDeviceType == 'MLCX':
                    mlc = np.array(seq[2].LeafJawPositions, dtype=np.single)
            return xjaws, yjaws, mlc

        for beam in ds.BeamSequence:
            if beam.TreatmentDeliveryType != "TREATMENT":
                continue
            imrt_beam = IMRTBeam()
            for cp in beam.ControlPointSequence:
                if int(cp.ControlPointIndex) == 0:
                    imrt_beam.dicom_energy_label = _label_text(cp.NominalBeamEnergy)
                    ip = cp.IsocenterPosition
                    imrt_beam.iso = np.array([ip[0], ip[1], ip[2]], dtype=np.single)
                    cumulative_mu = 0.0
                    ca = float(cp.BeamLimitingDeviceAngle)
                    ta = float(cp.PatientSupportAngle)
                    ga = float(cp.GantryAngle)
                    xjaws, yjaws, mlc = apertures(cp)
                    continue
                mu = float(cp.CumulativeMetersetWeight) - cumulative_mu
                if mlc is not None:
                    leaves = np.reshape(mlc, (2, n_pairs))
                else:
                    leaves = np.zeros((2, n_pairs), dtype=np.single)
                    if xjaws is not None:
                        leaves[0, :] = xjaws[0]
                        leaves[1, :] = xjaws[1]
                    if yjaws is not None:
                        outside = (offsets < yjaws[0]) | (offsets > yjaws[1])
                        leaves[:, outside] = 0.0
                imrt_beam.addControlPoint(IMRTControlPoint(imrt_beam.iso, mu * total_mu, self._mlc_array(leaves),
                                                           ga, ca, ta, xjaws, yjaws))
                if "GantryAngle" in cp:
                    ga = float(cp.GantryAngle)
                if "BeamLimitingDevicePositionSequence" in cp:
                    xjaws, yjaws, mlc = apertures(cp)
                cumulative_mu = float(cp.CumulativeMetersetWeight)
            self.addBeam(imrt_beam)

    def addSquareField(self, dicom_energy_label='6', dimx=10, dimy=10, mu=100, gantry_angle=0.0,
                       collimator_angle=0.0, table_angle=0.0):
        """Add a beam of one control point: a dimx x dimy cm field at the
        isocentre (0, 0, 0), leaf pairs whose centre lies outside +-dimy/2
        closed (plan_imrt.py:353-382)."""
        imrt_beam = IMRTBeam()
        imrt_beam.dicom_energy_label = dicom_energy_label
        iso = np.array([0.0, 0.0, 0.0], dtype=np.single)
        dimx *= 10.0
        dimy *= 10.0
        bm = self.beam_models[0]
        leaves = np.zeros((2, bm.n_mlc_pairs), dtype=np.single)
        leaves[0, :] = -dimx / 2.0
        leaves[1, :] = dimx / 2.0
        outside = (bm.mlc_offsets < -(dimy / 2.0)) | (bm.mlc_offsets > (dimy / 2.0))
        leaves[:, outside] = 0.0
        jawx = np.array([-dimx / 2.0, dimx / 2.0], dtype=np.single)
        jawy = np.array([-dimy / 2.0, dimy / 2.0], dtype=np.single)
        imrt_beam.addControlPoint(IMRTControlPoint(iso, mu, self._mlc_array(leaves), gantry_angle, collimator_angle,
                                                   table_angle, jawx, jawy))
        self.addBeam(imrt_beam)


def _label_text(value):
    """DICOM NominalBeamEnergy (DS) as an energy label.  pydicom's str() of a
    DS gives the file's characters ('6' for the usual 6 MV beam); the decoder
    here keeps the number, so an integral value reads as its integer text."""
    if isinstance(value, (bytes, bytearray)):
        value = value.decode(errors="replace")
    if isinstance(value, str):
        return value.strip()
    f = float(value)
    return str(int(f)) if f == int(f) else repr(f)
