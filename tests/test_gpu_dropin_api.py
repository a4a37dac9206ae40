"""The reference's other call sites of the dose path (GPU): the WET cache
(computeIMPTWET, computeIMPTPlanFromWET; plan_impt.py:208-257) and
writeWETNRRD (plan_impt.py:259-287) give exactly what computeIMPTPlan
computes beam by beam."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_wet_cache_and_wet_nrrd(tmp_path):
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from DoseCUDA import IMPTPlan
    from DoseCUDA.imageio import read_nrrd
    from DoseCUDA.workloads import _beam, _grid, grid_spots
    print("[gpu-dropin] WET cache, plan from WET, WET NRRD", flush=True)
    g = _grid((30, 36, 32), 2.5)
    rng = np.random.default_rng(4)
    g.HU = rng.uniform(-200.0, 400.0, size=(30, 36, 32)).astype(np.float32)
    plan = IMPTPlan()
    plan.n_fractions = 2
    c = np.arange(-9.0, 9.0 + 1e-6, 3.0)
    for ga, ca in ((0.0, 0.0), (110.0, 20.0)):
        sl = grid_spots([18, 26], c, c, disk_radius=9.0)
        sl[:, 2] = rng.uniform(0.2, 1.0, size=sl.shape[0]).astype(np.float32)
        plan.addBeam(_beam(ga, sl, couch=ca))

    g.computeIMPTPlan(plan)
    dose, beam_doses = g.dose.copy(), [d.copy() for d in g.beam_doses]
    assert float(dose.max()) > 0.0
    wets = g.computeIMPTWET(plan)
    assert len(wets) == 2
    total = np.zeros(g.size, dtype=np.single)
    for k, (w, beam) in enumerate(zip(wets, plan.beam_list)):
        d = g.computeIMPTPlanFromWET(plan, w, beam)
        assert np.array_equal(d * plan.n_fractions, beam_doses[k]), "beam %d" % k
        total += d
    total *= plan.n_fractions
    assert np.array_equal(total, dose)

    g.writeWETNRRD(plan, str(tmp_path / "wet.nrrd"))
    for k, w in enumerate(wets):
        vol, origin, spacing = read_nrrd(str(tmp_path / ("wet_beam%02d.nrrd" % (k + 1))))
        assert np.array_equal(vol, w.voxel_data)
        assert np.allclose(origin, g.origin) and np.allclose(spacing, g.spacing)


def test_reference_phantom_script(tmp_path, machine_dir):
    """The reference's own IMPT example, tests/test_phantom_impt.py: the
    default cube phantom (138^3 @ 3 mm), one beam with a ring of 98 single
    spots of increasing energy id, computeIMPTPlan, then the PHYSICAL dose and
    the CT written as NRRD.  Dose within 1e-3 of D_max of the oracle (gamma
    2%/2 mm 1.0), files equal to the arrays."""
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from DoseCUDA import IMPTBeam, IMPTDoseGrid, IMPTPlan
    from DoseCUDA.imageio import read_nrrd
    from oracle import impt_oracle as O
    print("[gpu-dropin] reference phantom script", flush=True)
    dose = IMPTDoseGrid()
    plan = IMPTPlan()
    beam = IMPTBeam()
    dose.createCubePhantom()
    beam.dicom_rangeshifter_label = '0'
    n_spots = 98
    for energy_id in range(n_spots):
        theta = 2.0 * 3.14159 * energy_id / n_spots
        beam.addSingleSpot(100.0 * np.cos(theta), 100.0 * np.sin(theta), 0.2, energy_id)
    plan.addBeam(beam)
    dose.computeIMPTPlan(plan)
    dose.writeDoseNRRD(str(tmp_path / "cube_impt_dose.nrrd"), individual_beams=False, dose_type="PHYSICAL")
    dose.writeCTNRRD(str(tmp_path / "cube_phantom_ct.nrrd"))

    d, origin, spacing = read_nrrd(str(tmp_path / "cube_impt_dose.nrrd"))
    assert np.array_equal(d, dose.dose) and np.allclose(origin, dose.origin) and np.allclose(spacing, 3.0)
    ct, _, _ = read_nrrd(str(tmp_path / "cube_phantom_ct.nrrd"))
    assert np.array_equal(ct, dose.HU)

    ref, _ = O.compute_plan(dose, plan, machine_dir)
    m = O.compare(dose.dose, ref)
    gamma = O.gamma_pass_rate(dose.dose, ref, 3.0)
    print("[gpu-dropin]   max|dD|/Dmax %.2e, gamma %.4f, D_max %.4g" % (m["max_rel_to_dmax"], gamma, float(ref.max())),
          flush=True)
    assert float(ref.max()) > 0.0 and m["max_rel_to_dmax"] <= 1e-3 and gamma == 1.0


def test_plan_and_ct_through_dicom(tmp_path):
    """A CT series and an RT Ion Plan written as DICOM and read back
    (loadCTDCM, writePlanDicom, readPlanDicom; plan.py:43-57,
    plan_impt.py:353-674) give the dose of the in-memory plan exactly: the
    files hold the same HU, spots, angles and isocentres, and the dose path
    sorts spots by energy id and position, so the layer order of the file does
    not matter."""
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from DoseCUDA import IMPTDoseGrid, IMPTPlan, dicomio
    from DoseCUDA.workloads import _beam, _grid, grid_spots
    print("[gpu-dropin] CT series and plan through DICOM", flush=True)
    g = _grid((24, 30, 28), 2.5)
    rng = np.random.default_rng(6)
    g.HU = np.round(rng.uniform(-300.0, 600.0, size=(24, 30, 28))).astype(np.float32)
    ct_dir = tmp_path / "ct"
    ct_dir.mkdir()
    for k in range(g.HU.shape[0]):
        ds = dicomio.Dataset()
        ds.SOPClassUID = dicomio.CT_IMAGE_STORAGE
        ds.SOPInstanceUID = dicomio.generate_uid()
        ds.Modality = "CT"
        ds.PatientName, ds.PatientID = "Phantom^Head", "PH1"
        ds.StudyInstanceUID, ds.SeriesInstanceUID, ds.FrameOfReferenceUID = "1.2.3", "1.2.3.1", "1.2.3.9"
        ds.InstanceNumber = g.HU.shape[0] - k
        ds.ImagePositionPatient = [float(g.origin[0]), float(g.origin[1]), float(g.origin[2] + k * g.spacing[2])]
        ds.PixelSpacing = [float(g.spacing[1]), float(g.spacing[0])]
        ds.Rows, ds.Columns = g.HU.shape[1], g.HU.shape[2]
        ds.SamplesPerPixel, ds.BitsAllocated, ds.BitsStored, ds.HighBit, ds.PixelRepresentation = 1, 16, 16, 15, 1
        ds.RescaleSlope, ds.RescaleIntercept = 1.0, -1024.0
        ds.PixelData = (g.HU[k].astype(np.int32) + 1024).astype("<i2").tobytes()
        meta = dicomio.Dataset()
        meta.MediaStorageSOPClassUID = dicomio.CT_IMAGE_STORAGE
        meta.MediaStorageSOPInstanceUID = ds.SOPInstanceUID
        dicomio.dcmwrite(str(ct_dir / ("ct%03d.dcm" % k)), ds, meta)

    plan = IMPTPlan()
    plan.n_fractions = 2
    c = np.arange(-9.0, 9.0 + 1e-6, 3.0)
    for ga, ca, rs in ((0.0, 0.0, "0"), (250.0, 15.0, "4")):
        sl = grid_spots([20, 31, 26], c, c, disk_radius=9.0)
        sl[:, 2] = rng.uniform(0.2, 1.0, size=sl.shape[0]).astype(np.float32)
        b = _beam(ga, sl, couch=ca)
        b.dicom_rangeshifter_label = rs
        b.iso = np.array([1.25, -2.5, 0.75], dtype=np.single)
        plan.addBeam(b)
    plan.writePlanDicom(str(ct_dir), str(tmp_path / "plan.dcm"))
    back = IMPTPlan()
    back.readPlanDicom(str(tmp_path / "plan.dcm"))

    h = IMPTDoseGrid()
    h.loadCTDCM(str(ct_dir))
    assert np.array_equal(h.HU, g.HU) and np.allclose(h.origin, g.origin) and np.allclose(h.spacing, g.spacing)
    g.computeIMPTPlan(plan)
    h.computeIMPTPlan(back)
    assert float(g.dose.max()) > 0.0
    assert np.array_equal(h.dose, g.dose)
    for a, b in zip(g.beam_doses, h.beam_doses):
        assert np.array_equal(a, b)
