"""Photon plan API on the CPU (no dose): square fields, the RT Plan reader and
the control-point bookkeeping of plan_imrt.py:161-382.  Expected values are
worked out here from the reader's rules (plan_imrt.py:273-351), including
its quirks: jaws are read from the second and leaves from the third item of
the device position sequence, the meterset of a control point is the
cumulative-weight step times the first referenced beam's meterset, and the
aperture in force before a control point is the one used."""

import numpy as np

from DoseCUDA import dicomio
from DoseCUDA.plan_imrt import IMRTPlan, IMRTBeam


def _item(**kw):
    it = dicomio.Dataset()
    for k, v in kw.items():
        setattr(it, k, v)
    return it


def test_square_field_layout():
    plan = IMRTPlan("VarianTrueBeamHF")
    plan.addSquareField('10', 6.0, 4.0, mu=75, gantry_angle=30.0, collimator_angle=15.0, table_angle=5.0)
    beam = plan.beam_list[0]
    assert isinstance(beam, IMRTBeam) and beam.dicom_energy_label == '10' and beam.n_cps == 1
    cp = beam.cp_list[0]
    bm = plan.beam_models[0]
    assert cp.mlc.shape == (bm.n_mlc_pairs, 4) and cp.mlc.dtype == np.float32
    open_ = (bm.mlc_offsets >= -20.0) & (bm.mlc_offsets <= 20.0)
    np.testing.assert_array_equal(cp.mlc[open_, 0], -30.0)
    np.testing.assert_array_equal(cp.mlc[open_, 1], 30.0)
    np.testing.assert_array_equal(cp.mlc[~open_, :2], 0.0)
    np.testing.assert_array_equal(cp.mlc[:, 2], bm.mlc_offsets)
    np.testing.assert_array_equal(cp.mlc[:, 3], bm.mlc_widths)
    assert (cp.mu, cp.ga, cp.ca, cp.ta) == (75, 30.0, 15.0, 5.0)
    np.testing.assert_array_equal(cp.xjaws, [-30, 30])
    np.testing.assert_array_equal(cp.yjaws, [-20, 20])
    np.testing.assert_array_equal(cp.iso, [0, 0, 0])


def _write_rt_plan(path, n_pairs):
    rng = np.random.default_rng(3)
    leaves0 = np.concatenate([rng.uniform(-40, -5, n_pairs), rng.uniform(5, 40, n_pairs)]).round(1)
    leaves1 = np.concatenate([rng.uniform(-30, -2, n_pairs), rng.uniform(2, 30, n_pairs)]).round(1)

    def bldps(items):
        return [_item(RTBeamLimitingDeviceType=t, LeafJawPositions=list(v)) for t, v in items]

    cps1 = [_item(ControlPointIndex=0, NominalBeamEnergy=6.0, IsocenterPosition=[1.5, -2.0, 3.0],
                  BeamLimitingDeviceAngle=10.0, PatientSupportAngle=0.0, GantryAngle=40.0,
                  CumulativeMetersetWeight=0.0,
                  BeamLimitingDevicePositionSequence=bldps([("ASYMX", [-50, 50]), ("ASYMY", [-60, 60]),
                                                            ("MLCX", leaves0)])),
            _item(ControlPointIndex=1, CumulativeMetersetWeight=0.4, GantryAngle=45.0,
                  BeamLimitingDevicePositionSequence=bldps([("ASYMX", [-50, 50]), ("ASYMY", [-60, 60]),
                                                            ("MLCX", leaves1)])),
            _item(ControlPointIndex=2, CumulativeMetersetWeight=1.0)]
    # jaws only, typed 'X' / 'Y': both read from the second item (the Y jaws)
    cps2 = [_item(ControlPointIndex=0, NominalBeamEnergy=10.0, IsocenterPosition=[0.0, 0.0, 0.0],
                  BeamLimitingDeviceAngle=0.0, PatientSupportAngle=15.0, GantryAngle=180.0,
                  CumulativeMetersetWeight=0.0,
                  BeamLimitingDevicePositionSequence=bldps([("X", [-70, 70]), ("Y", [-35, 45])])),
            _item(ControlPointIndex=1, CumulativeMetersetWeight=1.0)]
    setup = _item(TreatmentDeliveryType="SETUP", BeamNumber=9,
                  ControlPointSequence=[_item(ControlPointIndex=0, NominalBeamEnergy=6.0)])
    ds = dicomio.Dataset()
    ds.SOPClassUID = dicomio.RT_PLAN_STORAGE
    ds.Modality = "RTPLAN"
    ds.FractionGroupSequence = [_item(FractionGroupNumber=1, NumberOfFractionsPlanned=30,
                                      ReferencedBeamSequence=[_item(BeamMeterset=200.0), _item(BeamMeterset=999.0)])]
    ds.BeamSequence = [_item(BeamNumber=1, TreatmentDeliveryType="TREATMENT", ControlPointSequence=cps1), setup,
                       _item(BeamNumber=2, TreatmentDeliveryType="TREATMENT", ControlPointSequence=cps2)]
    meta = dicomio.Dataset()
    meta.MediaStorageSOPClassUID = dicomio.RT_PLAN_STORAGE
    meta.MediaStorageSOPInstanceUID = "1.2.3.4"
    dicomio.dcmwrite(str(path), ds, meta)
    return leaves0, leaves1


def test_read_plan_dicom(tmp_path):
    plan = IMRTPlan("VarianTrueBeamHF")
    bm = plan.beam_models[0]
    n = bm.n_mlc_pairs
    leaves0, leaves1 = _write_rt_plan(tmp_path / "rtplan.dcm", n)
    plan.readPlanDicom(str(tmp_path / "rtplan.dcm"))
    assert plan.n_fractions == 30.0 and plan.n_beams == 2
    b1, b2 = plan.beam_list
    assert b1.dicom_energy_label == '6' and b2.dicom_energy_label == '10'
    np.testing.assert_allclose(b1.iso, [1.5, -2.0, 3.0])
    assert b1.n_cps == 2 and b2.n_cps == 1
    c0, c1 = b1.cp_list
    # meterset of each segment: cumulative-weight step x the first referenced beam's 200 MU
    assert abs(c0.mu - 0.4 * 200) < 1e-9 and abs(c1.mu - 0.6 * 200) < 1e-9
    assert (c0.ga, c0.ca, c0.ta) == (40.0, 10.0, 0.0)
    assert c1.ga == 45.0   # the gantry angle of control point 1 goes with the segment that starts there
    # segment k uses the aperture of control point k
    np.testing.assert_allclose(c0.mlc[:, 0], leaves0[:n], atol=1e-5)
    np.testing.assert_allclose(c0.mlc[:, 1], leaves0[n:], atol=1e-5)
    np.testing.assert_allclose(c1.mlc[:, :2], np.stack([leaves1[:n], leaves1[n:]], axis=1), atol=1e-5)
    np.testing.assert_array_equal(c0.mlc[:, 2], bm.mlc_offsets)
    assert c0.xjaws is None and c0.yjaws is None   # 'ASYMX' / 'ASYMY' are neither 'X' nor 'Y'
    # jaws-only beam: both jaw pairs read from the second item (Y = -35, 45); leaves at the x jaws,
    # pairs whose centre lies outside the y jaws closed
    (c,) = b2.cp_list
    np.testing.assert_array_equal(c.xjaws, [-35, 45])
    np.testing.assert_array_equal(c.yjaws, [-35, 45])
    inside = (bm.mlc_offsets >= -35) & (bm.mlc_offsets <= 45)
    np.testing.assert_array_equal(c.mlc[inside, 0], -35)
    np.testing.assert_array_equal(c.mlc[inside, 1], 45)
    np.testing.assert_array_equal(c.mlc[~inside, :2], 0)
    assert abs(c.mu - 200.0) < 1e-9 and (c.ga, c.ta) == (180.0, 15.0)
