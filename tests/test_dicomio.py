"""DICOM Part 10 files (DoseCUDA.dicomio) and the RT Ion Plan reader / writer
(IMPTPlan.readPlanDicom / writePlanDicom).

pydicom, which the reference uses for these (plan_impt.py:17-60, 344-674), is
not installed, and the reference holds no DICOM files, so parity with pydicom
is unpinned.  What pins the code here: byte layouts assembled by hand in this
file from DICOM PS3.5 / PS3.10 (explicit VR little endian with defined- and
undefined-length sequences and items, implicit VR little endian, explicit VR
big endian, deflate), read back through the plan reader; an independent
structural walk of what the writer emits (ascending tags, even lengths,
header forms per VR, meta group length); and writer -> reader round trips of
spot lists, angles, isocentres and dose references.
"""

import os
import struct
import zlib

import numpy as np
import pytest

from DoseCUDA import IMPTBeam, IMPTPlan, dicomio
from DoseCUDA.plan_impt import create_dose_reference, get_roi_num

LONG = {"OB", "OD", "OF", "OL", "OV", "OW", "SQ", "SV", "UC", "UN", "UR", "UT", "UV"}


# --- hand assembly (PS3.5 7.1.2 / 7.1.3 / 7.5) ---------------------------------------------------

def el(g, e, vr, value, endian="<"):
    """Explicit-VR element; value bytes are padded to even length."""
    if len(value) % 2:
        value += b"\0" if vr in ("UI", "OB", "UN") else b" "
    head = struct.pack(endian + "HH", g, e) + vr.encode()
    if vr in LONG:
        return head + b"\0\0" + struct.pack(endian + "I", len(value)) + value
    return head + struct.pack(endian + "H", len(value)) + value


def sq(g, e, items, undefined=False, item_undefined=False):
    body = b""
    for it in items:
        if item_undefined:
            body += struct.pack("<HHI", 0xFFFE, 0xE000, 0xFFFFFFFF) + it + struct.pack("<HHI", 0xFFFE, 0xE00D, 0)
        else:
            body += struct.pack("<HHI", 0xFFFE, 0xE000, len(it)) + it
    if undefined:
        return (struct.pack("<HH", g, e) + b"SQ\0\0" + struct.pack("<I", 0xFFFFFFFF) + body
                + struct.pack("<HHI", 0xFFFE, 0xE0DD, 0))
    return struct.pack("<HH", g, e) + b"SQ\0\0" + struct.pack("<I", len(body)) + body


def fl(*v):
    return struct.pack("<%df" % len(v), *v)


def part10(dataset, ts=dicomio.EXPLICIT_VR_LE, sop=dicomio.RT_ION_PLAN_STORAGE):
    meta = (el(0x0002, 0x0001, "OB", b"\x00\x01") + el(0x0002, 0x0002, "UI", sop.encode())
            + el(0x0002, 0x0003, "UI", b"1.2.3.4.5") + el(0x0002, 0x0010, "UI", ts.encode()))
    return b"\0" * 128 + b"DICM" + el(0x0002, 0x0000, "UL", struct.pack("<I", len(meta))) + meta + dataset


def control_point(energy, xy, weights, gantry=b"90", couch=b"0", iso=b"1\\2.5\\-3"):
    return (el(0x300A, 0x0114, "DS", energy) + el(0x300A, 0x011E, "DS", gantry)
            + el(0x300A, 0x0122, "DS", couch) + el(0x300A, 0x012C, "DS", iso)
            + el(0x300A, 0x0392, "IS", str(len(weights)).encode())
            + el(0x300A, 0x0394, "FL", fl(*xy)) + el(0x300A, 0x0396, "FL", fl(*weights)))


def hand_plan():
    """Two beams: one with range shifter '4' (undefined-length sequences and
    items), one without (defined lengths), 4 control points each."""
    cps1 = [control_point(b"70.2", [-5.0, 2.5, 7.25, -1.0], [0.5, 1.25]),
            control_point(b"70.2", [-5.0, 2.5, 7.25, -1.0], [0.0, 0.0]),
            control_point(b"73", [0.0, 0.0], [2.0]),
            control_point(b"73", [0.0, 0.0], [0.0])]
    beam1 = (el(0x300A, 0x00C0, "IS", b"1") + el(0x300A, 0x00C2, "LO", b"LAO")
             + el(0x300A, 0x00C3, "ST", b"left anterior oblique")
             + el(0x300A, 0x0312, "IS", b"1")
             + sq(0x300A, 0x0314, [el(0x300A, 0x0316, "IS", b"1") + el(0x300A, 0x0318, "SH", b"4")],
                  undefined=True, item_undefined=True)
             + sq(0x300A, 0x03A8, cps1, undefined=True, item_undefined=True)
             + el(0x4001, 0x0010, "LO", b"RAYSEARCHLABS 2.0") + el(0x4001, 0x1002, "UN", b"Constant 1.1"))
    cps2 = [control_point(b"71.1", [3.0, -4.0, 1.0, 1.0, -2.0, 0.5], [1.0, 2.0, 3.0], gantry=b"270.5",
                          couch=b"15", iso=b"0\\0\\0"),
            control_point(b"71.1", [3.0, -4.0, 1.0, 1.0, -2.0, 0.5], [0.0, 0.0, 0.0], gantry=b"270.5",
                          couch=b"15", iso=b"0\\0\\0")]
    beam2 = (el(0x300A, 0x00C0, "IS", b"2") + el(0x300A, 0x0312, "IS", b"0")
             + sq(0x300A, 0x03A8, cps2))
    fg = el(0x300A, 0x0071, "IS", b"1") + el(0x300A, 0x0078, "IS", b"25")
    ds = (el(0x0008, 0x0016, "UI", dicomio.RT_ION_PLAN_STORAGE.encode()) + el(0x0008, 0x0060, "CS", b"RTPLAN")
          + el(0x0010, 0x0010, "PN", b"Doe^John") + sq(0x300A, 0x0070, [fg])
          + sq(0x300A, 0x03A2, [beam1, beam2], undefined=True))
    return part10(ds)


def test_reader_hand_assembled_plan(tmp_path):
    p = tmp_path / "plan.dcm"
    p.write_bytes(hand_plan())
    ds = dicomio.dcmread(str(p))
    assert ds.file_meta.TransferSyntaxUID == dicomio.EXPLICIT_VR_LE and ds.Modality == "RTPLAN"
    assert ds.PatientName == "Doe^John"
    b1 = ds.IonBeamSequence[0]
    assert b1[(0x4001, 0x0010)].value == "RAYSEARCHLABS 2.0" and b1[(0x4001, 0x1002)].value == b"Constant 1.1"
    assert b1.IonControlPointSequence[0].IsocenterPosition == [1.0, 2.5, -3.0]
    assert not hasattr(ds.IonBeamSequence[1], "BeamName") and "BeamName" in b1

    plan = IMPTPlan()
    plan.readPlanDicom(str(p))
    assert plan.n_fractions == 25.0 and plan.n_beams == 2
    e70, e73, e71 = (plan.beam_models[0].energyIDFromLabel(v) for v in (70.2, 73.0, 71.1))
    assert (e70, e73, e71) == (0, 3, 1)
    a, b = plan.beam_list
    assert a.dicom_rangeshifter_label == "4" and b.dicom_rangeshifter_label == "0"
    assert (a.gantry_angle, a.couch_angle, b.gantry_angle, b.couch_angle) == (90.0, 0.0, 270.5, 15.0)
    assert a.iso.dtype == np.single and np.array_equal(a.iso, [1.0, 2.5, -3.0])
    assert (a.BeamName, a.BeamDescription) == ("LAO", "left anterior oblique")
    assert (b.BeamName, b.BeamDescription) == ("PBS_Beam2", "PBS_Beam 2")   # addBeam's defaults
    # even control points only; float32 rows (x, y, mu, energy id)
    want_a = np.array([[-5.0, 2.5, 0.5, e70], [7.25, -1.0, 1.25, e70], [0.0, 0.0, 2.0, e73]], np.single)
    want_b = np.array([[3.0, -4.0, 1.0, e71], [1.0, 1.0, 2.0, e71], [-2.0, 0.5, 3.0, e71]], np.single)
    assert a.spot_list.dtype == np.single and np.array_equal(a.spot_list, want_a) and a.n_spots == 3
    assert np.array_equal(b.spot_list, want_b) and b.n_spots == 3


def test_reader_other_transfer_syntaxes(tmp_path):
    # implicit VR little endian: tag, 32-bit length, value; VRs from the dictionary
    def iel(g, e, value):
        if len(value) % 2:
            value += b" "
        return struct.pack("<HHI", g, e, len(value)) + value
    item = iel(0x3006, 0x0022, b"7") + iel(0x3006, 0x0026, b"PTV")
    ds = (iel(0x0010, 0x0020, b"ID42") + iel(0x0020, 0x0013, b"12")
          + struct.pack("<HHI", 0x3006, 0x0020, 0xFFFFFFFF)           # undefined-length sequence
          + struct.pack("<HHI", 0xFFFE, 0xE000, len(item)) + item
          + struct.pack("<HHI", 0xFFFE, 0xE0DD, 0)
          + iel(0x0019, 0x1001, b"\x01\x02\x03\x04"))                  # private: no dictionary VR -> bytes
    p = tmp_path / "imp.dcm"
    p.write_bytes(part10(ds, ts=dicomio.IMPLICIT_VR_LE, sop=dicomio.RT_STRUCTURE_SET_STORAGE))
    r = dicomio.dcmread(str(p))
    assert r.PatientID == "ID42" and r.InstanceNumber == 12
    assert get_roi_num(r, "PTV") == 7 and get_roi_num(r, "CTV") is None
    assert r[(0x0019, 0x1001)].value == b"\x01\x02\x03\x04"

    # explicit VR big endian (data set only; the meta group stays little endian)
    be = el(0x0028, 0x0010, "US", struct.pack(">H", 512), ">") + el(0x300A, 0x0394, "FL", struct.pack(">2f", 1.5, -2.0), ">")
    q = tmp_path / "be.dcm"
    q.write_bytes(part10(be, ts=dicomio.EXPLICIT_VR_BE))
    r = dicomio.dcmread(str(q))
    assert r[(0x0028, 0x0010)].value == 512 and r.ScanSpotPositionMap == [1.5, -2.0]

    # deflated explicit VR little endian: raw deflate of the data set
    raw = el(0x0010, 0x0020, "LO", b"DEFLATED") + el(0x300A, 0x012C, "DS", b"1.25\\-7\\3e2")
    c = zlib.compressobj(9, zlib.DEFLATED, -15)
    d = tmp_path / "df.dcm"
    d.write_bytes(part10(c.compress(raw) + c.flush(), ts=dicomio.DEFLATED_EXPLICIT_VR_LE))
    r = dicomio.dcmread(str(d))
    assert r.PatientID == "DEFLATED" and r.IsocenterPosition == [1.25, -7.0, 300.0]

    # a bare data set (no preamble, no meta group) needs force
    bare = tmp_path / "bare.dcm"
    bare.write_bytes(el(0x0010, 0x0020, "LO", b"BARE"))
    with pytest.raises(dicomio.InvalidDicomError):
        dicomio.dcmread(str(bare))
    assert dicomio.dcmread(str(bare), force=True).PatientID == "BARE"


def walk(buf, pos, end, out, depth=0):
    """Independent structural walk of explicit VR little endian elements:
    records (depth, tag, VR, length) and checks ascending tags and even lengths."""
    last = -1
    while pos < end:
        g, e = struct.unpack_from("<HH", buf, pos)
        tag = (g << 16) | e
        assert tag > last, "tags out of order at %d" % pos
        last = tag
        vr = buf[pos + 4:pos + 6].decode()
        if vr in LONG:
            assert buf[pos + 6:pos + 8] == b"\0\0"
            (n,), hl = struct.unpack_from("<I", buf, pos + 8), 12
        else:
            (n,), hl = struct.unpack_from("<H", buf, pos + 6), 8
        assert n != 0xFFFFFFFF and n % 2 == 0
        out.append((depth, tag, vr, n))
        pos += hl
        if vr == "SQ":
            q = pos
            while q < pos + n:
                ig, ie, ilen = struct.unpack_from("<HHI", buf, q)
                assert (ig, ie) == (0xFFFE, 0xE000)
                walk(buf, q + 8, q + 8 + ilen, out, depth + 1)
                q += 8 + ilen
            assert q == pos + n
        pos += n
    assert pos == end


def make_ct_dir(d, n=3):
    os.makedirs(d, exist_ok=True)
    for i in (3, 1, 2)[:n]:
        ct = dicomio.Dataset()
        ct.SOPClassUID = dicomio.CT_IMAGE_STORAGE
        ct.SOPInstanceUID = dicomio.generate_uid()
        ct.Modality = "CT"
        ct.PatientName = "Doe^Jane" if i == 1 else "Other^Slice"
        ct.PatientID = "P%d" % i
        ct.StudyInstanceUID = "1.2.826.0.1.%d" % i
        ct.FrameOfReferenceUID = "1.2.826.0.9.%d" % i
        ct.InstanceNumber = i
        ct.PixelData = bytes(8)
        meta = dicomio.Dataset()
        meta.MediaStorageSOPClassUID = dicomio.CT_IMAGE_STORAGE
        meta.MediaStorageSOPInstanceUID = ct.SOPInstanceUID
        dicomio.dcmwrite(os.path.join(d, "slice%d.dcm" % i), ct, meta)


def demo_plan():
    plan = IMPTPlan()
    rng = np.random.default_rng(11)
    for gantry, couch, rs in ((0.0, 0.0, "0"), (120.0, 10.0, "4")):
        b = IMPTBeam()
        b.gantry_angle, b.couch_angle, b.dicom_rangeshifter_label = gantry, couch, rs
        b.iso = np.array([1.5, -2.25, 12.3], np.single)      # 12.3 needs the 16-character DS limit
        rows = [np.column_stack([rng.uniform(-20, 20, (6, 2)), rng.uniform(0.1, 1, 6), np.full(6, e)])
                for e in (20, 35, 21)]
        b.spot_list = np.vstack(rows).astype(np.single)
        b.n_spots = len(b.spot_list)
        plan.addBeam(b)
    plan.n_fractions = 3
    return plan


def test_writer_structure_and_round_trip(tmp_path):
    ct_dir = str(tmp_path / "ct")
    make_ct_dir(ct_dir)
    plan = demo_plan()
    out = str(tmp_path / "plan.dcm")
    plan.writePlanDicom(ct_dir, out)

    buf = open(out, "rb").read()
    assert buf[:128] == b"\0" * 128 and buf[128:132] == b"DICM"
    (glen,) = struct.unpack_from("<I", buf, 140)
    assert buf[132:140] == b"\x02\x00\x00\x00UL\x04\x00"
    meta_end = 144 + glen
    assert struct.unpack_from("<H", buf, meta_end)[0] != 0x0002    # the group length covers the meta group
    rec = []
    walk(buf, 132, len(buf), rec)
    vr_of = {t: v for _, t, v, _ in rec}
    for tag, (vr, kw) in dicomio.TAG_INFO.items():
        if tag in vr_of:
            assert vr_of[tag] == vr, kw
    n_cp = sum(1 for d, t, _, _ in rec if t == dicomio.KEYWORD_TAG["ControlPointIndex"])
    assert n_cp == 2 * 3 * 2      # 2 beams x 3 layers x 2 control points

    ds = dicomio.dcmread(out)
    assert ds.file_meta.MediaStorageSOPInstanceUID == ds.SOPInstanceUID
    assert (ds.PatientName, ds.PatientID) == ("Doe^Jane", "P1")          # lowest InstanceNumber
    assert ds.NumberOfBeams == 2 and ds.FractionGroupSequence[0].NumberOfFractionsPlanned == 3
    assert [r.ReferencedBeamNumber for r in ds.FractionGroupSequence[0].ReferencedBeamSequence] == [1, 2]
    dr = ds.DoseReferenceSequence[0]
    assert (dr.ReferencedROINumber, dr.TargetPrescriptionDose, dr.DoseReferenceDescription) == (1, 6.0, "PTV Prescription")
    assert dr[(0x4001, 0x1011)].value == b"6.0\0"       # odd-length UN value padded with NUL
    beam = ds.IonBeamSequence[1]
    assert beam.RangeShifterSequence[0].RangeShifterID == "4" and beam.VirtualSourceAxisDistances == [1340.0, 1930.0]
    cps = beam.IonControlPointSequence
    assert beam.NumberOfControlPoints == 6 and [c.ControlPointIndex for c in cps] == list(range(6))
    energies = [c.NominalBeamEnergy for c in cps]
    labels = plan.beam_models[1].energy_labels
    assert energies == [labels[35], labels[35], labels[21], labels[21], labels[20], labels[20]]
    mu = plan.beam_list[1].spot_list[:, 2]
    per_layer = [np.sum(mu[plan.beam_list[1].spot_list[:, 3] == e]) for e in (35, 21, 20)]
    cum = [c.CumulativeMetersetWeight for c in cps]
    assert np.allclose(cum, [0, per_layer[0], per_layer[0], sum(per_layer[:2]), sum(per_layer[:2]), sum(per_layer)],
                       rtol=1e-6)
    assert all(w == 0.0 for w in cps[1].ScanSpotMetersetWeights)
    sig = plan.beam_models[1].reference_sigmas[35] * 2.0 * np.sqrt(2.0 * np.log(2.0)) * 10.0
    assert np.allclose(cps[0].ScanningSpotSize, [sig, sig], rtol=1e-6)

    back = IMPTPlan()
    back.readPlanDicom(out)
    assert back.n_fractions == 3.0 and back.n_beams == 2
    for a, b in zip(plan.beam_list, back.beam_list):
        order = np.argsort(-a.spot_list[:, 3], kind="stable")      # layers by descending energy id
        assert np.array_equal(b.spot_list, a.spot_list[order])
        assert np.array_equal(b.iso, a.iso) and b.iso.dtype == np.single
        assert (b.gantry_angle, b.couch_angle) == (a.gantry_angle, a.couch_angle)
        assert (b.dicom_rangeshifter_label, b.BeamName, b.BeamDescription) == (
            a.dicom_rangeshifter_label, a.BeamName, a.BeamDescription)


def test_dose_references_from_structure_set(tmp_path):
    rs = dicomio.Dataset()
    rs.SOPInstanceUID = "1.2.3.99"
    rois = []
    for num, name in ((4, "BODY"), (9, "PTV_high"), (11, "PTV_low")):
        r = dicomio.Dataset()
        r.ROINumber, r.ROIName = num, name
        rois.append(r)
    rs.StructureSetROISequence = rois
    meta = dicomio.Dataset()
    meta.MediaStorageSOPClassUID = dicomio.RT_STRUCTURE_SET_STORAGE
    meta.MediaStorageSOPInstanceUID = rs.SOPInstanceUID
    rs_path = str(tmp_path / "rs.dcm")
    dicomio.dcmwrite(rs_path, rs, meta)

    ct_dir = str(tmp_path / "ct")
    make_ct_dir(ct_dir, 1)
    plan = demo_plan()
    plan.addPrescription(60.0, "PTV_high", 0.05)
    plan.addPrescription(54.0, "PTV_low", 0.02)
    out = str(tmp_path / "plan.dcm")
    plan.writePlanDicom(ct_dir, out, rt_struct_file=rs_path)
    ds = dicomio.dcmread(out)
    assert [(d.DoseReferenceNumber, d.ReferencedROINumber, d.TargetPrescriptionDose)
            for d in ds.DoseReferenceSequence] == [(1, 9, 60.0), (2, 11, 54.0)]
    assert ds.ReferencedStructureSetSequence[0].ReferencedSOPInstanceUID == "1.2.3.99"

    plan.addPrescription(50.0, "GTV", 0.0)
    with pytest.raises(ValueError, match="GTV"):
        plan.writePlanDicom(ct_dir, out, rt_struct_file=rs_path)
    # without a structure set the ROI numbers are the prescriptions' positions
    d2 = create_dose_reference(dicomio.Dataset(), plan.Prescriptions)
    assert [d.ReferencedROINumber for d in d2.DoseReferenceSequence] == [1, 2, 3]

    plan.beam_list[0].dicom_rangeshifter_label = "7"
    with pytest.raises(ValueError, match="rangeshifter ID 7"):
        plan.writePlanDicom(ct_dir, out)


def test_value_encoding_rules():
    ds = dicomio.Dataset()
    for v in (12.300000190734863, -0.000123456789012345, 1e-300, 123456789.123456789, float(np.float32(1 / 3))):
        ds.IsocenterPosition = [v]
        raw = dicomio._encode_value("DS", [v])
        assert len(raw.rstrip(b" ")) <= 16 and len(raw) % 2 == 0
        assert np.float32(float(raw)) == np.float32(v)
    with pytest.raises(ValueError):
        dicomio._encode_value("IS", [2.5])
    assert dicomio._encode_value("UI", "1.2.3") == b"1.2.3\0"
    assert dicomio._encode_value("CS", ["A", "B"]) == b"A\\B " and dicomio._encode_value("CS", "AB") == b"AB"
    assert dicomio._encode_value("DS", None) == b"" and dicomio._encode_value("FL", [1.0, 2.0]) == fl(1.0, 2.0)
    with pytest.raises(AttributeError):
        ds.NotADicomKeyword = 1
    with pytest.raises(AttributeError):
        _ = ds.PatientName
    assert dicomio.generate_uid().startswith("2.25.") and len(dicomio.generate_uid()) <= 64


# --- CT series and RT Dose (plan.py:43-57, 99-124, 133-171) ---------------------------------------

def write_ds(path, ds, sop, ts=dicomio.EXPLICIT_VR_LE):
    meta = dicomio.Dataset()
    meta.MediaStorageSOPClassUID = sop
    meta.MediaStorageSOPInstanceUID = ds.get("SOPInstanceUID", "1.2.3")
    meta.TransferSyntaxUID = ts
    dicomio.dcmwrite(str(path), ds, meta)


def ct_slice(z, stored, series="1.2.5.1", slope=1.0, intercept=-1024.0, signed=False):
    ds = dicomio.Dataset()
    ds.SOPClassUID = dicomio.CT_IMAGE_STORAGE
    ds.SOPInstanceUID = dicomio.generate_uid()
    ds.Modality = "CT"
    ds.SeriesInstanceUID = series
    ds.ImagePositionPatient = [-12.5, 30.0, z]
    ds.PixelSpacing = [0.8, 0.6]             # row spacing (y), column spacing (x)
    ds.Rows, ds.Columns = stored.shape
    ds.SamplesPerPixel = 1
    ds.BitsAllocated, ds.BitsStored, ds.HighBit = 16, 16, 15
    ds.PixelRepresentation = 1 if signed else 0
    ds.RescaleSlope, ds.RescaleIntercept = slope, intercept
    ds.PixelData = stored.astype("<i2" if signed else "<u2").tobytes()
    return ds


def test_load_ct_series(tmp_path):
    from DoseCUDA import DoseGrid
    rng = np.random.default_rng(5)
    zs = [17.5, 10.0, 15.0, 12.5]
    stored = {z: rng.integers(0, 3000, size=(3, 5)) for z in zs}
    stored[10.0][0, 0] = 0                  # -1024 HU: clipped to -1000
    for n, z in enumerate(zs):               # file names out of z order, both little-endian syntaxes
        ts = dicomio.IMPLICIT_VR_LE if n % 2 else dicomio.EXPLICIT_VR_LE
        write_ds(tmp_path / ("im%d.dcm" % n), ct_slice(z, stored[z]), dicomio.CT_IMAGE_STORAGE, ts)
    write_ds(tmp_path / "other.dcm", ct_slice(0.0, stored[10.0], series="1.2.9"), dicomio.CT_IMAGE_STORAGE)
    (tmp_path / "notes.txt").write_text("not DICOM")
    g = DoseGrid()
    g.loadCTDCM(str(tmp_path))
    want = np.stack([stored[z] for z in sorted(zs)]).astype(np.float64) - 1024.0
    assert g.HU.dtype == np.single and np.array_equal(g.HU, np.clip(want, -1000, None).astype(np.single))
    assert g.HU[0, 0, 0] == -1000.0
    assert np.array_equal(g.origin, [-12.5, 30.0, 10.0]) and np.allclose(g.spacing, [0.6, 0.8, 2.5], rtol=0, atol=1e-12)
    assert tuple(g.size) == (4, 3, 5)

    # signed storage, fractional slope
    d2 = tmp_path / "signed"
    d2.mkdir()
    s = np.array([[-300, 0, 7], [5, -2, 1000]])
    for n, z in enumerate((0.0, 3.0)):
        write_ds(d2 / ("s%d.dcm" % n), ct_slice(z, s + n, slope=0.5, intercept=-10.0, signed=True),
                 dicomio.CT_IMAGE_STORAGE)
    g.loadCTDCM(str(d2))
    assert np.array_equal(g.HU, np.stack([s * 0.5 - 10.0, (s + 1) * 0.5 - 10.0]).astype(np.single))
    assert g.spacing[2] == 3.0
    (tmp_path / "empty").mkdir()
    with pytest.raises(RuntimeError):
        g.loadCTDCM(str(tmp_path / "empty"))


def rt_dose():
    ds = dicomio.Dataset()
    ds.SOPClassUID = "1.2.840.10008.5.1.4.1.1.481.2"
    ds.SOPInstanceUID = "1.2.3.77"
    ds.Modality = "RTDOSE"
    ds.SeriesDescription = "TPS dose"
    ds.ImagePositionPatient = [-5.0, -3.0, 11.0]
    ds.PixelSpacing = [2.0, 2.0]
    ds.Rows, ds.Columns, ds.NumberOfFrames = 4, 6, 3
    ds.GridFrameOffsetVector = [0.0, 2.5, 5.0]
    ds.SamplesPerPixel, ds.BitsAllocated, ds.BitsStored, ds.HighBit, ds.PixelRepresentation = 1, 16, 16, 15, 0
    ds.DoseUnits, ds.DoseType, ds.DoseSummationType, ds.DoseGridScaling = "GY", "PHYSICAL", "PLAN", 1e-3
    ds.ReferencedRTPlanSequence = [_item([("ReferencedSOPClassUID", dicomio.RT_ION_PLAN_STORAGE),
                                          ("ReferencedSOPInstanceUID", "1.2.3.4")])]
    ds.PixelData = bytes(2 * 4 * 6 * 3)
    ds.add_new((0x3005, 0x1001), "UN", b"vendor")    # implicit VR: kept as raw bytes
    return ds


def _item(pairs):
    it = dicomio.Dataset()
    for k, v in pairs:
        setattr(it, k, v)
    return it


@pytest.mark.parametrize("ts", [dicomio.EXPLICIT_VR_LE, dicomio.IMPLICIT_VR_LE])
def test_rt_dose_resample_and_write(tmp_path, ts):
    from DoseCUDA import DoseGrid
    ref_path = tmp_path / "ref.dcm"
    write_ds(ref_path, rt_dose(), "1.2.840.10008.5.1.4.1.1.481.2", ts)
    g = DoseGrid()
    g.createCubePhantom(size=[10, 9, 8], spacing=3.0)
    g.HU = np.random.default_rng(2).uniform(-1000, 1500, size=(10, 9, 8)).astype(np.single)
    hu0, origin0, spacing0 = g.HU.copy(), g.origin.copy(), g.spacing.copy()
    g.resampleCTfromReferenceDose(str(ref_path))
    assert g.HU.shape == (3, 4, 6) and tuple(g.size) == (3, 4, 6) and g.HU.dtype == np.single
    assert np.array_equal(g.origin, [-5.0, -3.0, 11.0]) and np.array_equal(g.spacing, [2.0, 2.0, 2.5])
    want = imageio_resample(hu0, origin0, spacing0)
    assert np.array_equal(g.HU, want)

    rng = np.random.default_rng(9)
    g.dose = rng.uniform(0, 60, size=(3, 4, 6)).astype(np.single)
    g.beam_doses = [g.dose * 0.25, g.dose * 0.75]
    out = str(tmp_path / "dose.dcm")
    g.writeDoseDCM(out, str(ref_path))
    r = dicomio.dcmread(out)
    assert r.file_meta.TransferSyntaxUID == ts
    assert np.array_equal(dicomio.pixel_array(r), np.array(g.dose * 500.0 * 1.1, dtype=np.uint16))
    assert (r.SeriesDescription, r.DoseSummationType, r.DoseType, r.DoseGridScaling) == (
        "TPS dose_DoseCUDA", "PLAN", "EFFECTIVE", 0.002)
    uid = r.ReferencedRTPlanSequence[0].ReferencedSOPInstanceUID
    assert uid.startswith("2.25.") and uid != "1.2.3.4"
    assert r[(0x3005, 0x1001)].value == b"vendor" and r.GridFrameOffsetVector == [0.0, 2.5, 5.0]
    g.writeDoseDCM(out, str(ref_path), dose_type="PHYSICAL", individual_beams=True)
    for i, bd in enumerate(g.beam_doses):
        r = dicomio.dcmread(out.replace(".dcm", "_beam%02i.dcm" % (i + 1)))
        assert r.DoseSummationType == "BEAM" and r.DoseType == "PHYSICAL"
        assert np.array_equal(dicomio.pixel_array(r), np.array(bd * 500.0, dtype=np.uint16))
    with pytest.raises(Exception, match=".dcm extension"):
        g.writeDoseDCM(str(tmp_path / "d.nrrd"), str(ref_path))
    with pytest.raises(Exception, match="Unknown dose type"):
        g.writeDoseDCM(out, str(ref_path), dose_type="RBE")
    g.dose = np.zeros((3, 4, 5), np.single)
    with pytest.raises(ValueError):
        g.writeDoseDCM(out, str(ref_path))


def imageio_resample(hu, origin, spacing):
    from DoseCUDA import imageio
    return np.array(imageio.resample_linear(hu, origin, spacing, np.array([-5.0, -3.0, 11.0]),
                                            np.array([2.0, 2.0, 2.5]), (6, 4, 3), default=-1000.0), dtype=np.single)
