"""DICOM Part 10 files without pydicom: the subset the IMPT API's DICOM
readers and writers need.

The reference reads and writes RT Ion Plans, reads CT series, RT Structure
Sets and RT Dose files and writes RT Dose files through pydicom and
SimpleITK (DoseCUDA/plan_impt.py:17-60, 344-674; plan.py:43-57, 99-124,
133-171), neither of which is installed here.  This module restates the file format of DICOM PS3.5 /
PS3.10 for that use:

* ``dcmread``: a preamble + ``DICM`` file (or, with ``force``, a bare data
  set), the file meta group in explicit VR little endian, then the data set in
  explicit VR little endian, implicit VR little endian, deflated explicit VR
  little endian or explicit VR big endian.  Sequences and items of defined or
  undefined length, encapsulated pixel data (kept as raw bytes), ``UN``
  elements of undefined length (read as implicit-VR sequences, PS3.5 6.2.2).
* ``dcmwrite``: preamble, ``DICM``, the file meta group with its group
  length, then the data set in explicit (default) or implicit VR little
  endian, elements in ascending tag order, sequences and items with defined
  lengths, values padded to even length (``UI``, ``OB`` and ``UN`` with NUL,
  text with a space).
* ``pixel_array``: native 8/16/32-bit single-sample pixel data as numpy.
* ``Dataset``: elements by tag, with attribute access by keyword for the
  keywords of ``_DICTIONARY`` (the attributes the plan reader / writer, the
  structure-set lookup, the CT series reader and the RT Dose reader / writer
  touch).  Reading a missing keyword raises ``AttributeError`` as pydicom
  does, so ``hasattr`` tests work unchanged.

Decoded values: ``DS`` -> float, ``IS`` -> int, other text VRs -> str, binary
numbers -> int / float; several values (backslash-separated text, several
binary numbers) -> list; an empty value -> '' (text) or None (binary);
``SQ`` -> list of Dataset; ``OB`` / ``OW`` / ``UN`` / other -> bytes.
"""

import math
import struct
import uuid
import zlib

IMPLICIT_VR_LE = "1.2.840.10008.1.2"
EXPLICIT_VR_LE = "1.2.840.10008.1.2.1"
DEFLATED_EXPLICIT_VR_LE = "1.2.840.10008.1.2.1.99"
EXPLICIT_VR_BE = "1.2.840.10008.1.2.2"
RT_ION_PLAN_STORAGE = "1.2.840.10008.5.1.4.1.1.481.8"
RT_PLAN_STORAGE = "1.2.840.10008.5.1.4.1.1.481.5"
RT_STRUCTURE_SET_STORAGE = "1.2.840.10008.5.1.4.1.1.481.3"
CT_IMAGE_STORAGE = "1.2.840.10008.5.1.4.1.1.2"
# UUID-derived root (PS3.5 B.2): "2.25." + the decimal value of a UUID
IMPLEMENTATION_CLASS_UID = "2.25.223914627125826349755283436578412155627"
IMPLEMENTATION_VERSION_NAME = "DOSEHIP_1"

_UNDEFINED = 0xFFFFFFFF
_ITEM = 0xFFFEE000
_ITEM_END = 0xFFFEE00D
_SEQ_END = 0xFFFEE0DD
_PIXEL_DATA = 0x7FE00010

# VRs whose explicit-VR header has 2 reserved bytes and a 32-bit length
_LONG_VRS = frozenset(("OB", "OD", "OF", "OL", "OV", "OW", "SQ", "SV", "UC", "UN", "UR", "UT", "UV"))
# text VRs with backslash-separated values / a single value
_TEXT_MULTI = frozenset(("AE", "AS", "CS", "DA", "DS", "DT", "IS", "LO", "PN", "SH", "TM", "UC", "UI"))
_TEXT_SINGLE = frozenset(("LT", "ST", "UT", "UR"))
_BINARY = {"US": "H", "SS": "h", "UL": "I", "SL": "i", "FL": "f", "FD": "d", "UV": "Q", "SV": "q"}
_KNOWN_VRS = _LONG_VRS | _TEXT_MULTI | _TEXT_SINGLE | frozenset(_BINARY) | frozenset(("AT",))

# (group, element, VR, keyword) of the attributes used by the plan readers
# (RT Ion Plan and RT Plan) / writer, the structure-set lookup, the CT series
# reader and the RT Dose reader / writer.
_DICTIONARY = """
0002 0000 UL FileMetaInformationGroupLength
0002 0001 OB FileMetaInformationVersion
0002 0002 UI MediaStorageSOPClassUID
0002 0003 UI MediaStorageSOPInstanceUID
0002 0010 UI TransferSyntaxUID
0002 0012 UI ImplementationClassUID
0002 0013 SH ImplementationVersionName
0008 0005 CS SpecificCharacterSet
0008 0012 DA InstanceCreationDate
0008 0013 TM InstanceCreationTime
0008 0016 UI SOPClassUID
0008 0018 UI SOPInstanceUID
0008 0020 DA StudyDate
0008 0030 TM StudyTime
0008 0060 CS Modality
0008 0070 LO Manufacturer
0008 103E LO SeriesDescription
0008 1150 UI ReferencedSOPClassUID
0008 1155 UI ReferencedSOPInstanceUID
0010 0010 PN PatientName
0010 0020 LO PatientID
0012 0062 CS PatientIdentityRemoved
0012 0063 LO DeidentificationMethod
0018 0050 DS SliceThickness
0018 5100 CS PatientPosition
0020 000D UI StudyInstanceUID
0020 000E UI SeriesInstanceUID
0020 0011 IS SeriesNumber
0020 0013 IS InstanceNumber
0020 0032 DS ImagePositionPatient
0020 0037 DS ImageOrientationPatient
0020 0052 UI FrameOfReferenceUID
0028 0002 US SamplesPerPixel
0028 0004 CS PhotometricInterpretation
0028 0008 IS NumberOfFrames
0028 0009 AT FrameIncrementPointer
0028 0010 US Rows
0028 0011 US Columns
0028 0030 DS PixelSpacing
0028 0100 US BitsAllocated
0028 0101 US BitsStored
0028 0102 US HighBit
0028 0103 US PixelRepresentation
0028 1052 DS RescaleIntercept
0028 1053 DS RescaleSlope
3004 0002 CS DoseUnits
3004 0004 CS DoseType
3004 000A CS DoseSummationType
3004 000C DS GridFrameOffsetVector
3004 000E DS DoseGridScaling
3006 0020 SQ StructureSetROISequence
3006 0022 IS ROINumber
3006 0026 LO ROIName
3006 0084 IS ReferencedROINumber
300A 0002 SH RTPlanLabel
300A 0003 LO RTPlanName
300A 0006 DA RTPlanDate
300A 0007 TM RTPlanTime
300A 0009 LO TreatmentProtocols
300A 000A CS PlanIntent
300A 000C CS RTPlanGeometry
300A 0010 SQ DoseReferenceSequence
300A 0012 IS DoseReferenceNumber
300A 0014 CS DoseReferenceStructureType
300A 0015 CS NominalBeamEnergyUnit
300A 0016 LO DoseReferenceDescription
300A 0020 CS DoseReferenceType
300A 0026 DS TargetPrescriptionDose
300A 0028 DS TargetUnderdoseVolumeFraction
300A 0070 SQ FractionGroupSequence
300A 0071 IS FractionGroupNumber
300A 0078 IS NumberOfFractionsPlanned
300A 0080 IS NumberOfBeams
300A 0086 DS BeamMeterset
300A 00A0 IS NumberOfBrachyApplicationSetups
300A 00B0 SQ BeamSequence
300A 00B2 SH TreatmentMachineName
300A 00B3 CS PrimaryDosimeterUnit
300A 00B6 SQ BeamLimitingDeviceSequence
300A 00B8 CS RTBeamLimitingDeviceType
300A 00BC IS NumberOfLeafJawPairs
300A 00C0 IS BeamNumber
300A 00C2 LO BeamName
300A 00C3 ST BeamDescription
300A 00C4 CS BeamType
300A 00C6 CS RadiationType
300A 00CE CS TreatmentDeliveryType
300A 00D0 IS NumberOfWedges
300A 00E0 IS NumberOfCompensators
300A 00ED IS NumberOfBoli
300A 00F0 IS NumberOfBlocks
300A 00F9 LO AccessoryCode
300A 010C DS CumulativeDoseReferenceCoefficient
300A 010E DS FinalCumulativeMetersetWeight
300A 0110 IS NumberOfControlPoints
300A 0111 SQ ControlPointSequence
300A 0112 IS ControlPointIndex
300A 0114 DS NominalBeamEnergy
300A 011A SQ BeamLimitingDevicePositionSequence
300A 011C DS LeafJawPositions
300A 011E DS GantryAngle
300A 011F CS GantryRotationDirection
300A 0120 DS BeamLimitingDeviceAngle
300A 0121 CS BeamLimitingDeviceRotationDirection
300A 0122 DS PatientSupportAngle
300A 0123 CS PatientSupportRotationDirection
300A 0125 DS TableTopEccentricAngle
300A 0126 CS TableTopEccentricRotationDirection
300A 0128 DS TableTopVerticalPosition
300A 0129 DS TableTopLongitudinalPosition
300A 012A DS TableTopLateralPosition
300A 012C DS IsocenterPosition
300A 0134 DS CumulativeMetersetWeight
300A 0140 FL TableTopPitchAngle
300A 0142 CS TableTopPitchRotationDirection
300A 0144 FL TableTopRollAngle
300A 0146 CS TableTopRollRotationDirection
300A 014A FL GantryPitchAngle
300A 014C CS GantryPitchRotationDirection
300A 0180 SQ PatientSetupSequence
300A 0182 IS PatientSetupNumber
300A 0308 CS ScanMode
300A 0309 CS ModulatedScanModeType
300A 030A FL VirtualSourceAxisDistances
300A 030C SQ SnoutSequence
300A 030D FL SnoutPosition
300A 030F SH SnoutID
300A 0312 IS NumberOfRangeShifters
300A 0314 SQ RangeShifterSequence
300A 0316 IS RangeShifterNumber
300A 0318 SH RangeShifterID
300A 0320 CS RangeShifterType
300A 0330 IS NumberOfLateralSpreadingDevices
300A 0332 SQ LateralSpreadingDeviceSequence
300A 0334 IS LateralSpreadingDeviceNumber
300A 0336 SH LateralSpreadingDeviceID
300A 0338 CS LateralSpreadingDeviceType
300A 033C FL LateralSpreadingDeviceWaterEquivalentThickness
300A 0340 IS NumberOfRangeModulators
300A 0342 SQ RangeModulatorSequence
300A 0344 IS RangeModulatorNumber
300A 0346 SH RangeModulatorID
300A 0348 CS RangeModulatorType
300A 034A LO RangeModulatorDescription
300A 0350 CS PatientSupportType
300A 0352 SH PatientSupportID
300A 035A FL MetersetRate
300A 0360 SQ RangeShifterSettingsSequence
300A 0362 LO RangeShifterSetting
300A 0370 SQ LateralSpreadingDeviceSettingsSequence
300A 0372 LO LateralSpreadingDeviceSetting
300A 0374 FL IsocenterToLateralSpreadingDeviceDistance
300A 0380 SQ RangeModulatorSettingsSequence
300A 0390 SH ScanSpotTuneID
300A 0392 IS NumberOfScanSpotPositions
300A 0394 FL ScanSpotPositionMap
300A 0396 FL ScanSpotMetersetWeights
300A 0398 FL ScanningSpotSize
300A 039A IS NumberOfPaintings
300A 03A2 SQ IonBeamSequence
300A 03A8 SQ IonControlPointSequence
300C 0002 SQ ReferencedRTPlanSequence
300C 0004 SQ ReferencedBeamSequence
300C 0006 IS ReferencedBeamNumber
300C 0050 SQ ReferencedDoseReferenceSequence
300C 0051 IS ReferencedDoseReferenceNumber
300C 0060 SQ ReferencedStructureSetSequence
300C 0102 IS ReferencedLateralSpreadingDeviceNumber
300C 0104 IS ReferencedRangeModulatorNumber
300E 0002 CS ApprovalStatus
7FE0 0010 OW PixelData
"""

TAG_INFO = {}      # tag -> (VR, keyword)
KEYWORD_TAG = {}   # keyword -> tag
for _line in _DICTIONARY.strip().splitlines():
    _g, _e, _vr, _kw = _line.split()
    _t = (int(_g, 16) << 16) | int(_e, 16)
    TAG_INFO[_t] = (_vr, _kw)
    KEYWORD_TAG[_kw] = _t
del _line, _g, _e, _vr, _kw, _t


class InvalidDicomError(Exception):
    pass


def make_tag(tag):
    """(group, element) tuple or 32-bit int -> 32-bit int."""
    if isinstance(tag, tuple):
        return (int(tag[0]) << 16) | int(tag[1])
    return int(tag)


def generate_uid():
    """A new UUID-derived UID (PS3.5 B.2)."""
    return "2.25.%d" % uuid.uuid4().int


class DataElement:
    __slots__ = ("tag", "VR", "value")

    def __init__(self, tag, VR, value):
        self.tag = make_tag(tag)
        self.VR = VR
        self.value = value

    def __repr__(self):
        return "(%04X,%04X) %s %r" % (self.tag >> 16, self.tag & 0xFFFF, self.VR, self.value)


class Dataset:
    """Elements by tag; keywords of ``TAG_INFO`` read and write as attributes."""

    def __init__(self):
        object.__setattr__(self, "_elems", {})

    def __getattr__(self, name):
        tag = KEYWORD_TAG.get(name)
        elems = object.__getattribute__(self, "_elems")
        if tag is None or tag not in elems:
            raise AttributeError("Dataset has no attribute %r" % name)
        return elems[tag].value

    def __setattr__(self, name, value):
        tag = KEYWORD_TAG.get(name)
        if tag is None:
            if name[:1].isupper():
                raise AttributeError("%r is not a DICOM keyword known to this module" % name)
            object.__setattr__(self, name, value)
            return
        self._elems[tag] = DataElement(tag, TAG_INFO[tag][0], value)

    def __delattr__(self, name):
        tag = KEYWORD_TAG.get(name)
        if tag is None:
            object.__delattr__(self, name)
        elif tag in self._elems:
            del self._elems[tag]
        else:
            raise AttributeError(name)

    def __contains__(self, key):
        tag = KEYWORD_TAG.get(key) if isinstance(key, str) else make_tag(key)
        return tag is not None and tag in self._elems

    def __getitem__(self, tag):
        return self._elems[make_tag(tag)]

    def __setitem__(self, tag, elem):
        t = make_tag(tag)
        if not isinstance(elem, DataElement) or elem.tag != t:
            raise ValueError("Dataset[%r] needs a DataElement with that tag" % (tag,))
        self._elems[t] = elem

    def __delitem__(self, tag):
        del self._elems[make_tag(tag)]

    def __len__(self):
        return len(self._elems)

    def __iter__(self):
        return iter(self.elements())

    def add_new(self, tag, VR, value):
        t = make_tag(tag)
        self._elems[t] = DataElement(t, VR, value)

    def get(self, keyword, default=None):
        try:
            return getattr(self, keyword)
        except AttributeError:
            return default

    def elements(self):
        return [self._elems[t] for t in sorted(self._elems)]

    def __repr__(self):
        return "Dataset(%s)" % ", ".join(repr(e) for e in self.elements())


# ---------------------------------------------------------------------------
# Reading

def _text(vr, raw):
    s = raw.decode("latin-1")
    if vr == "UI":
        s = s.rstrip("\0 ")
    if vr in _TEXT_SINGLE:
        return s.rstrip(" \0")
    vals = [v.strip(" \0") for v in s.split("\\")]
    if vr == "DS":
        vals = [float(v) if v else "" for v in vals]
    elif vr == "IS":
        vals = [int(float(v)) if v else "" for v in vals]
    if len(vals) == 1:
        return vals[0]
    return vals


def _decode(vr, raw, endian):
    if vr in _TEXT_MULTI or vr in _TEXT_SINGLE:
        return _text(vr, raw)
    if vr in _BINARY:
        fmt = _BINARY[vr]
        n = len(raw) // struct.calcsize(fmt)
        if n == 0:
            return None
        vals = struct.unpack(endian + fmt * n, raw[:n * struct.calcsize(fmt)])
        return vals[0] if n == 1 else list(vals)
    if vr == "AT":
        n = len(raw) // 4
        vals = [(g << 16) | e for g, e in zip(*[iter(struct.unpack(endian + "H" * (2 * n), raw[:4 * n]))] * 2)]
        return vals[0] if n == 1 else vals
    return bytes(raw)


class _Reader:
    def __init__(self, buf, explicit, little):
        self.buf = buf
        self.explicit = explicit
        self.endian = "<" if little else ">"

    def _u16(self, pos):
        return struct.unpack_from(self.endian + "H", self.buf, pos)[0]

    def _u32(self, pos):
        return struct.unpack_from(self.endian + "I", self.buf, pos)[0]

    def tag_at(self, pos):
        return (self._u16(pos) << 16) | self._u16(pos + 2)

    def header(self, pos, explicit):
        """(tag, VR, length, header bytes) of the element at pos."""
        tag = self.tag_at(pos)
        if tag in (_ITEM, _ITEM_END, _SEQ_END):
            return tag, None, self._u32(pos + 4), 8
        if explicit:
            vr = self.buf[pos + 4:pos + 6].decode("latin-1")
            if vr in _LONG_VRS:
                return tag, vr, self._u32(pos + 8), 12
            if vr not in _KNOWN_VRS:
                raise InvalidDicomError("unknown VR %r at offset %d" % (vr, pos))
            return tag, vr, self._u16(pos + 6), 8
        vr = TAG_INFO.get(tag, ("UN", None))[0]
        if (tag & 0xFFFF) == 0 and tag not in TAG_INFO:
            vr = "UL"  # group length
        return tag, vr, self._u32(pos + 4), 8

    def dataset(self, pos, end, explicit, stop_before_pixels=False):
        """Parse elements from pos up to end or an item delimiter; returns
        (Dataset, position after it)."""
        ds = Dataset()
        while pos < end:
            tag, vr, length, hl = self.header(pos, explicit)
            if tag == _ITEM_END:
                return ds, pos + hl
            if tag in (_ITEM, _SEQ_END):
                raise InvalidDicomError("unexpected item tag at offset %d" % pos)
            if tag == _PIXEL_DATA and stop_before_pixels:
                return ds, end
            pos += hl
            if vr == "SQ" or (length == _UNDEFINED and tag != _PIXEL_DATA):
                # an undefined-length UN holds implicit VR little endian (PS3.5 6.2.2)
                inner_explicit = explicit and vr != "UN"
                items, pos = self.sequence(pos, length, inner_explicit)
                ds._elems[tag] = DataElement(tag, "SQ", items)
                continue
            if length == _UNDEFINED:  # encapsulated pixel data: keep the fragments' bytes
                start = pos
                while True:
                    t, _, ln, h = self.header(pos, explicit)
                    pos += h
                    if t == _SEQ_END:
                        break
                    pos += ln
                ds._elems[tag] = DataElement(tag, vr, bytes(self.buf[start:pos - 8]))
                continue
            if pos + length > len(self.buf):
                raise InvalidDicomError("element (%04X,%04X) runs past the end of the file"
                                        % (tag >> 16, tag & 0xFFFF))
            raw = self.buf[pos:pos + length]
            pos += length
            ds._elems[tag] = DataElement(tag, vr, _decode(vr, raw, self.endian))
        return ds, pos

    def sequence(self, pos, length, explicit):
        items = []
        end = len(self.buf) if length == _UNDEFINED else pos + length
        while pos < end:
            tag, _, ilen, hl = self.header(pos, explicit)
            pos += hl
            if tag == _SEQ_END:
                break
            if tag != _ITEM:
                raise InvalidDicomError("expected an item tag at offset %d" % (pos - hl))
            if ilen == _UNDEFINED:
                ds, pos = self.dataset(pos, len(self.buf), explicit)
            else:
                ds, _ = self.dataset(pos, pos + ilen, explicit)
                pos += ilen
            items.append(ds)
        return items, pos


def _looks_explicit(buf, pos):
    vr = bytes(buf[pos + 4:pos + 6])
    return len(vr) == 2 and vr.isalpha() and vr.isupper() and vr.decode() in _KNOWN_VRS


def dcmread(path, force=False, stop_before_pixels=False):
    """Read a DICOM file; returns a Dataset with the file meta group as its
    ``file_meta`` attribute (an empty Dataset when the file has none)."""
    with open(path, "rb") as fh:
        buf = fh.read()
    pos = 0
    if len(buf) >= 132 and buf[128:132] == b"DICM":
        pos = 132
    elif buf[:4] == b"DICM":
        pos = 4
    elif not force:
        raise InvalidDicomError("%s: no DICM prefix (pass force=True to read a bare data set)" % path)
    meta = Dataset()
    if len(buf) >= pos + 8 and struct.unpack_from("<H", buf, pos)[0] == 0x0002:
        r = _Reader(buf, True, True)
        while pos + 8 <= len(buf) and struct.unpack_from("<H", buf, pos)[0] == 0x0002:
            tag, vr, length, hl = r.header(pos, _looks_explicit(buf, pos))
            raw = buf[pos + hl:pos + hl + length]
            meta._elems[tag] = DataElement(tag, vr, _decode(vr, raw, "<"))
            pos += hl + length
    ts = meta.get("TransferSyntaxUID")
    if ts is None:
        explicit, little = _looks_explicit(buf, pos), True
    elif ts == IMPLICIT_VR_LE:
        explicit, little = False, True
    elif ts == EXPLICIT_VR_BE:
        explicit, little = True, False
    elif ts == DEFLATED_EXPLICIT_VR_LE:
        buf = buf[:pos] + zlib.decompress(buf[pos:], -15)
        explicit, little = True, True
    else:  # explicit VR little endian, and the encapsulated (compressed-pixel) syntaxes
        explicit, little = True, True
    ds, _ = _Reader(buf, explicit, little).dataset(pos, len(buf), explicit, stop_before_pixels)
    object.__setattr__(ds, "file_meta", meta)
    return ds


# ---------------------------------------------------------------------------
# Writing (explicit VR little endian)

def _format_ds(v):
    """Decimal string of at most 16 characters (PS3.5 6.2, VR DS)."""
    x = float(v)
    if not math.isfinite(x):
        raise ValueError("DS value %r is not finite" % (v,))
    s = repr(x)
    if len(s) <= 16:
        return s
    for prec in range(15, 0, -1):
        s = "%.*g" % (prec, x)
        if len(s) <= 16:
            return s
    raise ValueError("DS value %r does not fit 16 characters" % (v,))


def _format_is(v):
    x = float(v)
    if x != int(x):
        raise ValueError("IS value %r is not an integer" % (v,))
    return str(int(x))


def _values(value):
    if value is None:
        return []
    if isinstance(value, (list, tuple)):
        return list(value)
    if hasattr(value, "tolist") and not isinstance(value, (bytes, str)):
        v = value.tolist()
        return v if isinstance(v, list) else [v]
    return [value]


def _encode_value(vr, value):
    if vr in _TEXT_MULTI or vr in _TEXT_SINGLE:
        vals = _values(value)
        if vr == "DS":
            parts = ["" if v == "" or v is None else _format_ds(v) for v in vals]
        elif vr == "IS":
            parts = ["" if v == "" or v is None else _format_is(v) for v in vals]
        else:
            parts = ["" if v is None else str(v) for v in vals]
        raw = "\\".join(parts).encode("ascii")
        pad = b"\0" if vr == "UI" else b" "
        return raw + pad if len(raw) % 2 else raw
    if vr in _BINARY:
        vals = _values(value)
        return struct.pack("<" + _BINARY[vr] * len(vals), *vals)
    if vr == "AT":
        out = b""
        for t in _values(value):
            t = make_tag(t)
            out += struct.pack("<HH", t >> 16, t & 0xFFFF)
        return out
    raw = b"" if value is None else bytes(value)
    return raw + b"\0" if len(raw) % 2 else raw


def _encode_element(elem, explicit):
    t = elem.tag
    head = struct.pack("<HH", t >> 16, t & 0xFFFF)
    if elem.VR == "SQ":
        body = b"".join(struct.pack("<HHI", 0xFFFE, 0xE000, len(b)) + b
                        for b in (_encode_dataset(item, explicit) for item in (elem.value or [])))
    else:
        body = _encode_value(elem.VR, elem.value)
    if not explicit:
        return head + struct.pack("<I", len(body)) + body
    vr = elem.VR.encode("ascii")
    if elem.VR in _LONG_VRS:
        return head + vr + b"\0\0" + struct.pack("<I", len(body)) + body
    if len(body) > 0xFFFF:
        raise ValueError("(%04X,%04X) %s value of %d bytes exceeds the 16-bit length field"
                         % (t >> 16, t & 0xFFFF, elem.VR, len(body)))
    return head + vr + struct.pack("<H", len(body)) + body


def _encode_dataset(ds, explicit=True):
    return b"".join(_encode_element(e, explicit) for e in ds.elements())


def dcmwrite(path, ds, file_meta):
    """Write ``ds`` as a Part 10 file.  The file meta group gets its group
    length, version and implementation identifiers filled in and is written
    in explicit VR little endian; the data set follows in the meta group's
    TransferSyntaxUID, explicit VR little endian (the default) or implicit VR
    little endian."""
    meta = Dataset()
    for e in file_meta.elements():
        meta[e.tag] = e
    ts = meta.get("TransferSyntaxUID", EXPLICIT_VR_LE)
    if ts not in (EXPLICIT_VR_LE, IMPLICIT_VR_LE):
        raise ValueError("dcmwrite writes explicit or implicit VR little endian, not %s" % ts)
    meta.TransferSyntaxUID = ts
    meta.FileMetaInformationVersion = b"\x00\x01"
    if "ImplementationClassUID" not in meta:
        meta.ImplementationClassUID = IMPLEMENTATION_CLASS_UID
    if "ImplementationVersionName" not in meta:
        meta.ImplementationVersionName = IMPLEMENTATION_VERSION_NAME
    if "FileMetaInformationGroupLength" in meta:
        del meta.FileMetaInformationGroupLength
    body = _encode_dataset(meta)
    meta.FileMetaInformationGroupLength = len(body)
    with open(path, "wb") as fh:
        fh.write(b"\0" * 128 + b"DICM")
        fh.write(_encode_dataset(meta))
        fh.write(_encode_dataset(ds, ts == EXPLICIT_VR_LE))


def pixel_array(ds):
    """Native (uncompressed) pixel data as a numpy array (frames, rows,
    columns) or (rows, columns), one sample per pixel, stored values (no
    rescale).  Encapsulated (compressed) pixel data raises NotImplementedError."""
    import numpy as np

    ts = ds.file_meta.get("TransferSyntaxUID", EXPLICIT_VR_LE) if "file_meta" in ds.__dict__ else EXPLICIT_VR_LE
    if ts not in (EXPLICIT_VR_LE, IMPLICIT_VR_LE, EXPLICIT_VR_BE, DEFLATED_EXPLICIT_VR_LE):
        raise NotImplementedError("pixel data in transfer syntax %s (compressed) is not supported" % ts)
    if int(ds.get("SamplesPerPixel", 1)) != 1:
        raise NotImplementedError("only one sample per pixel is supported")
    bits = int(ds.BitsAllocated)
    signed = int(ds.get("PixelRepresentation", 0)) == 1
    kinds = {8: ("i1", "u1"), 16: ("i2", "u2"), 32: ("i4", "u4")}
    if bits not in kinds:
        raise NotImplementedError("BitsAllocated %d is not supported" % bits)
    dt = np.dtype((">" if ts == EXPLICIT_VR_BE else "<") + kinds[bits][0 if signed else 1])
    rows, cols = int(ds.Rows), int(ds.Columns)
    frames = int(ds.get("NumberOfFrames", 1) or 1)
    n = rows * cols * frames
    raw = ds.PixelData
    if len(raw) < n * dt.itemsize:
        raise ValueError("PixelData holds %d bytes for %d pixels of %d bits" % (len(raw), n, bits))
    arr = np.frombuffer(raw[:n * dt.itemsize], dtype=dt).astype(dt.newbyteorder("="))
    return arr.reshape(frames, rows, cols) if "NumberOfFrames" in ds else arr.reshape(rows, cols)
