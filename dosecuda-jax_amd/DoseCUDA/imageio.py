"""NRRD volumes and linear CT resampling without SimpleITK.

The reference reads and writes NRRD files and resamples CTs through SimpleITK
(DoseCUDA/plan.py:33-41, 59-97, 173-212; plan_impt.py:259-287), which is not
installed here.  This module covers what those calls do for the dose path:

* ``write_nrrd`` / ``read_nrrd``: NRRD0004 volumes with an attached header,
  axis order x, y, z in the file and (z, y, x) in numpy (as
  ``sitk.GetImageFromArray`` / ``GetArrayFromImage``), origin and spacing in
  mm, identity directions (the reference never sets any); raw or gzip
  encoding, little or big endian on reading.
* ``resample_linear``: ``sitk.ResampleImageFilter`` with its defaults as the
  reference uses it (identity transform and directions, linear
  interpolation, a default value outside the input): an output voxel's
  continuous input index ``(o + i s' - o_in) / s_in`` per axis must lie in
  [-0.5, n - 0.5), else it gets the default; inside, the value is trilinear
  in double precision with neighbours beyond the first / last voxel clamped
  to it (ITK's LinearInterpolateImageFunction), cast to the input dtype.
"""

import gzip
import os

import numpy as np

_TYPES = {  # NRRD type names -> numpy dtypes
    "float": np.float32, "double": np.float64,
    "short": np.int16, "int16": np.int16, "signed short": np.int16, "short int": np.int16,
    "ushort": np.uint16, "uint16": np.uint16, "unsigned short": np.uint16,
    "int": np.int32, "int32": np.int32, "signed int": np.int32,
    "uint": np.uint32, "uint32": np.uint32, "unsigned int": np.uint32,
    "char": np.int8, "int8": np.int8, "signed char": np.int8,
    "uchar": np.uint8, "uint8": np.uint8, "unsigned char": np.uint8,
}
_NAMES = {np.dtype(np.float32): "float", np.dtype(np.float64): "double", np.dtype(np.int16): "short",
          np.dtype(np.uint16): "ushort", np.dtype(np.int32): "int", np.dtype(np.uint32): "uint",
          np.dtype(np.int8): "signed char", np.dtype(np.uint8): "uchar"}


def _vec(v):
    return "(" + ",".join(repr(float(x)) for x in v) + ")"


def write_nrrd(path, volume, origin, spacing, encoding="raw"):
    """Write a (z, y, x) volume with origin / spacing (x, y, z) in mm."""
    vol = np.ascontiguousarray(volume)
    if vol.ndim != 3:
        raise ValueError("write_nrrd: volume must be 3-D (z, y, x), got shape %s" % (vol.shape,))
    if vol.dtype not in _NAMES:
        raise ValueError("write_nrrd: unsupported dtype %s" % vol.dtype)
    if encoding not in ("raw", "gzip"):
        raise ValueError("write_nrrd: encoding must be 'raw' or 'gzip'")
    nz, ny, nx = vol.shape
    sx, sy, sz = (float(s) for s in spacing)
    header = ["NRRD0004",
              "# written by DoseCUDA.imageio",
              "type: %s" % _NAMES[vol.dtype],
              "dimension: 3",
              "space: left-posterior-superior",
              "sizes: %d %d %d" % (nx, ny, nz),
              "space directions: %s %s %s" % (_vec((sx, 0, 0)), _vec((0, sy, 0)), _vec((0, 0, sz))),
              "kinds: domain domain domain",
              "endian: little",
              "encoding: %s" % encoding,
              "space origin: %s" % _vec(origin)]
    data = vol.astype(vol.dtype.newbyteorder("<"), copy=False).tobytes()
    if encoding == "gzip":
        data = gzip.compress(data)
    with open(path, "wb") as f:
        f.write(("\n".join(header) + "\n\n").encode("ascii"))
        f.write(data)


def _parse_vector(text):
    text = text.strip()
    if text == "none":
        return None
    return [float(x) for x in text.strip("()").split(",")]


def read_nrrd(path):
    """Read a 3-D NRRD volume: returns (volume (z, y, x), origin (x, y, z),
    spacing (x, y, z)) with float64 origin / spacing."""
    with open(path, "rb") as f:
        raw = f.read()
    if not raw.startswith(b"NRRD000"):
        raise ValueError("%s: not an NRRD file" % path)
    end = raw.find(b"\n\n")
    if end < 0:
        raise ValueError("%s: NRRD header not terminated" % path)
    fields = {}
    for line in raw[:end].decode("ascii", "replace").split("\n")[1:]:
        if not line or line.startswith("#"):
            continue
        if ":=" in line:            # key/value pairs
            continue
        key, _, val = line.partition(":")
        fields[key.strip().lower()] = val.strip()
    if "data file" in fields or "datafile" in fields:
        raise NotImplementedError("%s: detached NRRD data files are not supported" % path)
    if int(fields.get("dimension", 0)) != 3:
        raise ValueError("%s: only 3-D volumes are supported" % path)
    dtype = _TYPES.get(fields.get("type", "").lower())
    if dtype is None:
        raise ValueError("%s: unsupported NRRD type %r" % (path, fields.get("type")))
    sizes = [int(v) for v in fields["sizes"].split()]
    spacing = [1.0, 1.0, 1.0]
    if "space directions" in fields:
        vecs = [_parse_vector(t) for t in fields["space directions"].replace(") (", ");(").split(";")]
        for a, v in enumerate(vecs):
            if v is None:
                continue
            off = [abs(x) for b, x in enumerate(v) if b != a]
            if any(x > 1e-9 * max(1.0, abs(v[a])) for x in off) or v[a] <= 0:
                raise NotImplementedError("%s: only positive axis-aligned space directions are supported" % path)
            spacing[a] = v[a]
    elif "spacings" in fields:
        spacing = [float(v) for v in fields["spacings"].split()]
    origin = [0.0, 0.0, 0.0]
    if "space origin" in fields:
        origin = _parse_vector(fields["space origin"])
    enc = fields.get("encoding", "raw").lower()
    data = raw[end + 2:]
    if enc in ("gzip", "gz"):
        data = gzip.decompress(data)
    elif enc != "raw":
        raise NotImplementedError("%s: NRRD encoding %r is not supported" % (path, enc))
    big = fields.get("endian", "little").lower() == "big"
    dt = np.dtype(dtype).newbyteorder(">" if big else "<")
    n = sizes[0] * sizes[1] * sizes[2]
    if len(data) < n * dt.itemsize:
        raise ValueError("%s: %d data bytes for %d voxels" % (path, len(data), n))
    vol = np.frombuffer(data[:n * dt.itemsize], dtype=dt).astype(dtype).reshape(sizes[2], sizes[1], sizes[0])
    return vol, np.array(origin, dtype=np.float64), np.array(spacing, dtype=np.float64)


def _axis_weights(n_in, o_in, s_in, n_out, o_out, s_out):
    """Per output index: lower neighbour, upper neighbour (clamped), upper
    weight, and whether the continuous index lies in [-0.5, n_in - 0.5)."""
    ci = (o_out + np.arange(n_out, dtype=np.float64) * s_out - o_in) / s_in
    inside = (ci >= -0.5) & (ci < n_in - 0.5)
    base = np.floor(ci)
    w = ci - base
    lo = np.clip(base, 0, n_in - 1).astype(np.int64)
    hi = np.clip(base + 1, 0, n_in - 1).astype(np.int64)
    return lo, hi, w, inside


def resample_linear(volume, origin, spacing, new_origin, new_spacing, new_size, default=-1000.0):
    """Resample a (z, y, x) volume onto the grid new_origin / new_spacing /
    new_size (all x, y, z), as sitk.ResampleImageFilter with its defaults."""
    vol = np.asarray(volume)
    nz, ny, nx = vol.shape
    nxo, nyo, nzo = (int(v) for v in new_size)
    wx = _axis_weights(nx, float(origin[0]), float(spacing[0]), nxo, float(new_origin[0]), float(new_spacing[0]))
    wy = _axis_weights(ny, float(origin[1]), float(spacing[1]), nyo, float(new_origin[1]), float(new_spacing[1]))
    wz = _axis_weights(nz, float(origin[2]), float(spacing[2]), nzo, float(new_origin[2]), float(new_spacing[2]))
    out = np.empty((nzo, nyo, nxo), dtype=vol.dtype)
    ok_xy = wy[3][:, None] & wx[3][None, :]
    for k in range(nzo):
        if not wz[3][k]:
            out[k] = default
            continue
        # trilinear = product of per-axis weights, summed over the 8 corners (double precision)
        acc = np.zeros((nyo, nxo), dtype=np.float64)
        for zc, wzc in ((wz[0][k], 1.0 - wz[2][k]), (wz[1][k], wz[2][k])):
            sl = vol[zc].astype(np.float64)
            for yc, wyc in ((wy[0], 1.0 - wy[2]), (wy[1], wy[2])):
                rows = sl[yc]
                for xc, wxc in ((wx[0], 1.0 - wx[2]), (wx[1], wx[2])):
                    acc += (wzc * wyc[:, None] * wxc[None, :]) * rows[:, xc]
        out[k] = np.where(ok_xy, acc, default).astype(vol.dtype)
    return out


def check_extension(path, ext, what):
    if not path.endswith(ext):
        raise Exception("%s path must have %s extension" % (what, ext))
    d = os.path.dirname(path)
    if d and not os.path.isdir(d):
        raise FileNotFoundError("%s: directory %s does not exist" % (path, d))
