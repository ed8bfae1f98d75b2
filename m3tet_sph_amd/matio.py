"""MATLAB Level-5 MAT-file reader and writer (the format of MATLAB's ``save``,
v6 uncompressed and v7 zlib-compressed), for the snapshot and restart files
of the time loop:

* ``write_output_file``  <- mfiles_M3TET/write_output_file.m:27-31
  (``<outdir>/<prefix>_Data_<iplot, 4 digits>.mat`` with istep, time, dt, VAR)
* ``write_restart_file`` <- mfiles_M3TET/write_restart_file.m:28-31
  (``<outdir>/<prefix>_RestartFile<irestart>.mat`` with VAR, dt, time, istep,
  iplot)
* ``load_restart_file``  <- the reading side of load_restart_data.m:118-150

so that GPU states can be compared with, or fed to, the reference's MATLAB
pre- and post-processing.

Layout written (Level 5): a 128-byte header (116 bytes of text, 8 bytes of
subsystem offset, version 0x0100, endian mark 'IM'), then one data element
per variable: an 8-byte tag (type, byte count; elements of at most 4 bytes
use the packed 4-byte form) and the data padded to 8 bytes.  A variable is a
miMATRIX element holding array flags (class, logical bit), dimensions, name
and the real part in column-major order; structs add the field-name length
and names and hold one miMATRIX per field; cells hold one miMATRIX per
element.  With ``compress`` every variable is wrapped in a miCOMPRESSED
element (zlib of the whole miMATRIX element), as ``save`` does by default.

Values map as: numpy numeric / bool arrays -> numeric / logical arrays (a
1-D array becomes a column, as the reference's VAR fields are), Python
scalars -> 1x1 double, str -> char row, dict -> 1x1 struct, list -> cell
row.  Reading returns numpy arrays with their MATLAB shape, str for char
rows, dict for 1x1 structs (a list of dicts for struct arrays) and numpy
object arrays for cells.
"""
from __future__ import annotations

import os
import struct
import time
import zlib

import numpy as np

# data types (miXXX)
MI_INT8, MI_UINT8, MI_INT16, MI_UINT16, MI_INT32, MI_UINT32 = 1, 2, 3, 4, 5, 6
MI_SINGLE, MI_DOUBLE, MI_INT64, MI_UINT64 = 7, 9, 12, 13
MI_MATRIX, MI_COMPRESSED, MI_UTF8, MI_UTF16, MI_UTF32 = 14, 15, 16, 17, 18
MI_DTYPE = {MI_INT8: "<i1", MI_UINT8: "<u1", MI_INT16: "<i2", MI_UINT16: "<u2", MI_INT32: "<i4",
            MI_UINT32: "<u4", MI_SINGLE: "<f4", MI_DOUBLE: "<f8", MI_INT64: "<i8", MI_UINT64: "<u8"}
# array classes (mxXXX_CLASS)
MX_CELL, MX_STRUCT, MX_OBJECT, MX_CHAR, MX_SPARSE = 1, 2, 3, 4, 5
MX_NUMERIC = {6: "<f8", 7: "<f4", 8: "<i1", 9: "<u1", 10: "<i2", 11: "<u2", 12: "<i4", 13: "<u4",
              14: "<i8", 15: "<u8"}
CLASS_OF = {np.dtype("f8"): (6, MI_DOUBLE), np.dtype("f4"): (7, MI_SINGLE), np.dtype("i1"): (8, MI_INT8),
            np.dtype("u1"): (9, MI_UINT8), np.dtype("i2"): (10, MI_INT16), np.dtype("u2"): (11, MI_UINT16),
            np.dtype("i4"): (12, MI_INT32), np.dtype("u4"): (13, MI_UINT32), np.dtype("i8"): (14, MI_INT64),
            np.dtype("u8"): (15, MI_UINT64)}
F_LOGICAL, F_GLOBAL, F_COMPLEX = 0x0200, 0x0400, 0x0800
NAMELEN = 64


class MatError(ValueError):
    pass


# --------------------------------------------------------------------- writer

def _element(mtype, payload):
    n = len(payload)
    if n <= 4:
        return struct.pack("<HH", mtype, n) + payload + b"\0" * (4 - n)
    return struct.pack("<II", mtype, n) + payload + b"\0" * ((8 - n % 8) % 8)


def _matrix(name, value):
    """Bytes of one miMATRIX element (tag included)."""
    body = _matrix_body(name, value)
    return struct.pack("<II", MI_MATRIX, len(body)) + body


def _flags(cls, flag_bits=0):
    return _element(MI_UINT32, struct.pack("<II", cls | flag_bits, 0))


def _dims(shape):
    shape = tuple(int(s) for s in shape)
    if len(shape) < 2:
        shape = shape + (1,) * (2 - len(shape))
    return _element(MI_INT32, struct.pack("<%di" % len(shape), *shape))


def _name(name):
    return _element(MI_INT8, name.encode("ascii"))


def _matrix_body(name, v):
    if isinstance(v, dict):
        keys = list(v.keys())
        for k in keys:
            if not k or len(k) >= NAMELEN or not k.isidentifier():
                raise MatError(f"invalid struct field name {k!r}")
        names = b"".join(k.encode("ascii").ljust(NAMELEN, b"\0") for k in keys)
        out = _flags(MX_STRUCT) + _dims((1, 1)) + _name(name)
        out += _element(MI_INT32, struct.pack("<i", NAMELEN)) + _element(MI_INT8, names)
        return out + b"".join(_matrix("", v[k]) for k in keys)
    if isinstance(v, (list, tuple)):
        out = _flags(MX_CELL) + _dims((1, len(v))) + _name(name)
        return out + b"".join(_matrix("", x) for x in v)
    if isinstance(v, str):
        codes = np.array([ord(c) for c in v], dtype="<u2")
        if codes.size and codes.max() > 0xFFFF:
            raise MatError("characters outside the basic plane")
        return _flags(MX_CHAR) + _dims((1, len(v))) + _name(name) + _element(MI_UINT16, codes.tobytes())
    a = np.asarray(v)
    if a.dtype == object:
        out = _flags(MX_CELL) + _dims(a.shape if a.ndim >= 2 else (a.size, 1)) + _name(name)
        return out + b"".join(_matrix("", x) for x in a.ravel(order="F"))
    if np.iscomplexobj(a):
        raise MatError("complex arrays are not written")
    if a.ndim == 0:
        a = a.reshape(1, 1)
    elif a.ndim == 1:
        a = a.reshape(-1, 1)
    if a.dtype == np.bool_:
        data = a.astype("<u1")
        return _flags(9, F_LOGICAL) + _dims(a.shape) + _name(name) + _element(MI_UINT8,
                                                                             data.tobytes(order="F"))
    if isinstance(v, (int, float)) or a.dtype.kind in "iu" and a.dtype.itemsize > 8:
        a = a.astype("<f8")
    a = a.astype(a.dtype.newbyteorder("<"), copy=False)
    key = np.dtype(a.dtype.str.lstrip("<>="))
    if key not in CLASS_OF:
        raise MatError(f"unsupported dtype {a.dtype}")
    cls, mtype = CLASS_OF[key]
    return _flags(cls) + _dims(a.shape) + _name(name) + _element(mtype, a.tobytes(order="F"))


def savemat(path, variables, compress=True):
    """Write ``variables`` (name -> value) as a Level-5 MAT-file."""
    text = ("MATLAB 5.0 MAT-file, Platform: GLNXA64, Created on: "
            + time.strftime("%a %b %d %H:%M:%S %Y")).encode("ascii")
    header = text.ljust(116, b" ") + b"\0" * 8 + struct.pack("<H", 0x0100) + b"IM"
    with open(path, "wb") as f:
        f.write(header)
        for name, value in variables.items():
            if not name.isidentifier() or len(name) >= NAMELEN:
                raise MatError(f"invalid variable name {name!r}")
            el = _matrix(name, value if not isinstance(value, (int, float)) else float(value))
            if compress:
                z = zlib.compress(el)
                f.write(struct.pack("<II", MI_COMPRESSED, len(z)) + z)
            else:
                f.write(el)


# --------------------------------------------------------------------- reader

class _Stream:
    def __init__(self, buf, pos=0, end=None):
        self.b, self.p, self.end = buf, pos, len(buf) if end is None else end

    def more(self):
        return self.p + 8 <= self.end

    def element(self):
        """(type, payload bytes) of the next data element, padding skipped."""
        w0, w1 = struct.unpack_from("<II", self.b, self.p)
        if w0 >> 16:                      # packed small element
            mtype, n = w0 & 0xFFFF, w0 >> 16
            data = self.b[self.p + 4:self.p + 4 + n]
            self.p += 8
            return mtype, data
        mtype, n = w0, w1
        start = self.p + 8
        if start + n > self.end:
            raise MatError("truncated data element")
        data = self.b[start:start + n]
        self.p = start + n + ((8 - n % 8) % 8 if mtype != MI_COMPRESSED else 0)
        return mtype, data


def _numbers(mtype, data):
    if mtype not in MI_DTYPE:
        raise MatError(f"unexpected data type {mtype}")
    return np.frombuffer(data, dtype=MI_DTYPE[mtype])


def _text(mtype, data):
    if mtype in (MI_UTF8, MI_INT8, MI_UINT8):
        return data.decode("utf-8")
    if mtype in (MI_UTF16, MI_UINT16):
        return data.decode("utf-16-le")
    if mtype == MI_UTF32:
        return data.decode("utf-32-le")
    raise MatError(f"unexpected character type {mtype}")


def _parse_matrix(data):
    """(name, value) of a miMATRIX payload."""
    if len(data) == 0:
        return "", np.zeros((0, 0))
    s = _Stream(data)
    t, fl = s.element()
    flags = _numbers(t, fl).astype(np.uint32)
    cls, bits = int(flags[0] & 0xFF), int(flags[0])
    t, dm = s.element()
    dims = tuple(int(x) for x in _numbers(t, dm))
    t, nm = s.element()
    name = nm.decode("ascii") if len(nm) else ""
    n = int(np.prod(dims)) if dims else 0
    if cls in MX_NUMERIC:
        t, re = s.element()
        vals = _numbers(t, re).astype(MX_NUMERIC[cls])
        if bits & F_COMPLEX:
            t, im = s.element()
            vals = vals + 1j * _numbers(t, im).astype(MX_NUMERIC[cls])
        if vals.size != n:
            raise MatError(f"array {name!r}: {vals.size} values for dims {dims}")
        out = vals.reshape(dims, order="F")
        return name, out.astype(bool) if bits & F_LOGICAL else out
    if cls == MX_CHAR:
        t, ch = s.element()
        txt = _text(t, ch)
        if len(dims) == 2 and dims[0] == 1:
            return name, txt
        return name, np.array(list(txt), dtype="<U1").reshape(dims, order="F")
    if cls == MX_CELL:
        out = np.empty(n, dtype=object)
        for i in range(n):
            t, el = s.element()
            out[i] = _parse_matrix(el)[1]
        return name, out.reshape(dims, order="F")
    if cls == MX_STRUCT:
        t, fl_len = s.element()
        flen = int(_numbers(t, fl_len)[0])
        t, fn = s.element()
        fields = [fn[i:i + flen].split(b"\0", 1)[0].decode("ascii") for i in range(0, len(fn), flen)]
        recs = []
        for _ in range(n):
            rec = {}
            for f in fields:
                t, el = s.element()
                rec[f] = _parse_matrix(el)[1]
            recs.append(rec)
        if n == 1:
            return name, recs[0]
        return name, recs
    if cls == MX_SPARSE:
        raise MatError(f"sparse array {name!r} is not read")
    raise MatError(f"array class {cls} of {name!r} is not read")


def loadmat(path):
    """Variables of a Level-5 MAT-file as a dict (name -> value)."""
    with open(path, "rb") as f:
        buf = f.read()
    if len(buf) < 128:
        raise MatError("not a MAT-file (short header)")
    if buf[126:128] != b"IM":
        raise MatError("only little-endian Level-5 MAT-files are read")
    if struct.unpack_from("<H", buf, 124)[0] != 0x0100:
        raise MatError("not a Level-5 MAT-file")
    s = _Stream(buf, 128)
    out = {}
    while s.more():
        t, data = s.element()
        if t == MI_COMPRESSED:
            inner = zlib.decompress(data)
            t, data = _Stream(inner).element()
        if t != MI_MATRIX:
            raise MatError(f"unexpected top-level element type {t}")
        name, value = _parse_matrix(data)
        out[name] = value
    return out


# ------------------------------------------------------ time-loop file names

def num2str_d(i, n):
    """Zero-padded decimal (num2str_d.m)."""
    return str(int(i)).zfill(n)


def write_output_file(SETTINGS, VAR, COMM, time_, istep, dt, iplot):
    """save([outdir '/' prefix '_Data_' num2str_d(iplot,4) '.mat'],'istep','time','dt','VAR')."""
    path = os.path.join(SETTINGS["outdir"], f"{COMM['prefix']}_Data_{num2str_d(iplot, 4)}.mat")
    savemat(path, {"istep": float(istep), "time": float(time_), "dt": float(dt), "VAR": dict(VAR)})
    return path


def write_restart_file(SETTINGS, VAR, COMM, istep, iplot, time_, dt, irestart):
    """save([outdir '/' prefix '_RestartFile' num2str(irestart)],'VAR','dt','time','istep','iplot')."""
    path = os.path.join(SETTINGS["outdir"], f"{COMM['prefix']}_RestartFile{int(irestart)}.mat")
    savemat(path, {"VAR": dict(VAR), "dt": float(dt), "time": float(time_), "istep": float(istep),
                   "iplot": float(iplot)})
    return path


def load_restart_file(path):
    """(VAR, time, istep, iplot) from a restart or output file; iplot is None
    for output files, which do not hold it."""
    d = loadmat(path)
    if "VAR" not in d:
        raise MatError(f"{path}: no VAR")
    VAR = {k: (np.asarray(v).ravel(order="F") if isinstance(v, np.ndarray) else v) for k, v in d["VAR"].items()}
    scal = lambda k: float(np.asarray(d[k]).ravel()[0]) if k in d else None
    ip = scal("iplot")
    return VAR, scal("time"), int(scal("istep")), None if ip is None else int(ip)
