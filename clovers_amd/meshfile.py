"""STL and PLY triangle meshes for scene files.

The reference turns an ``STL`` / ``PLY`` scene object into a list of
``Triangle::from_coordinates`` hitables (clovers/src/objects/stl.rs:38-96,
clovers/src/objects/ply.rs:37-103), each corner transformed as

    p' = (R * p) * scale + center,   R = Rotation3::from_euler_angles(
                                              rx.to_radians(), ry.to_radians(), rz.to_radians())

and the scene assembly flattens the list into one top-level object per
triangle (clovers-cli/src/scenefile.rs:81-95).  The file formats come from
third-party crates that are not part of the reference tree: stl_io 0.10.0
(ASCII or binary STL; faces in file order, corners in file order) and ply-rs
0.1.3 (ASCII and binary PLY).  The rotation is nalgebra 0.34.1's
``Rotation3::from_euler_angles(roll, pitch, yaw)`` = Rz(yaw) Ry(pitch) Rx(roll),
whose entries are the published ZYX products of the angles' sines and
cosines.  Everything is evaluated in float32 in the reference's operation
order, with sinf/cosf taken from the platform C library, which is what
Rust's f32::sin_cos calls.
"""
import ctypes
import ctypes.util
import os
import struct

import numpy as np

_f32 = np.float32
_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.sinf.restype = ctypes.c_float
_libm.sinf.argtypes = [ctypes.c_float]
_libm.cosf.restype = ctypes.c_float
_libm.cosf.argtypes = [ctypes.c_float]
_DEG = _f32(np.pi) / _f32(180.0)   # f32::to_radians multiplies by PI / 180 in f32


def _sin_cos(a):
    return _f32(_libm.sinf(float(a))), _f32(_libm.cosf(float(a)))


def euler_rotation(degrees):
    """Rotation3::from_euler_angles of the three angles (degrees) as a float32
    3x3 matrix, row-major; entry formulas and product order as nalgebra's."""
    roll, pitch, yaw = (_f32(d) * _DEG for d in degrees)
    sr, cr = _sin_cos(roll)
    sp, cp = _sin_cos(pitch)
    sy, cy = _sin_cos(yaw)
    return np.array([
        [cy * cp, cy * sp * sr - sy * cr, cy * sp * cr + sy * sr],
        [sy * cp, sy * sp * sr + cy * cr, sy * sp * cr - cy * sr],
        [-sp, cp * sr, cp * cr],
    ], dtype=_f32)


def transform(corners, rotation_deg, scale, center):
    """corners [n, 3, 3] float32 -> (R p) * scale + center in float32, summing
    each row of R p left to right (nalgebra's matrix-vector product)."""
    R = euler_rotation(rotation_deg)
    p = np.asarray(corners, _f32)
    with np.errstate(over="ignore", invalid="ignore"):
        rp = np.stack([(R[r, 0] * p[..., 0] + R[r, 1] * p[..., 1]) + R[r, 2] * p[..., 2] for r in range(3)], -1)
        out = rp * _f32(scale) + np.asarray(center, _f32)
    return out.astype(_f32)


# ------------------------------------------------------------------- STL
def read_stl(path):
    """Triangle corners [n, 3, 3] float32 of an ASCII or binary STL file."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) >= 84:
        n = struct.unpack_from("<I", data, 80)[0]
        if 84 + 50 * n == len(data):
            rec = np.frombuffer(data, dtype=np.dtype([("normal", "<f4", 3), ("v", "<f4", (3, 3)), ("attr", "<u2")]),
                                count=n, offset=84)
            return np.array(rec["v"], dtype=_f32)
    text = data.decode("ascii", errors="strict")
    if not text.lstrip().lower().startswith("solid"):
        raise ValueError("%s: not an STL file" % path)
    verts = []
    tris = []
    for line in text.splitlines():
        t = line.split()
        if not t:
            continue
        key = t[0].lower()
        if key == "vertex":
            verts.append([_f32(float(v)) for v in t[1:4]])
        elif key == "endloop":
            if len(verts) != 3:
                raise ValueError("%s: a facet with %d vertices" % (path, len(verts)))
            tris.append(verts)
            verts = []
    return np.array(tris, dtype=_f32).reshape(-1, 3, 3)


# ------------------------------------------------------------------- PLY
_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2",
    "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}


def _ply_header(f):
    if f.readline().strip() != b"ply":
        raise ValueError("not a PLY file")
    fmt = None
    elements = []
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header without end_header")
        t = line.decode("ascii").split()
        if not t or t[0] in ("comment", "obj_info"):
            continue
        if t[0] == "format":
            fmt = t[1]
        elif t[0] == "element":
            elements.append((t[1], int(t[2]), []))
        elif t[0] == "property":
            if t[1] == "list":
                elements[-1][2].append((t[4], "list", _PLY_TYPES[t[2]], _PLY_TYPES[t[3]]))
            else:
                elements[-1][2].append((t[2], "scalar", _PLY_TYPES[t[1]], None))
        elif t[0] == "end_header":
            return fmt, elements


def _read_binary_element(buf, pos, count, props, endian):
    """values of one element as {name: array or list of arrays}, and the new offset"""
    if all(kind == "scalar" for _, kind, _, _ in props):
        dt = np.dtype([(name, endian + t) for name, _, t, _ in props])
        rec = np.frombuffer(buf, dtype=dt, count=count, offset=pos)
        return {name: rec[name] for name, _, _, _ in props}, pos + dt.itemsize * count
    # lists: try the common case of one list whose rows all have the same length
    if len(props) == 1 and count > 0:
        name, _, ct, it = props[0]
        cdt = np.dtype(endian + ct)
        k = int(np.frombuffer(buf, dtype=cdt, count=1, offset=pos)[0])
        dt = np.dtype([("n", endian + ct), ("v", endian + it, (k,))])
        if pos + dt.itemsize * count <= len(buf):
            rec = np.frombuffer(buf, dtype=dt, count=count, offset=pos)
            if np.all(rec["n"] == k):
                return {name: rec["v"]}, pos + dt.itemsize * count
    out = {name: [] for name, _, _, _ in props}
    for _ in range(count):
        for name, kind, t, it in props:
            if kind == "scalar":
                dt = np.dtype(endian + t)
                out[name].append(np.frombuffer(buf, dtype=dt, count=1, offset=pos)[0])
                pos += dt.itemsize
            else:
                cdt, idt = np.dtype(endian + t), np.dtype(endian + it)
                n = int(np.frombuffer(buf, dtype=cdt, count=1, offset=pos)[0])
                pos += cdt.itemsize
                out[name].append(np.frombuffer(buf, dtype=idt, count=n, offset=pos))
                pos += idt.itemsize * n
    return out, pos


def read_ply(path):
    """Triangle corners [n, 3, 3] float32 of a PLY file: the vertex element's
    x, y, z (integer or float properties, ply.rs:110-129) and the first three
    entries of every face's vertex_indices (ply.rs:76-81; signed lists use
    the absolute value, ply.rs:131-153)."""
    with open(path, "rb") as f:
        fmt, elements = _ply_header(f)
        body = f.read()
    data = {}
    if fmt == "ascii":
        rows = iter(body.decode("ascii").split("\n"))
        for name, count, props in elements:
            vals = {p[0]: [] for p in props}
            for _ in range(count):
                t = next(rows).split()
                while not t:
                    t = next(rows).split()
                k = 0
                for pname, kind, typ, it in props:
                    if kind == "scalar":
                        vals[pname].append(np.array(t[k]).astype(typ))
                        k += 1
                    else:
                        n = int(t[k])
                        vals[pname].append(np.array(t[k + 1:k + 1 + n]).astype(it))
                        k += 1 + n
            data[name] = (props, vals)
    elif fmt in ("binary_little_endian", "binary_big_endian"):
        endian = "<" if fmt == "binary_little_endian" else ">"
        pos = 0
        for name, count, props in elements:
            vals, pos = _read_binary_element(body, pos, count, props, endian)
            data[name] = (props, vals)
    else:
        raise ValueError("%s: unknown PLY format %r" % (path, fmt))
    vprops, v = data["vertex"]
    kinds = {p[0]: p for p in vprops}
    for c in "xyz":
        if kinds[c][1] != "scalar" or kinds[c][2] not in ("i4", "u4", "f4", "f8"):
            raise NotImplementedError("PLY: unsupported vertex property format %r" % (kinds[c],))  # ply.rs:118-127
    verts = np.stack([np.asarray(v[c]).astype(_f32) for c in "xyz"], -1)
    fprops, fv = data["face"]
    fk = {p[0]: p for p in fprops}["vertex_indices"]
    if fk[1] != "list" or fk[3] in ("i1", "f4", "f8"):
        raise NotImplementedError("PLY: unsupported face index format %r" % (fk,))
    idx = fv["vertex_indices"]
    if isinstance(idx, np.ndarray):
        first3 = idx[:, :3]
    else:
        first3 = np.array([np.asarray(r)[:3] for r in idx])
    first3 = np.abs(first3.astype(np.int64))
    return verts[first3]


def load_mesh(path):
    ext = os.path.splitext(path)[1].lower()
    if ext == ".stl":
        return read_stl(path)
    if ext == ".ply":
        return read_ply(path)
    raise ValueError("unknown mesh file type %r" % path)


def mesh_triangles(init, root=None):
    """STLInit / PLYInit (stl.rs:26-36, ply.rs:25-35) -> world corners [n, 3, 3]
    float32; `path` is relative to `root` (the reference reads it relative to
    the working directory)."""
    path = init["path"]
    if root is not None and not os.path.isabs(path):
        path = os.path.join(root, path)
    corners = load_mesh(path)
    return transform(corners, init["rotation"], init["scale"], init["center"])
