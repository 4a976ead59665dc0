"""STL / PLY meshes of scene files (clovers/src/objects/stl.rs, ply.rs) on the
CPU: file readers by round trips through files written here, the corner
transform by closed forms and an independent rotation library, and the
reference's own mesh files and scenes where the reference tree is present."""
import os
import struct

import numpy as np
import pytest

import clovers_amd
from clovers_amd import abi
from clovers_amd.meshfile import euler_rotation, read_ply, read_stl, transform
from clovers_amd.scenefile import load_scene

REF = "/root/reference"
has_ref = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "scenes")), reason="reference tree not present")


def random_tris(n, seed=1):
    return np.random.default_rng(seed).normal(size=(n, 3, 3)).astype(np.float32)


def write_binary_stl(path, tris):
    with open(path, "wb") as f:
        f.write(b"solid but binary".ljust(80, b" "))     # binary files may start with "solid" too
        f.write(struct.pack("<I", len(tris)))
        for t in tris:
            f.write(struct.pack("<3f", 0, 0, 0) + t.astype("<f4").tobytes() + b"\0\0")


def test_stl_binary_and_ascii_round_trip(tmp_path):
    t = random_tris(37)
    write_binary_stl(tmp_path / "a.stl", t)
    assert np.array_equal(read_stl(tmp_path / "a.stl"), t)
    lines = ["solid x"]
    for tri in t:
        lines += ["facet normal 0 0 1", " outer loop"] + ["  vertex %.9g %.9g %.9g" % tuple(v) for v in tri] + \
                 [" endloop", "endfacet"]
    (tmp_path / "b.stl").write_text("\n".join(lines + ["endsolid x"]) + "\n")
    assert np.array_equal(read_stl(tmp_path / "b.stl"), t)


def write_ply(path, verts, faces, fmt, vtype="float", ltype=("uchar", "int")):
    np_t = {"float": "f4", "double": "f8", "int": "i4", "uchar": "u1", "ushort": "u2", "uint": "u4"}
    e = "<" if fmt == "binary_little_endian" else ">"
    head = ["ply", "format %s 1.0" % fmt, "comment test", "element vertex %d" % len(verts)]
    head += ["property %s %s" % (vtype, c) for c in "xyz"] + ["property float confidence"]
    head += ["element face %d" % len(faces), "property list %s %s vertex_indices" % ltype, "end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        if fmt == "ascii":
            for v in verts:
                f.write(("%r %r %r 1\n" % tuple(float(x) for x in v)).encode())
            for fc in faces:
                f.write((" ".join(str(x) for x in [len(fc)] + list(fc)) + "\n").encode())
            return
        for v in verts:
            f.write(np.asarray(v, e + np_t[vtype]).tobytes() + np.float32(1).astype(e + "f4").tobytes())
        for fc in faces:
            f.write(np.array([len(fc)], e + np_t[ltype[0]]).tobytes() + np.asarray(fc, e + np_t[ltype[1]]).tobytes())


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
def test_ply_round_trip(tmp_path, fmt):
    rng = np.random.default_rng(3)
    verts = rng.normal(size=(20, 3)).astype(np.float32)
    faces = rng.integers(0, 20, size=(30, 3))
    write_ply(tmp_path / "m.ply", verts, faces, fmt)
    assert np.array_equal(read_ply(tmp_path / "m.ply"), verts[faces])


def test_ply_index_rules(tmp_path):
    """signed lists use |i| (ply.rs:150-151); a polygon contributes its first
    three corners only (ply.rs:76-81); integer coordinates are converted"""
    verts = np.arange(30, dtype=np.int32).reshape(10, 3)
    faces = [[0, -1, 2], [3, 4, 5, 6], [-9, 8, 7]]
    path = tmp_path / "p.ply"
    with open(path, "wb") as f:
        f.write(b"ply\nformat binary_little_endian 1.0\nelement vertex 10\nproperty int x\nproperty int y\n"
                b"property int z\nelement face 3\nproperty list uchar int vertex_indices\nend_header\n")
        f.write(verts.astype("<i4").tobytes())
        for fc in faces:
            f.write(bytes([len(fc)]) + np.asarray(fc, "<i4").tobytes())
    want = verts.astype(np.float32)[[[0, 1, 2], [3, 4, 5], [9, 8, 7]]]
    assert np.array_equal(read_ply(path), want)


def test_ply_unsupported_property_types(tmp_path):
    write_ply(tmp_path / "u.ply", np.zeros((3, 3)), [[0, 1, 2]], "binary_little_endian", vtype="uchar")
    with pytest.raises(NotImplementedError):   # ply.rs:118-127: char/short coordinates are unimplemented!()
        read_ply(tmp_path / "u.ply")


def test_euler_rotation_against_an_independent_library():
    from scipy.spatial.transform import Rotation
    for ang in ([0, 0, 0], [90, 0, 0], [0, 90, 0], [0, 0, 90], [-90, 0, 0], [0, 200, 0], [0, 0, -30], [12, -34, 56]):
        R = euler_rotation(ang)
        # from_euler_angles(roll, pitch, yaw) = Rz(yaw) Ry(pitch) Rx(roll) = extrinsic x, then y, then z
        want = Rotation.from_euler("xyz", ang, degrees=True).as_matrix()
        assert np.allclose(R, want, atol=2e-7), ang
    # roll = +90 degrees turns +y into +z (right-handed)
    assert np.allclose(euler_rotation([90, 0, 0]) @ [0, 1, 0], [0, 0, 1], atol=1e-7)


def test_transform_scale_and_offset_exact():
    t = random_tris(5)
    out = transform(t, [0, 0, 0], 2.5, [1, -2, 3])
    assert np.array_equal(out, t * np.float32(2.5) + np.array([1, -2, 3], np.float32))


def test_mesh_objects_in_scene_files(tmp_path):
    t = random_tris(6)
    write_binary_stl(tmp_path / "m.stl", t)
    sf = {"camera": {"look_from": [0, 0, 5], "look_at": [0, 0, 0], "up": [0, 1, 0], "vertical_fov": 40,
                     "aperture": 0, "focus_distance": 5},
          "materials": [{"name": "g", "kind": "Dielectric"}],
          "objects": [{"kind": "STL", "path": "m.stl", "scale": 2, "center": [0, 1, 0], "rotation": [0, 0, 30],
                       "material": "g", "priority": True}]}
    s = load_scene(sf, mesh_root=str(tmp_path))
    o = s.objects
    assert len(o) == 6 and np.all(o["kind"] == abi.OBJ_TRIANGLE) and np.all(o["priority"] == 1)
    assert np.array_equal(o["p"].reshape(-1, 3, 3), transform(t, [0, 0, 30], 2, [0, 1, 0]))
    assert s.materials[o["material"][0]]["kind"] == abi.MAT_DIELECTRIC
    clovers_amd.validate(s)
    # a missing shared material is a panic in the reference (no fallback for meshes)
    sf["objects"][0]["material"] = "nope"
    with pytest.raises(ValueError):
        load_scene(sf, mesh_root=str(tmp_path))
    # a mesh inside a list is "recursing too hard" (scenefile.rs:86-88)
    sf["objects"] = [{"kind": "ObjectList", "objects": [{"kind": "STL", "path": "m.stl", "scale": 1,
                                                         "center": [0, 0, 0], "rotation": [0, 0, 0]}]}]
    with pytest.raises(ValueError):
        load_scene(sf, mesh_root=str(tmp_path))


@has_ref
def test_reference_mesh_files():
    prism = read_stl(os.path.join(REF, "stl/prism.stl"))
    teapot = read_stl(os.path.join(REF, "stl/teapot.stl"))
    bunny = read_ply(os.path.join(REF, "ply/bunny.binary.ply"))
    assert prism.shape == (8, 3, 3) and teapot.shape == (9438, 3, 3) and bunny.shape == (69451, 3, 3)
    for m in (prism, teapot, bunny):
        assert np.all(np.isfinite(m))
    # the header's vertex count bounds the indices the faces use
    assert len(np.unique(bunny.reshape(-1, 3), axis=0)) <= 35947


@has_ref
@pytest.mark.parametrize("name", ["teapot", "bunny", "dispersive", "boing", "one_weekend"])
def test_reference_scenes_load(name):
    s = load_scene(os.path.join(REF, "scenes", name + ".json"), mesh_root=REF)
    info = clovers_amd.validate(s)
    assert info["n_prims"] >= len(s.objects) and info["n_lights"] >= 1
