"""The library's own scene-file reader (clv_scene_file_parse, C++) against
the Python restatement (clovers_amd.scenefile / meshfile), which the other
scene-file tests pin: the descriptors must agree field for field -- geometry
and mesh corners bit for bit, colours to float32 rounding -- and errors must
carry the same status.  No device needed."""
import json
import os

import numpy as np
import pytest

import clovers_amd
from clovers_amd import abi, synth
from clovers_amd.scenefile import load_scene
from test_meshfile import random_tris, write_binary_stl, write_ply

REF = "/root/reference"
has_ref = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "scenes")), reason="reference tree not present")


def same(py, nat):
    o1, o2 = py.objects, nat.objects
    assert len(o1) == len(o2)
    for f in ("kind", "material", "priority", "xform", "p"):
        assert np.array_equal(o1[f], o2[f]), f
    m1 = np.array(py.materials, dtype=abi.MATERIAL_DTYPE).reshape(-1)
    m2 = nat.materials
    for f in abi.MATERIAL_DTYPE.names:
        if m1[f].dtype.kind == "f":
            assert np.allclose(m1[f], m2[f], rtol=3e-7, atol=0), f
        else:
            assert np.array_equal(m1[f], m2[f]), f
    x1 = np.array(py.xforms, dtype=abi.XFORM_DTYPE).reshape(-1) if py.xforms else np.zeros(0, abi.XFORM_DTYPE)
    assert x1.tobytes() == nat.xforms.tobytes()
    assert np.allclose(py.background, nat.background, rtol=3e-7, atol=0)
    for k, v in py.camera.items():
        assert np.allclose(v, nat.camera[k]), k


@pytest.mark.parametrize("name", ["cornell", "spheres", "moving", "prism", "features", "media"])
def test_synthetic_scenes_agree(name):
    js = {"cornell": synth.cornell_json, "spheres": synth.spheres_json, "moving": lambda: synth.spheres_json(3, True),
          "prism": synth.prism_json, "features": synth.features_json, "media": synth.media_json}[name]
    js = js()
    nat = clovers_amd.load_scene_file(json.dumps(js))
    same(load_scene(js), nat)
    clovers_amd.validate(nat)


def test_meshes_and_colours_agree(tmp_path):
    t = random_tris(9)
    write_binary_stl(tmp_path / "b.stl", t)
    lines = ["solid x"]
    for tri in t:
        lines += ["facet normal 0 0 1", "outer loop"] + ["vertex %.9g %.9g %.9g" % tuple(v) for v in tri] + \
                 ["endloop", "endfacet"]
    (tmp_path / "a.stl").write_text("\n".join(lines + ["endsolid"]) + "\n")
    rng = np.random.default_rng(5)
    verts = rng.normal(size=(12, 3)).astype(np.float32)
    faces = rng.integers(0, 12, size=(15, 3))
    for fmt in ("ascii", "binary_little_endian", "binary_big_endian"):
        write_ply(tmp_path / ("m_%s.ply" % fmt), verts, faces, fmt)
    mesh = lambda kind, path, rot: {"kind": kind, "path": path, "scale": 1.7, "center": [0.5, -1, 2],
                                    "rotation": rot, "material": "g"}
    sf = {"camera": {"look_from": [0, 0, 5], "look_at": [0, 0, 0], "up": [0, 1, 0], "vertical_fov": 40,
                     "aperture": 0.1, "focus_distance": 5},
          "background_color": {"red": 0.2, "green": 0.5, "blue": 0.9},
          "materials": [{"name": "g", "kind": "Metal", "fuzz": 0.1, "albedo": {"kind": "SolidColor", "color": {"hex": "#b87333"}}},
                        {"name": "c", "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": {"lin_srgb": [0.1, 0.5, 0.9]}}},
                        {"name": "d", "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": {"xyz_d65": [0.3, 0.3, 0.2]}}},
                        {"name": "e", "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": {"srgb": [0.3, 0.6, 0.2]}}},
                        {"name": "f", "kind": "DiffuseLight", "emit": {"kind": "IlluminantD50", "intensity": 3}}],
          "objects": [mesh("STL", "a.stl", [10, 20, 30]), mesh("STL", "b.stl", [-90, 0, 0]),
                      mesh("PLY", "m_ascii.ply", [0, 200, 0]), mesh("PLY", "m_binary_little_endian.ply", [0, 0, -30]),
                      mesh("PLY", "m_binary_big_endian.ply", [45, 45, 45]),
                      {"kind": "Sphere", "center": [0, 0, 0], "radius": 1, "material": "c", "priority": True},
                      {"kind": "Quad", "q": [0, 0, 0], "u": [1, 0, 0], "v": [0, 1, 0], "material": "d"},
                      {"kind": "Quad", "q": [0, 0, 0], "u": [1, 0, 0], "v": [0, 1, 0], "material": "e"},
                      {"kind": "Quad", "q": [0, 0, 0], "u": [1, 0, 0], "v": [0, 1, 0], "material": "f"}]}
    nat = clovers_amd.load_scene_file(json.dumps(sf), mesh_root=str(tmp_path))
    same(load_scene(sf, mesh_root=str(tmp_path)), nat)
    assert len(nat.objects) == 9 + 9 + 3 * 15 + 4


@pytest.mark.parametrize("mutate,status", [
    (lambda s: s["objects"].append({"kind": "ConstantMedium", "boundary": {"kind": "Quad", "q": [0, 0, 0],
                                                                             "u": [1, 0, 0], "v": [0, 1, 0]}}),
     abi.EUNSUPPORTED),
    (lambda s: s["objects"].append({"kind": "ConstantMedium", "priority": True,
                                    "boundary": {"kind": "Sphere", "center": [0, 0, 0], "radius": 1}}), abi.EINVAL),
    (lambda s: s["objects"].append({"kind": "Translate", "offset": [0, 0, 0],
                                    "object": {"kind": "ObjectList", "objects": []}}), abi.EUNSUPPORTED),
    (lambda s: s["objects"].append({"kind": "ObjectList", "objects": [{"kind": "ObjectList", "objects": []}]}),
     abi.EINVAL),
    (lambda s: s["materials"].append({"name": "x", "kind": "Isotropic", "albedo": {"kind": "SurfaceChecker"}}),
     abi.EUNSUPPORTED),
    (lambda s: s["materials"].append({"name": "x", "kind": "ConeLight", "spread": 3}), abi.EINVAL),
    (lambda s: s.pop("camera"), abi.EINVAL),
    (lambda s: s["objects"].append({"kind": "STL", "path": "/nonexistent.stl", "scale": 1, "center": [0, 0, 0],
                                    "rotation": [0, 0, 0]}), abi.EINVAL),
])
def test_errors(mutate, status):
    sf = synth.cornell_json()
    mutate(sf)
    with pytest.raises(clovers_amd.ClvError) as e:
        clovers_amd.load_scene_file(json.dumps(sf))
    assert e.value.status == status


def test_malformed_json():
    for bad in ["{", '{"camera": }', '{"a": [1, 2,]}', '{"a": 1} x', '{"a": "\\q"}']:
        with pytest.raises(clovers_amd.ClvError) as e:
            clovers_amd.load_scene_file(bad)
        assert e.value.status == abi.EINVAL, bad


def test_json_strings_and_numbers():
    sf = synth.cornell_json()
    sf["materials"][0]["name"] = "lämp \"1\" \\ \U0001F4A1"
    for o in sf["objects"]:
        if o.get("material") == "light":
            o["material"] = sf["materials"][0]["name"]
    sf["objects"][0]["q"] = [1e-3, -2.5E+2, 0.1]
    same(load_scene(sf), clovers_amd.load_scene_file(json.dumps(sf)))
    same(load_scene(sf), clovers_amd.load_scene_file(json.dumps(sf, ensure_ascii=True)))


@has_ref
@pytest.mark.parametrize("name", ["cornell", "colorchecker", "dispersive", "one_weekend", "boing", "teapot", "bunny",
                                  "spatial_checker_smoke"])
def test_reference_scenes_agree(name):
    path = os.path.join(REF, "scenes", name + ".json")
    nat = clovers_amd.load_scene_file(path, mesh_root=REF)
    same(load_scene(path, mesh_root=REF), nat)
    clovers_amd.validate(nat)
