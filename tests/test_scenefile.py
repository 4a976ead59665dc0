"""Scene files -> descriptor (clovers-cli/src/scenefile.rs, clovers/src/objects.rs,
clovers/src/colorinit.rs) and the host scene build (clv_scene_validate, no device).

Colour conversions are pinned by closed forms: palette's sRGB EOTF,
sRGB->XYZ(D65) from primaries, Bradford D65->E.  Their last-ulp agreement with
palette's f32 code is not pinned (no reference test covers it)."""
import numpy as np
import pytest

import clovers_amd
from clovers_amd import abi, synth
from clovers_amd.color import (M_D65_TO_E, M_SRGB_TO_XYZ_D65, WHITE_D65, color_init_to_xyz_e, srgb_eotf,
                               srgb_to_xyz_e)
from clovers_amd.scenefile import load_scene


def test_srgb_eotf_closed_form():
    assert srgb_eotf(0.73) == pytest.approx(0.4919, abs=1e-4)
    assert srgb_eotf(25.0) == pytest.approx(((25.055) / 1.055) ** 2.4, rel=1e-12)  # ~2002, decoded above 1
    assert srgb_eotf(25.0) == pytest.approx(2002, rel=2e-3)
    assert srgb_eotf(0.04) == pytest.approx(0.04 / 12.92)


def test_srgb_matrix_maps_white_to_d65_and_bradford_to_e():
    assert np.allclose(M_SRGB_TO_XYZ_D65 @ np.ones(3), WHITE_D65, atol=1e-12)
    assert np.allclose(M_SRGB_TO_XYZ_D65[1], [0.2126, 0.7152, 0.0722], atol=2e-4)  # BT.709 luma
    assert np.allclose(M_D65_TO_E @ WHITE_D65, np.ones(3), atol=1e-12)
    assert np.allclose(srgb_to_xyz_e([1, 1, 1]), np.ones(3), atol=1e-12)   # neutral -> E neutral
    g = srgb_to_xyz_e([0.5, 0.5, 0.5])
    assert np.allclose(g, srgb_eotf(0.5), atol=1e-12)


def test_typed_colours():
    assert np.allclose(color_init_to_xyz_e({"xyz_e": [0.1, 0.2, 0.3]}), [0.1, 0.2, 0.3])
    assert np.allclose(color_init_to_xyz_e({"lin_srgb": [1, 1, 1]}), [1, 1, 1], atol=1e-6)
    assert np.allclose(color_init_to_xyz_e({"hex": "#ffffff"}), [1, 1, 1], atol=1e-6)
    assert np.allclose(color_init_to_xyz_e([0.2, 0.3, 0.4]), color_init_to_xyz_e({"srgb": [0.2, 0.3, 0.4]}))
    with pytest.raises(NotImplementedError):
        color_init_to_xyz_e({"oklch": [0.5, 0.1, 30]})


def test_cornell_descriptor():
    s = synth.cornell()
    o = s.objects
    assert len(o) == 8
    assert list(o["kind"]) == [abi.OBJ_QUAD] * 6 + [abi.OBJ_BOXY] * 2
    assert list(o["priority"]) == [0, 0, 0, 0, 0, 1, 0, 0]
    assert len(s.xforms) == 2
    x = s.xforms[0]
    assert x["n_ops"] == 2
    assert x["ops"][0]["kind"] == abi.XF_TRANSLATE and list(x["ops"][0]["v"]) == [265, 0, 295]
    assert x["ops"][1]["kind"] == abi.XF_ROTATE_Y and x["ops"][1]["v"][0] == 15
    # shared materials + the default appended last (scenefile.rs:58-59)
    assert len(s.materials) == 5
    assert s.materials[-1]["kind"] == abi.MAT_LAMBERTIAN and np.allclose(s.materials[-1]["color"], 0.5)
    lamp = s.materials[0]
    assert lamp["kind"] == abi.MAT_DIFFUSE_LIGHT and lamp["texture"] == abi.TEX_ILLUMINANT_D65
    assert lamp["intensity"] == np.float32(0.15)
    grey = s.materials[3]
    assert np.allclose(grey["color"], srgb_eotf(0.73), atol=1e-6)   # legacy arrays are encoded sRGB


def test_material_defaults_and_fallback():
    sf = synth.cornell_json()
    sf["objects"].append({"kind": "Sphere", "center": [1, 2, 3], "radius": 1, "material": "no such material"})
    sf["objects"].append({"kind": "Sphere", "center": [1, 2, 3], "radius": 1,
                          "material": {"kind": "Dielectric"}})
    sf["objects"].append({"kind": "Sphere", "center": [1, 2, 3], "radius": 1,
                          "material": {"kind": "Metal", "fuzz": 3.0, "albedo": {"kind": "SolidColor", "color": [1, 1, 1]}}})
    s = load_scene(sf)
    o = s.objects
    assert o["material"][8] == len(sf["materials"])          # unknown name -> default material
    d = s.materials[o["material"][9]]
    assert d["kind"] == abi.MAT_DIELECTRIC and d["refractive_index"] == np.float32(1.5)
    assert np.allclose(d["color"], 1.0)                       # Xyz<E>(1,1,1) default, not sRGB
    m = s.materials[o["material"][10]]
    assert m["fuzz"] == np.float32(3.0)                        # not clamped on the scene-file path


def test_object_list_flattening_and_priority():
    sf = synth.cornell_json()
    sf["objects"] = sf["objects"][:6] + [{
        "kind": "ObjectList", "priority": True,
        "objects": [{"kind": "Sphere", "center": [0, 0, 0], "radius": 1},
                    {"kind": "Triangle", "q": [0, 0, 0], "u": [1, 0, 0], "v": [0, 1, 0]}]}]
    s = load_scene(sf)
    assert len(s.objects) == 8
    assert list(s.objects["priority"][6:]) == [1, 1]
    assert s.objects["kind"][7] == abi.OBJ_TRIANGLE_QUV


def test_unsupported_inputs_raise():
    sf = synth.cornell_json()
    sf["objects"].append({"kind": "ConstantMedium", "density": 0.1,
                          "boundary": {"kind": "Quad", "q": [0, 0, 0], "u": [1, 0, 0], "v": [0, 1, 0]}})
    with pytest.raises(NotImplementedError):
        load_scene(sf)
    sf = synth.cornell_json()
    sf["objects"].append({"kind": "ConstantMedium", "density": 0.1,
                          "texture": {"kind": "SurfaceChecker"},
                          "boundary": {"kind": "Sphere", "center": [0, 0, 0], "radius": 1}})
    with pytest.raises(NotImplementedError):
        load_scene(sf)
    sf = synth.cornell_json()
    sf["objects"].append({"kind": "ConstantMedium", "priority": True,
                          "boundary": {"kind": "Sphere", "center": [0, 0, 0], "radius": 1}})
    with pytest.raises(ValueError):
        load_scene(sf)
    sf = synth.cornell_json()
    sf["objects"].append({"kind": "Translate", "offset": [0, 0, 0],
                          "object": {"kind": "ObjectList", "objects": []}})
    with pytest.raises(NotImplementedError):
        load_scene(sf)


def test_moving_sphere_checkers_and_cone_light():
    """MovingSphere (objects.rs:104-109: the file's times are not read), the
    two checker textures with their serde defaults (spatial_checker.rs:46-61,
    surface_checker.rs:44-59) and ConeLight (cone_light.rs: both fields required)."""
    sf = synth.cornell_json()
    sf["materials"] += [
        {"name": "sc", "kind": "Lambertian", "albedo": {"kind": "SpatialChecker"}},
        {"name": "uc", "kind": "Metal", "albedo": {"kind": "SurfaceChecker", "even": [1, 0, 0], "odd": {"xyz_e": [0.1, 0.2, 0.3]}}},
        {"name": "cone", "kind": "ConeLight", "spread": 12.5, "emit": {"kind": "SurfaceChecker", "density": 3}},
        {"name": "cone65", "kind": "ConeLight", "spread": 30, "emit": {"kind": "IlluminantD65", "intensity": 2}},
    ]
    sf["objects"] += [
        {"kind": "MovingSphere", "center_0": [1, 2, 3], "center_1": [4, 5, 6], "time_0": 0.3, "time_1": 0.7,
         "radius": 0.5, "material": "sc"},
        {"kind": "Quad", "q": [0, 0, 0], "u": [1, 0, 0], "v": [0, 1, 0], "material": "cone"},
    ]
    s = load_scene(sf)
    o = s.objects
    assert o["kind"][-2] == abi.OBJ_MOVING_SPHERE
    assert list(o["p"][-2][:7]) == [1, 2, 3, 4, 5, 6, 0.5]
    mats = {k: s.materials[o["material"][i]] for k, i in (("sc", -2), ("cone", -1))}
    sc = mats["sc"]
    assert sc["texture"] == abi.TEX_SPATIAL_CHECKER and sc["density"] == 1.0
    assert np.allclose(sc["color"], 0.8) and np.allclose(sc["odd_color"], 0.5)
    cone = mats["cone"]
    assert cone["kind"] == abi.MAT_CONE_LIGHT and cone["spread"] == np.float32(12.5)
    assert cone["texture"] == abi.TEX_SURFACE_CHECKER and cone["density"] == 3.0
    by_kind = {(int(m["kind"]), int(m["texture"])): m for m in s.materials}
    uc = by_kind[(abi.MAT_METAL, abi.TEX_SURFACE_CHECKER)]
    assert uc["density"] == 10.0 and np.allclose(uc["odd_color"], [0.1, 0.2, 0.3])
    assert np.allclose(uc["color"], srgb_to_xyz_e([1, 0, 0]), atol=1e-6)
    c65 = by_kind[(abi.MAT_CONE_LIGHT, abi.TEX_ILLUMINANT_D65)]
    assert c65["intensity"] == 2.0 and c65["spread"] == 30.0
    clovers_amd.validate(s)
    with pytest.raises(ValueError):
        load_scene({**sf, "materials": [{"name": "x", "kind": "ConeLight", "spread": 3}]})


def test_moving_sphere_is_not_a_light_sampling_target():
    sf = synth.cornell_json()
    sf["objects"].append({"kind": "MovingSphere", "center_0": [0, 0, 0], "center_1": [0, 1, 0], "radius": 1,
                          "priority": True})
    with pytest.raises(clovers_amd.ClvError) as e:     # MovingSphere::random is unimplemented!() in the reference
        clovers_amd.validate(load_scene(sf))
    assert e.value.status == abi.EINVAL


def test_validate_shapes():
    info = clovers_amd.validate(synth.cornell())
    assert info["n_objects"] == 8 and info["n_prims"] == 18 and info["n_lights"] == 1
    info = clovers_amd.validate(synth.spheres())
    assert info["n_prims"] == 488 and info["n_lights"] == 1
    info = clovers_amd.validate(synth.prism())
    assert info["n_prims"] == 16 and info["n_lights"] == 15


def test_validate_rejects_what_the_reference_panics_on():
    sf = synth.cornell_json()
    for o in sf["objects"]:
        o.pop("priority", None)
    with pytest.raises(clovers_amd.ClvError) as e:
        clovers_amd.validate(load_scene(sf))                  # empty MIS list: sah.rs:23 unwraps None
    assert e.value.status == abi.EINVAL
    sf = synth.cornell_json()
    sf["objects"][6]["priority"] = True                       # Translate{RotateY{Boxy}} has no random()
    with pytest.raises(clovers_amd.ClvError) as e:
        clovers_amd.validate(load_scene(sf))
    assert e.value.status == abi.EUNSUPPORTED


def test_validate_rejects_bad_indices():
    s = synth.cornell()
    objs = s.objects.copy()
    objs["material"][0] = 99
    s._obj_cache = objs
    with pytest.raises(clovers_amd.ClvError) as e:
        clovers_amd.validate(s)
    assert e.value.status == abi.EINVAL


def test_prism_triangles_wind_outwards():
    t = synth.prism_triangles().reshape(-1, 3, 3).astype(np.float64)
    c = t.reshape(-1, 3).mean(0)
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    out = t.mean(1) - c
    assert np.all((n * out).sum(1) > 0)
    assert len(t) == 12


def test_icosphere_is_closed_and_outward():
    V, T = synth.geodesic_icosphere(6)
    assert len(T) == 20 * 36
    # closed 2-manifold: every edge shared by exactly two triangles, in opposite directions
    e = np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]])
    fwd = {tuple(x) for x in e}
    assert len(fwd) == len(e)
    assert all((b, a) in fwd for a, b in fwd)
    P = V[T]
    n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    assert np.all((n * P.mean(1)).sum(1) > 0)
    assert len(V) == 10 * 36 + 2   # Euler: V = 10 f^2 + 2
