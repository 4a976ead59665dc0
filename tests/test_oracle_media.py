"""The oracle's restatement of ConstantMedium and Isotropic (CPU), pinned by
what the definitions fix:
  * free flight (constant_medium.rs:65-126): along a ray the medium is hit
    with probability 1 - exp(-density * chord) and at an exponentially
    distributed distance past the boundary's entry (rate = density), truncated
    at the exit;
    through a Sphere, a Boxy seen along its diagonal (RotateY 45 inside the
    boundary) and from a camera inside the boundary (entry clipped to the
    ray's start);
  * a medium of white Isotropic albedo (isotropic.rs: SpherePDF, scattering
    pdf 1/(4 pi)) under a uniform sky conserves radiance: every pixel sees the
    sky (a wrong phase-function weight or mixture pdf breaks the identity);
  * a medium never hit (density 0) leaves the image bit-identical to the
    scene without it (its draws do not come from the path's stream)."""
import numpy as np
import pytest

from clovers_amd import abi
from clovers_amd.scenefile import load_scene
from oracle import py_oracle as po
import special_scenes as S


def probe(scene, W=128, seed=3):
    return po.OracleScene(scene).primary_hits(abi.render_opts(W, W, 1, 1, seed=seed))


@pytest.mark.parametrize("sigma", [0.2, 0.5, 1.5])
def test_sphere_free_flight(sigma):
    ids, ts = probe(S.medium_probe({"kind": "Sphere", "center": [0, 0, 0], "radius": 1.0}, sigma))
    S.free_flight_check(ids, ts, 9.0, 2.0, sigma, "sphere sigma %.1f" % sigma)


def test_rotated_box_free_flight():
    box = {"kind": "RotateY", "angle": 45, "object": {"kind": "Boxy", "corner_0": [-1, -1, -1], "corner_1": [1, 1, 1]}}
    ids, ts = probe(S.medium_probe(box, 0.4))
    S.free_flight_check(ids, ts, 10.0 - np.sqrt(2.0), 2.0 * np.sqrt(2.0), 0.4, "box diagonal")


def test_translated_box_free_flight():
    box = {"kind": "Boxy", "corner_0": [-0.5, -0.5, -0.25], "corner_1": [0.5, 0.5, 0.75]}
    outer = lambda m: {"kind": "Translate", "offset": [0, 0, 2.0], "object": m}
    ids, ts = probe(S.medium_probe(box, 1.0, outer))
    S.free_flight_check(ids, ts, 10.0 - 2.75, 1.0, 1.0, "translated box")


def test_camera_inside_medium():
    """entry behind the camera: clipped to the ray's start (0.001)"""
    ids, ts = probe(S.medium_probe({"kind": "Sphere", "center": [0, 0, 0], "radius": 100.0}, 0.2))
    S.free_flight_check(ids, ts, 0.001, 110.0 - 0.001, 0.2, "inside")


def test_white_medium_furnace():
    s = S.medium_furnace(1.0)
    W = H = 24
    img, sq, m64 = po.OracleScene(s).render(abi.render_opts(W, H, 128, 64, seed=5), want_sq=True)
    g = float(s.background[0])
    want = S.spd_xyz((g, g, g))
    mask = S.sphere_mask(W, H)
    assert mask.sum() > 60
    for label, m in (("fog", mask), ("sky", ~mask)):
        mean = m64[m].mean(0)
        var = (sq[m] - m64[m] ** 2).mean(0)
        se = np.sqrt(np.maximum(var, 0) / (m.sum() * 128)) + 1e-9
        print(label, mean, want, se)
        assert np.all(np.abs(mean - want) < 4 * se + 2e-5 * want), (label, mean, want, se)
    # the fog pixels really scatter (their variance is well above the sky's)
    assert (sq[mask] - m64[mask] ** 2).mean() > 3 * max((sq[~mask] - m64[~mask] ** 2).mean(), 1e-12)


def test_absorbing_medium_darkens_by_transmittance():
    """A black medium (albedo 0) in front of an emitting wall: the wall seen
    through the sphere's centre is scaled by its transmittance exp(-2 sigma)."""
    sigma = 0.5
    wall = {"kind": "Quad", "priority": True, "q": [-5, -5, -3], "u": [10, 0, 0], "v": [0, 10, 0], "material": "lamp"}
    base = {"time_0": 0, "time_1": 1, "camera": S.MEDIUM_CAM, "background_color": [0, 0, 0],
            "materials": [{"name": "lamp", "kind": "DiffuseLight", "emit": {"kind": "SolidColor", "color": {"xyz_e": [1, 1, 1]}}}]}
    fog = {"kind": "ConstantMedium", "density": sigma, "texture": {"kind": "SolidColor", "color": {"xyz_e": [0, 0, 0]}},
           "boundary": {"kind": "Sphere", "center": [0, 0, 0], "radius": 1.0}}
    opts = abi.render_opts(16, 16, 256, 4, seed=9)
    with_fog, sq, m64 = po.OracleScene(load_scene(dict(base, objects=[fog, wall]))).render(opts, want_sq=True)
    without, _, m0 = po.OracleScene(load_scene(dict(base, objects=[wall]))).render(opts, want_sq=True)
    ratio = m64[..., 1].sum() / m0[..., 1].sum()
    t = np.exp(-2 * sigma)
    n = 16 * 16 * 256
    se = np.sqrt(t * (1 - t) / n)
    print("transmittance %.4f want %.4f (se %.4f)" % (ratio, t, se))
    assert abs(ratio - t) < 5 * se


def test_unhit_medium_changes_nothing():
    """density 0: -1/0 = -inf makes every free flight infinite, and the
    medium draws nothing from the path's stream, so the image is bit-identical
    to the scene without it"""
    zero = S.medium_furnace(0.0)
    empty = S.medium_furnace(0.0)
    empty._objects = [o[o["kind"] != abi.OBJ_MEDIUM_SPHERE] for o in empty._objects]
    empty._obj_cache = None
    assert len(empty.objects) == 1
    opts = abi.render_opts(12, 12, 8, 8, seed=2)
    a = po.OracleScene(zero).render(opts)[0]
    b = po.OracleScene(empty).render(opts)[0]
    assert np.array_equal(a, b)
    c = po.OracleScene(S.medium_furnace(1.0)).render(opts)[0]
    assert not np.array_equal(a, c)
