"""The oracle's restatement of MovingSphere, SpatialChecker, SurfaceChecker
and ConeLight, pinned by closed forms (CPU).

No reference test covers these; the checks are what the definitions fix:
  * checker textures: on a flat wall the cells are rectangles, so a pixel
    whose corners fall in one cell sees exactly the even or odd emission
    (spatial_checker.rs:74-88 by world position, surface_checker.rs:72-84 by the
    quad's / triangle's (alpha, beta) and the sphere's (u, v), sphere.rs:58-65);
  * ConeLight: lit exactly where the angle between the ray and the wall's
    normal is within the spread (cone_light.rs:40-53);
  * MovingSphere: an emitter sweeping across the view covers each pixel for
    the fraction of [0, 1) given by the swept disc (moving_sphere.rs:85-90,
    camera time uniform in [0, 1), sampler/random.rs:23); a sphere that does
    not move renders bit-identically to a Sphere."""
import numpy as np
import pytest

from clovers_amd import abi
from oracle import py_oracle as po
import special_scenes as S


def render(scene, W, H, spp, depth=2, seed=3):
    xyz, sq, m64 = po.OracleScene(scene).render(abi.render_opts(W, H, spp, depth, seed=seed), want_sq=True)
    return m64, sq


@pytest.mark.parametrize("kind,density", [("SpatialChecker", 2.0), ("SurfaceChecker", 40.0)])
def test_checker_wall(kind, density):
    W = H = 40
    img, sq = render(S.checker_wall(kind, density), W, H, 32)
    cls = S.checker_classes(W, H, kind, density)
    assert (cls == 0).sum() > 200 and (cls == 1).sum() > 200
    S.check_emission_classes(img, sq, 32, cls, {0: S.spd_xyz(S.EVEN), 1: S.spd_xyz(S.ODD)}, kind)


def test_checker_triangle_uses_alpha_beta():
    W = H = 40
    img, sq = render(S.checker_triangle(40.0), W, H, 32)
    cls = S.triangle_classes(W, H, 40.0)
    S.check_emission_classes(img, sq, 32, cls, {0: S.spd_xyz(S.EVEN), 1: S.spd_xyz(S.ODD), 2: np.zeros(3)}, "triangle")


def test_checker_sphere_uses_sphere_uv():
    W = H = 48
    img, sq = render(S.checker_sphere(8.0), W, H, 32)
    cls = S.sphere_uv_classes(W, H, 0.0, 1.0, 8.0)
    assert (cls == 0).sum() > 100 and (cls == 1).sum() > 100
    S.check_emission_classes(img, sq, 32, cls, {0: S.spd_xyz(S.EVEN), 1: S.spd_xyz(S.ODD)}, "sphere")


@pytest.mark.parametrize("spread", [10.0, 4.0])
def test_cone_light(spread):
    W = H = 40
    img, sq = render(S.cone_wall(spread), W, H, 16)
    cls = S.cone_classes(W, H, spread)
    assert (cls == 1).sum() > 20 and (cls == 0).sum() > 200
    S.check_emission_classes(img, sq, 16, cls, {1: S.spd_xyz(S.EVEN), 0: np.zeros(3)}, "cone")


def test_moving_sphere_motion_blur_closed_form():
    W, H, spp = 48, 24, 128
    img, sq = render(S.moving_lamp(), W, H, spp)
    S.motion_check(img, sq, spp, W, H)


def test_moving_sphere_at_rest_equals_sphere():
    a = S.moving_lamp(static=True)
    b = S.moving_lamp()
    objs = b.objects.copy()
    objs["p"][0][3:6] = objs["p"][0][0:3]        # centre_1 = centre_0: never moves
    b._obj_cache = objs
    o = abi.render_opts(32, 16, 8, 2, seed=9)
    ia, _, _ = po.OracleScene(a).render(o)
    ib, _, _ = po.OracleScene(b).render(o)
    assert np.array_equal(ia, ib)
    assert ia.max() > 0


def test_normal_map_closed_form():
    """normal_to_color (normals.rs:21-33) of the unit sphere seen from +z:
    the oriented normal of the hit point of each pixel's corner ray
    (draw_cpu.rs:148-166), and (0.5, 0.5, 0) for the camera-facing wall."""
    W, H = 40, 30
    sphere = S.checker_sphere(8.0)
    rgb = po.OracleScene(sphere).normal_map(abi.render_opts(W, H, 1, 1, seed=4))
    half = np.tan(np.radians(15.0))
    rows, cols = np.mgrid[0:H, 0:W]
    u, v = cols / W, (H - 1 - rows) / H
    d = np.stack([(2 * u - 1) * half * (W / H), (2 * v - 1) * half, -np.ones_like(u, dtype=np.float64)], -1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    o = np.array(S.CAM["look_from"], np.float64)
    b = d @ o
    disc = b * b - (o @ o - 1.0)
    hit = disc > 1e-3
    t = -b - np.sqrt(np.maximum(disc, 0))
    n = o + t[..., None] * d
    want = np.stack([0.5 - 0.5 * n[..., 0], 0.5 + 0.5 * n[..., 1], np.maximum(-n[..., 2], 0)], -1)
    assert np.allclose(rgb[hit], want[hit], atol=1e-5)
    assert np.all(rgb[disc < -1e-3] == 0)
    wall = po.OracleScene(S.emitter_wall({"kind": "DiffuseLight"})).normal_map(abi.render_opts(W, H, 1, 1))
    assert np.all(wall == np.array([0.5, 0.5, 0.0], np.float32))
