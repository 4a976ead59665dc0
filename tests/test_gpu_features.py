"""Checker textures, cone lights and moving spheres on the device: the closed
forms the oracle is pinned by (tests/test_oracle_features.py), bit-exact
invariants, and parity with the oracle on a room that mixes them all."""
import numpy as np
import pytest

from clovers_amd import abi, synth
from oracle import py_oracle as po
import special_scenes as S
from test_gpu_parity import block_z

pytestmark = pytest.mark.gpu


def render(r, scene, W, H, spp, depth=2, seed=5):
    sid = r.upload(scene)
    img, sq = r.render(sid, abi.render_opts(W, H, spp, depth, seed=seed), want_sq=True)
    r.free(sid)
    return img.astype(np.float64), sq


@pytest.mark.parametrize("kind,density", [("SpatialChecker", 2.0), ("SurfaceChecker", 40.0)])
def test_checker_wall(renderer, kind, density):
    W = H = 96
    img, sq = render(renderer, S.checker_wall(kind, density), W, H, 64)
    cls = S.checker_classes(W, H, kind, density)
    S.check_emission_classes(img, sq, 64, cls, {0: S.spd_xyz(S.EVEN), 1: S.spd_xyz(S.ODD)}, kind)


def test_checker_triangle(renderer):
    W = H = 96
    img, sq = render(renderer, S.checker_triangle(40.0), W, H, 64)
    cls = S.triangle_classes(W, H, 40.0)
    S.check_emission_classes(img, sq, 64, cls, {0: S.spd_xyz(S.EVEN), 1: S.spd_xyz(S.ODD), 2: np.zeros(3)}, "triangle")


def test_checker_sphere(renderer):
    W = H = 96
    img, sq = render(renderer, S.checker_sphere(8.0), W, H, 64)
    cls = S.sphere_uv_classes(W, H, 0.0, 1.0, 8.0)
    S.check_emission_classes(img, sq, 64, cls, {0: S.spd_xyz(S.EVEN), 1: S.spd_xyz(S.ODD)}, "sphere")


@pytest.mark.parametrize("spread", [10.0, 4.0])
def test_cone_light(renderer, spread):
    W = H = 96
    img, sq = render(renderer, S.cone_wall(spread), W, H, 32)
    cls = S.cone_classes(W, H, spread)
    S.check_emission_classes(img, sq, 32, cls, {1: S.spd_xyz(S.EVEN), 0: np.zeros(3)}, "cone")


def test_moving_sphere_motion_blur(renderer):
    W, H, spp = 96, 48, 256
    img, sq = render(renderer, S.moving_lamp(), W, H, spp)
    S.motion_check(img, sq, spp, W, H)


def test_moving_sphere_at_rest_equals_sphere(renderer):
    a = S.moving_lamp(static=True)
    b = S.moving_lamp()
    objs = b.objects.copy()
    objs["p"][0][3:6] = objs["p"][0][0:3]
    b._obj_cache = objs
    ia, _ = render(renderer, a, 64, 32, 16)
    ib, _ = render(renderer, b, 64, 32, 16)
    assert np.array_equal(ia, ib) and ia.max() > 0


_cache = {}


def features(renderer, moving_field=False):
    key = ("field" if moving_field else "room", id(renderer))
    if key not in _cache:
        s = synth.spheres(moving=True) if moving_field else synth.features()
        _cache[key] = (renderer.upload(s), po.OracleScene(s))
    return _cache[key]


@pytest.mark.parametrize("field", [False, True], ids=["room", "moving-field"])
def test_primary_hit_ids(renderer, field):
    sid, orc = features(renderer, field)
    W, H = (256, 192) if not field else (320, 180)
    o = abi.render_opts(W, H, 1, 8, seed=4)
    gi, gt = renderer.primary_hits(sid, o)
    oi, ot = orc.primary_hits(o)
    same = gi == oi
    both = (gi != abi.MISS) & (oi != abi.MISS)
    with np.errstate(invalid="ignore"):
        tie = both & (np.abs(gt - ot) <= 1e-5 * np.maximum(np.abs(ot), 1e-30))
    print("%d rays, %.6f equal, %d mismatches (%d t-ties)" % (gi.size, same.mean(), (~same).sum(), (~same & tie).sum()))
    assert same.mean() >= 0.9999 and np.all(tie[~same])
    if field:  # the rising spheres are hit at their positions at the ray's time
        assert np.any(np.isin(oi // 8, np.nonzero(synth.spheres(moving=True).objects["kind"] == abi.OBJ_MOVING_SPHERE)[0]))


@pytest.mark.parametrize("field", [False, True], ids=["room", "moving-field"])
def test_same_streams_same_paths(renderer, field):
    sid, orc = features(renderer, field)
    o = abi.render_opts(128, 96, 1, 8, seed=7)
    g, _ = renderer.render(sid, o)
    c, _, _ = orc.render(o)
    close = np.all(np.abs(g - c) <= 1e-3 * np.abs(c) + 1e-6, axis=-1)
    print("%.4f of pixels equal to float rounding" % close.mean())
    assert close.mean() > 0.95


@pytest.mark.parametrize("field", [False, True], ids=["room", "moving-field"])
def test_independent_streams_agree_within_error(renderer, field):
    sid, orc = features(renderer, field)
    W, H, spp = 128, 96, 64
    g, gsq = renderer.render(sid, abi.render_opts(W, H, spp, 8, seed=1001), want_sq=True)
    c, csq, _ = orc.render(abi.render_opts(W, H, spp, 8, seed=2002), want_sq=True)
    z, _, _ = block_z(g, gsq, c, csq, spp, spp, 8)
    print("%d block-channels, mean z^2 %.3f, max |z| %.2f" % (z.size, np.mean(z * z), np.abs(z).max()))
    assert np.mean(np.abs(z) > 3) <= 0.02
    assert np.abs(z).max() < 6
    assert 0.5 < np.mean(z * z) < 1.6


def test_plain_job_unchanged_beside_a_textured_one(renderer):
    """A batch holding a checker/cone/motion scene shades with the extended
    kernel; a plain scene's job in it must come out as when rendered alone."""
    c = renderer.upload(synth.cornell())
    f, _ = features(renderer)
    oc = abi.render_opts(64, 64, 8, 8, seed=4)
    outs = renderer.render_batch([(f, abi.render_opts(64, 48, 8, 8, seed=2)), (c, oc)])
    ref, _ = renderer.render(c, oc)
    assert np.array_equal(outs[1][0], ref)
    renderer.free(c)
