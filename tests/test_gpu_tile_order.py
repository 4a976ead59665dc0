"""Camera rays are laid out in 8x8 pixel tiles within each sample of a pass;
every path's random stream is keyed by (pixel, sample) and each pixel's
samples are summed in sample order, so the images must be bit-identical to
those of a context that lays them out row by row (CLV_TILE_ORDER=0, read when
a context is created)."""
import os

import numpy as np
import pytest

from clovers_amd import abi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def row_major():
    import clovers_amd
    old = os.environ.get("CLV_TILE_ORDER")
    os.environ["CLV_TILE_ORDER"] = "0"
    try:
        r = clovers_amd.Renderer(0)
    finally:
        if old is None:
            del os.environ["CLV_TILE_ORDER"]
        else:
            os.environ["CLV_TILE_ORDER"] = old
    yield r
    r.close()


def render_both(tiled, rows, scene, opts):
    out = []
    for r in (tiled, rows):
        sid = r.upload(scene)
        try:
            out.append(r.render(sid, opts, want_sq=True))
        finally:
            r.free(sid)
    return out


@pytest.mark.parametrize("W,H", [(96, 64), (101, 37), (7, 5)], ids=["tiles", "ragged", "one-tile"])
def test_cornell_identical(renderer, row_major, W, H):
    (a, asq), (b, bsq) = render_both(renderer, row_major, synth.cornell(), abi.render_opts(W, H, 8, 8, seed=5))
    assert np.array_equal(a, b)
    assert np.array_equal(asq, bsq)


def test_sphere_field_identical(renderer, row_major):
    (a, _), (b, _) = render_both(renderer, row_major, synth.spheres(), abi.render_opts(130, 75, 4, 16, seed=9))
    assert np.array_equal(a, b)


def test_uneven_batch_identical(renderer, row_major):
    """jobs of different sizes share passes; each job's tiles follow its own width and height"""
    outs = []
    for r in (renderer, row_major):
        s1, s2 = r.upload(synth.cornell()), r.upload(synth.spheres())
        try:
            outs.append(r.render_batch([(s1, abi.render_opts(45, 29, 6, 8, seed=1)),
                                        (s2, abi.render_opts(64, 48, 3, 10, seed=2)),
                                        (s1, abi.render_opts(17, 90, 5, 8, seed=3))]))
        finally:
            r.free(s2)
            r.free(s1)
    for (a, _), (b, _) in zip(*outs):
        assert np.array_equal(a, b)
