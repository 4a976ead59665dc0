"""The debug render modes on the device (clv_render_debug; draw_cpu.rs:148-207):
the normal map against the oracle's (itself pinned by a closed form), and the
count modes' ramp and per-ray counts against the counting traversal's totals."""
import numpy as np
import pytest

from clovers_amd import abi, synth
from clovers_amd.color import M_D65_TO_E, M_SRGB_TO_XYZ_D65
from oracle import py_oracle as po

pytestmark = pytest.mark.gpu

LIN_TO_E = M_D65_TO_E @ M_SRGB_TO_XYZ_D65


@pytest.mark.parametrize("name", ["features", "cornell", "moving"])
def test_normal_map_matches_oracle(renderer, name):
    s = {"features": synth.features, "cornell": synth.cornell, "moving": lambda: synth.spheres(moving=True)}[name]()
    o = abi.render_opts(160, 120, 1, 1, seed=6)
    sid = renderer.upload(s)
    img = renderer.render_debug(sid, o, abi.MODE_NORMAL_MAP)
    renderer.free(sid)
    want = np.einsum("ij,hwj->hwi", LIN_TO_E, po.OracleScene(s).normal_map(o).astype(np.float64))
    close = np.all(np.abs(img - want) <= 2e-5, axis=-1)
    print("%s: %.5f of pixels equal within 2e-5" % (name, close.mean()))
    # The rest are exact ties: a pixel-corner ray along an edge where two
    # surfaces meet (the Cornell room's wall/ceiling edges) hits both at the
    # same distance, and which one is reported depends on the order the tree
    # tests them in (as in the reference, whose own tree decides).  Such a
    # pixel shows the surface on the other side of the edge, which a
    # neighbouring pixel of the oracle's image shows too.
    assert close.mean() >= 0.995
    H, W = close.shape
    for y, x in zip(*np.nonzero(~close)):
        nb = want[max(y - 1, 0):y + 2, max(x - 1, 0):x + 2].reshape(-1, 3)
        assert np.any(np.all(np.abs(nb - img[y, x]) <= 2e-5, axis=-1)), (name, y, x, img[y, x], want[y, x])


def decode_counts(img):
    lin = np.einsum("ij,hwj->hwi", np.linalg.inv(LIN_TO_E), img.astype(np.float64))
    grey = np.all(np.abs(lin - lin[..., :1]) < 1e-5, -1)
    k = np.rint(lin[..., 0] * 255).astype(np.int64)
    ramp = np.full(img.shape[:2], -1)
    ramp[grey] = k[grey]
    for val, code in (((1, 1, 0), 256), ((1, 0.5, 0), 512), ((1, 0, 0), 1024)):
        ramp[np.all(np.abs(lin - val) < 1e-5, -1)] = code
    return ramp


@pytest.mark.parametrize("mode,key", [(abi.MODE_BVH_TEST_COUNT, "node_visits"),
                                      (abi.MODE_PRIMITIVE_TEST_COUNT, "prim_tests")])
def test_count_modes(renderer, mode, key):
    for scene, W, H in ((synth.cornell(), 96, 96), (synth.mesh(freq=24), 96, 54)):
        sid = renderer.upload(scene)
        img = renderer.render_debug(sid, abi.render_opts(W, H, 1, 1, seed=2), mode)
        total = renderer.counters()[key]
        renderer.free(sid)
        k = decode_counts(img)
        assert np.all(k >= 0), "a pixel outside the ramp's colours"
        small = k < 256
        print("mode %d: %d px, counts %d..%d, total %d" % (mode, k.size, k[small].min(), k[small].max(), total))
        assert k.min() >= (1 if mode == abi.MODE_BVH_TEST_COUNT else 0)   # the root is always fetched
        if small.all():
            assert k.sum() == total
        else:
            assert k[small].sum() + 256 * (~small).sum() <= total


def test_debug_mode_errors(renderer):
    import clovers_amd
    sid = renderer.upload(synth.cornell())
    with pytest.raises(clovers_amd.ClvError):
        renderer.render_debug(sid, abi.render_opts(8, 8, 1, 1), abi.MODE_PATH_TRACING)
    with pytest.raises(clovers_amd.ClvError):
        renderer.render_debug(sid, abi.render_opts(8, 8, 1, 1), 9)
    renderer.free(sid)
