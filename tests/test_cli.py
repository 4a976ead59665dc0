"""clovers-render, the command-line renderer over the C ABI (the reference's
`clovers render`, clovers-cli/src/render.rs:17-182): argument handling on the
CPU (every case here fails before a device is touched); the GPU test renders
a scene file to PNG and EXR and compares with the library called directly."""
import json
import os
import subprocess

import numpy as np
import pytest

import clovers_amd
from clovers_amd import abi, synth
from test_write import read_exr, read_png

CLI = os.path.join(os.path.dirname(clovers_amd.__file__), "clovers-render")


def run(*args, **kw):
    return subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=300, **kw)


@pytest.mark.parametrize("args,code", [
    ((), 2),
    (("-i", "x.json", "-m", "wireframe"), 2),
    (("-i", "x.json", "--sampler", "blue"), 2),
    (("-i", "x.json", "-f", "jpg"), 2),
    (("-i", "x.json", "-w", "abc"), 2),
    (("-i", "x.json", "--builder", "gpu"), 2),
    (("-i", "/nonexistent/scene.json"), 1),
])
def test_arguments(args, code):
    r = run(*args)
    assert r.returncode == code, (r.stdout, r.stderr)
    assert "error" in r.stderr


def test_scene_errors_are_reported(tmp_path):
    sf = synth.cornell_json()
    sf["objects"].append({"kind": "ConstantMedium", "boundary": {"kind": "Quad", "q": [0, 0, 0], "u": [1, 0, 0],
                                                                 "v": [0, 1, 0]}})
    p = tmp_path / "s.json"
    p.write_text(json.dumps(sf))
    r = run("-i", p)
    assert r.returncode == 1 and "ConstantMedium" in r.stderr


@pytest.mark.gpu
def test_render_scene_file(tmp_path, renderer):
    sf = synth.features_json()
    p = tmp_path / "features.json"
    p.write_text(json.dumps(sf))
    out = tmp_path / "out" / "features"
    r = run("-i", p, "-w", 96, "-h", 64, "-s", 8, "-d", 8, "-f", "png", "exr", "-o", out, "--seed", 3)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    img, _ = read_exr(str(out) + ".exr")
    sid = renderer.upload(clovers_amd.load_scene_file(str(p)))
    want, _ = renderer.render(sid, abi.render_opts(96, 64, 8, 8, seed=3))
    renderer.free(sid)
    assert np.array_equal(img, want)
    px, texts = read_png(str(out) + ".png")
    assert px.shape == (64, 96, 3) and px.max() > 0
    assert b"8 samples per pixel" in texts[b"Comment"]
    r = run("-i", p, "-w", 64, "-h", 48, "-m", "normal-map", "-f", "exr", "-o", tmp_path / "n", "--seed", 3)
    assert r.returncode == 0, r.stderr
    nm, _ = read_exr(str(tmp_path / "n") + ".exr")
    sid = renderer.upload(clovers_amd.load_scene_file(str(p)))
    want = renderer.render_debug(sid, abi.render_opts(64, 48, 1, 1, seed=3), abi.MODE_NORMAL_MAP)
    renderer.free(sid)
    assert np.array_equal(nm, want)
    # the scene BVH built on the GPU: the same image (the room has no ties between equally distant hits)
    r = run("-i", p, "-w", 96, "-h", 64, "-s", 8, "-d", 8, "-f", "exr", "-o", tmp_path / "d", "--seed", 3,
            "--builder", "device")
    assert r.returncode == 0, r.stderr
    dimg, _ = read_exr(str(tmp_path / "d") + ".exr")
    assert np.array_equal(dimg, img)
