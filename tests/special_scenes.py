"""Scenes with closed-form pixel expectations, for the full estimator.

For both, the per-sample estimator of draw_cpu.rs:130-142 averages
xyz(lambda_i) * L(lambda_i) over the 4 rotated wavelengths; with a uniform hero
each lambda_i is uniform on the 400 integer wavelengths, so the pixel
expectation is (1/400) * sum_lambda xyz(lambda) * L(lambda)."""
import numpy as np

from clovers_amd.scenefile import load_scene

CAM = {"look_from": [0, 0, 4], "look_at": [0, 0, 0], "up": [0, 1, 0], "vertical_fov": 30, "aperture": 0,
       "focus_distance": 10}


def direct_emitter():
    """A 10x10 emitter quad (SolidColor Xyz<E>(1,1,1), front face towards the
    camera) fills the view: every primary ray returns the emitter SPD."""
    return load_scene({
        "time_0": 0, "time_1": 1, "camera": CAM, "background_color": [0, 0, 0],
        "objects": [{"kind": "Quad", "priority": True, "q": [-5, -5, 0], "u": [10, 0, 0], "v": [0, 10, 0],
                     "material": "lamp"}],
        "materials": [{"name": "lamp", "kind": "DiffuseLight", "emit": {"kind": "SolidColor", "color": {"xyz_e": [1, 1, 1]}}}]})


def furnace(albedo=0.5, background=0.5):
    """Lambertian sphere of albedo Xyz<E>(a,a,a) under a uniform grey sky
    (legacy sRGB grey b -> Xyz<E>(g,g,g)); a convex object, so L = a * B on the
    sphere and B elsewhere.  The light-sampling target the reference requires
    is an emitting sphere out of view whose radiance equals the sky's, so it is
    indistinguishable from sky in every path (a tiny, far target would not do:
    with r/d ~ 1e-4 the f32 sphere test itself misses half of the directions
    aimed at it)."""
    from clovers_amd.color import srgb_eotf
    g = float(srgb_eotf(background))
    return load_scene({
        "time_0": 0, "time_1": 1, "camera": CAM, "background_color": [background] * 3,
        "objects": [{"kind": "Sphere", "center": [0, 0, 0], "radius": 1.0, "material": "diffuse"},
                    {"kind": "Sphere", "priority": True, "center": [0, 3, 0], "radius": 0.5, "material": "sky"}],
        "materials": [{"name": "diffuse", "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": {"xyz_e": [albedo] * 3}}},
                      {"name": "sky", "kind": "DiffuseLight", "emit": {"kind": "SolidColor", "color": {"xyz_e": [g] * 3}}}]})


def expected_pixel(spd_values, cie_rows):
    """(1/400) sum xyz(lambda) L(lambda): spd_values[400], cie_rows[400, 3]."""
    return (cie_rows * np.asarray(spd_values)[:, None]).sum(0) / 400.0


def sphere_mask(W, H, radius_px_frac=0.9):
    """Pixels whose footprint lies well inside the furnace sphere's silhouette."""
    # camera at z=4, vfov 30: the unit sphere's angular radius is asin(1/4)
    half = np.tan(np.radians(15.0))
    ang = np.arcsin(0.25)
    ys, xs = np.mgrid[0:H, 0:W]
    u = ((xs + 0.5) / W - 0.5) * 2 * half * (W / H)
    v = ((ys + 0.5) / H - 0.5) * 2 * half
    r = np.sqrt(u * u + v * v)
    return r < np.tan(ang) * radius_px_frac
