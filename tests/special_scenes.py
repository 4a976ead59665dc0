"""Scenes with closed-form pixel expectations, for the full estimator.

For both, the per-sample estimator of draw_cpu.rs:130-142 averages
xyz(lambda_i) * L(lambda_i) over the 4 rotated wavelengths; with a uniform hero
each lambda_i is uniform on the 400 integer wavelengths, so the pixel
expectation is (1/400) * sum_lambda xyz(lambda) * L(lambda)."""
import numpy as np

from oracle import py_oracle as po

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


# --------------------------------------------------------------------------
# Closed forms for textures, cone lights and motion.  The wall scenes put a
# 10 x 10 DiffuseLight / ConeLight quad at z = WALL_Z in front of CAM (it fills
# the view); the quad's pixel footprints are rectangles on the wall, so a
# pixel whose four corner rays land in one checker cell (or all inside / all
# outside the cone) sees exactly one emission spectrum.
WALL_Z = 0.25
EVEN, ODD = (1.0, 1.0, 1.0), (0.2, 0.2, 0.2)


def emitter_wall(material, z=WALL_Z):
    return load_scene({
        "time_0": 0, "time_1": 1, "camera": CAM, "background_color": [0, 0, 0],
        "objects": [{"kind": "Quad", "priority": True, "q": [-5, -5, z], "u": [10, 0, 0], "v": [0, 10, 0],
                     "material": "wall"}],
        "materials": [dict(material, name="wall")]})


def checker_wall(kind, density):
    """DiffuseLight wall whose emit texture is a checker of EVEN / ODD cells."""
    return emitter_wall({"kind": "DiffuseLight", "emit": {
        "kind": kind, "even": {"xyz_e": list(EVEN)}, "odd": {"xyz_e": list(ODD)}, "density": density}})


def cone_wall(spread):
    return emitter_wall({"kind": "ConeLight", "spread": spread,
                         "emit": {"kind": "SolidColor", "color": {"xyz_e": list(EVEN)}}})


def pixel_corner_rays(W, H):
    """unit directions of the rays through the 4 corners of every pixel of CAM
    (camera.rs:103-126 with aperture 0): [H, W, 4, 3], row 0 = top"""
    half = np.tan(np.radians(15.0))
    rows, cols = np.mgrid[0:H, 0:W]
    yp = H - 1 - rows
    out = np.zeros((H, W, 4, 3))
    for k, (du, dv) in enumerate(((0, 0), (1, 0), (0, 1), (1, 1))):
        u = (cols + du) / W
        v = (yp + dv) / H
        d = np.stack([(2 * u - 1) * half * (W / H), (2 * v - 1) * half, -np.ones_like(u, dtype=np.float64)], -1)
        out[:, :, k] = d / np.linalg.norm(d, axis=-1, keepdims=True)
    return out


def wall_points(W, H, z=WALL_Z):
    """world (x, y) of the pixel corners on the wall: [H, W, 4, 2]"""
    d = pixel_corner_rays(W, H)
    t = (z - CAM["look_from"][2]) / d[..., 2]
    return d[..., :2] * t[..., None]


def checker_classes(W, H, kind, density):
    """per pixel: 0 even, 1 odd, -1 mixed (the footprint straddles a cell edge)"""
    p = wall_points(W, H)
    if kind == "SpatialChecker":
        a, b = p[..., 0], p[..., 1]              # sin(d pi z) > 0 on the wall: the cell is set by x, y
        assert np.sin(density * np.pi * WALL_Z) > 0.1
    else:                                         # quad (u, v) = ((x + 5) / 10, (y + 5) / 10)
        a, b = (p[..., 0] + 5) / 10, (p[..., 1] + 5) / 10
    ca, cb = np.floor(density * a), np.floor(density * b)
    # keep a margin: corners within 1e-3 cells of an edge count as straddling
    fa, fb = density * a - ca, density * b - cb
    edge = (np.minimum(fa, 1 - fa) < 1e-3) | (np.minimum(fb, 1 - fb) < 1e-3)
    same = (ca == ca[..., :1]).all(-1) & (cb == cb[..., :1]).all(-1) & ~edge.any(-1)
    odd = ((ca[..., 0] + cb[..., 0]) % 2 == 1).astype(np.int64)   # sin(pi k) sign alternates per cell
    return np.where(same, odd, -1)


def cone_classes(W, H, spread):
    """per pixel: 1 lit, 0 dark, -1 mixed (within 0.05 degrees of the cone edge)"""
    d = pixel_corner_rays(W, H)
    ang = np.degrees(np.arccos(np.clip(-d[..., 2], -1, 1)))
    lit = (ang < spread - 0.05).all(-1)
    dark = (ang > spread + 0.05).all(-1)
    return np.where(lit, 1, np.where(dark, 0, -1))


# A lamp sphere sweeping across a black sky, seen from far away (z = 100,
# half-height 2 at z = 0).  Orthographically a pixel ray at (x, y) meets the
# sphere at time t iff |x - c(t)| < h(y) = sqrt(r^2 - y^2), c(t) = X0 + (X1 - X0) t
# for t uniform in [0, 1) (sampler/random.rs:23), so the pixel value is the
# emission times the fraction of time covered, averaged over its footprint.
# Perspective changes the silhouette by < 3e-4 relative at this distance.
MOTION = dict(X0=-1.5, X1=1.5, R=0.6, D=100.0, HALF=2.0)


def moving_lamp(static=False):
    m = MOTION
    fov = 2 * np.degrees(np.arctan(m["HALF"] / m["D"]))
    cam = {"look_from": [0, 0, m["D"]], "look_at": [0, 0, 0], "up": [0, 1, 0], "vertical_fov": fov,
           "aperture": 0, "focus_distance": 10}
    lamp = ({"kind": "Sphere", "center": [m["X0"], 0, 0], "radius": m["R"], "material": "lamp"} if static else
            {"kind": "MovingSphere", "center_0": [m["X0"], 0, 0], "center_1": [m["X1"], 0, 0], "radius": m["R"],
             "material": "lamp"})
    return load_scene({
        "camera": cam, "background_color": [0, 0, 0],
        "objects": [lamp, {"kind": "Sphere", "priority": True, "center": [0, 60, 0], "radius": 1.0, "material": "dark"}],
        "materials": [{"name": "lamp", "kind": "DiffuseLight", "emit": {"kind": "SolidColor", "color": {"xyz_e": [1, 1, 1]}}},
                      {"name": "dark", "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": {"xyz_e": [0, 0, 0]}}}]})


def motion_coverage(W, H, sub=16):
    """expected time-coverage fraction of each pixel of moving_lamp(): [H, W]"""
    m = MOTION
    hw, hh = m["HALF"] * W / H, m["HALF"]
    rows, cols = np.mgrid[0:H, 0:W]
    s = (np.arange(sub) + 0.5) / sub
    x = ((cols[..., None, None] + s[None, None, :, None]) / W * 2 - 1) * hw
    y = (((H - 1 - rows)[..., None, None] + s[None, None, None, :]) / H * 2 - 1) * hh
    h = np.sqrt(np.maximum(m["R"] ** 2 - y ** 2, 0.0))
    lo, hi = min(m["X0"], m["X1"]), max(m["X0"], m["X1"])
    cov = np.clip(np.minimum(x + h, hi) - np.maximum(x - h, lo), 0, None) / (hi - lo)
    return np.where(h > 0, cov, 0.0).mean((2, 3))


def sphere_uv_classes(W, H, center_z, radius, density):
    """Surface checker on a lamp sphere (centre (0, 0, center_z)) seen by CAM:
    per pixel 0 even, 1 odd, -1 mixed or off the sphere.  (u, v) of
    sphere.rs:58-65 at the corner and centre rays' hit points."""
    d = pixel_corner_rays(W, H)
    half = np.tan(np.radians(15.0))
    rows, cols = np.mgrid[0:H, 0:W]
    uc = (cols + 0.5) / W
    vc = (H - 1 - rows + 0.5) / H
    dc = np.stack([(2 * uc - 1) * half * (W / H), (2 * vc - 1) * half, -np.ones_like(uc, dtype=np.float64)], -1)
    dc /= np.linalg.norm(dc, axis=-1, keepdims=True)
    d = np.concatenate([d, dc[:, :, None]], 2)                     # [H, W, 5, 3]
    o = np.array(CAM["look_from"], np.float64) - np.array([0, 0, center_z])
    b = (d * o).sum(-1)
    disc = b * b - (o @ o - radius * radius)
    ok = disc > 1e-3
    t = -b - np.sqrt(np.maximum(disc, 0))
    p = (o + t[..., None] * d) / radius
    phi = np.arctan2(p[..., 2], p[..., 0])
    theta = np.arcsin(np.clip(p[..., 1], -1, 1))
    u = 1 - (phi + np.pi) / (2 * np.pi)
    v = (theta + np.pi / 2) / np.pi
    ca, cb = np.floor(density * u), np.floor(density * v)
    fa, fb = density * u - ca, density * v - cb
    edge = (np.minimum(fa, 1 - fa) < 2e-3) | (np.minimum(fb, 1 - fb) < 2e-3)
    same = (ca == ca[..., :1]).all(-1) & (cb == cb[..., :1]).all(-1) & ~edge.any(-1) & ok.all(-1)
    odd = ((ca[..., 0] + cb[..., 0]) % 2 == 1).astype(np.int64)
    return np.where(same, odd, -1)


def checker_sphere(density, center_z=0.0, radius=1.0):
    return load_scene({
        "camera": CAM, "background_color": [0, 0, 0],
        "objects": [{"kind": "Sphere", "priority": True, "center": [0, 0, center_z], "radius": radius, "material": "ball"}],
        "materials": [{"name": "ball", "kind": "DiffuseLight", "emit": {
            "kind": "SurfaceChecker", "even": {"xyz_e": list(EVEN)}, "odd": {"xyz_e": list(ODD)}, "density": density}}]})


def checker_triangle(density):
    """the lower-left half (x + y <= 0) of the wall as one scene-file Triangle with a surface checker"""
    return load_scene({
        "camera": CAM, "background_color": [0, 0, 0],
        "objects": [{"kind": "Triangle", "priority": True, "q": [-5, -5, WALL_Z], "u": [10, 0, 0], "v": [0, 10, 0],
                     "material": "wall"}],
        "materials": [{"name": "wall", "kind": "DiffuseLight", "emit": {
            "kind": "SurfaceChecker", "even": {"xyz_e": list(EVEN)}, "odd": {"xyz_e": list(ODD)}, "density": density}}]})


def triangle_classes(W, H, density):
    c = checker_classes(W, H, "SurfaceChecker", density)
    p = wall_points(W, H)
    inside = ((p[..., 0] + p[..., 1]) < -1e-2).all(-1)
    outside = ((p[..., 0] + p[..., 1]) > 1e-2).all(-1)
    return np.where(inside, c, np.where(outside, 2, -1))    # 2: background (black)


def check_emission_classes(img, sq, spp, classes, values, label):
    """Pixels of class k see emission spectrum values[k] (an XYZ expectation,
    or 0 for dark): every pure pixel is nearer its own class's value than any
    other's, and each class's mean lies within 4 standard errors of its value
    (exactly 0 for dark classes)."""
    img = img.astype(np.float64)
    keys = sorted(values)
    for k in keys:
        m = classes == k
        if not m.any():
            continue
        want = np.asarray(values[k], np.float64)
        if not np.any(want):
            assert np.all(img[m] == 0.0), label + ": class %d should be dark" % k
            continue
        mean = img[m].mean(0)
        var = (sq[m] - img[m] ** 2).mean(0)
        se = np.sqrt(np.maximum(var, 0) / (m.sum() * spp)) + 1e-9
        print("%s class %d: %d px, mean %s want %s (se %s)" % (label, k, m.sum(), mean, want, se))
        assert np.all(np.abs(mean - want) < 4 * se + 1e-6 * np.abs(want)), (label, k, mean, want, se)
        # nearest class value per pixel, on Y
        dist = np.stack([np.abs(img[m][:, 1] - np.asarray(values[j])[1]) for j in keys], -1)
        assert np.all(np.array(keys)[dist.argmin(-1)] == k), (label, k)


def spd_xyz(c):
    cie = np.array([po.wavelength_into_xyz(l) for l in range(380, 780)], np.float64)
    return expected_pixel([po.spectral_power(c, l) for l in range(380, 780)], cie)


def motion_check(img, sq, spp, W, H):
    cov = motion_coverage(W, H)
    e = spd_xyz((1.0, 1.0, 1.0))
    want = cov[..., None] * e
    # pixels no part of which is ever covered stay black (allow one pixel of slack for perspective)
    pad = np.pad(cov, 1)
    near = np.max([pad[1 + dy:1 + dy + H, 1 + dx:1 + dx + W] for dy in (-1, 0, 1) for dx in (-1, 0, 1)], 0)
    assert np.all(img[near == 0] == 0.0)
    m = cov > 0.02
    var = np.maximum(sq[m] - img[m] ** 2, 0)
    se = np.sqrt(var / spp) + 1e-12
    z = (img[m] - want[m]) / se
    print("motion: %d pixels, mean z^2 %.3f, max |z| %.2f" % (m.sum(), np.mean(z * z), np.abs(z).max()))
    assert 0.6 < np.mean(z * z) < 1.5 and np.abs(z).max() < 6
    # total: the swept disc integrates to pi r^2 whatever the motion
    tot = img.sum((0, 1))
    se_tot = np.sqrt(np.maximum(sq - img ** 2, 0).sum((0, 1)) / spp)
    assert np.all(np.abs(tot - want.sum((0, 1))) < 4 * se_tot), (tot, want.sum((0, 1)), se_tot)
    return cov


# --------------------------------------------------------------------------
# Constant media (objects/constant_medium.rs).  The probe camera looks down -z
# from z = 10 with a 0.05-degree field of view, so every primary ray runs
# within 5e-3 of the axis; camera rays are unit length (camera.rs:115-119), so
# a ray's parameter is its world distance.  A lamp sphere far out of view is
# the light-sampling target the reference requires.
MEDIUM_CAM = {"look_from": [0, 0, 10], "look_at": [0, 0, 0], "up": [0, 1, 0], "vertical_fov": 0.05,
              "aperture": 0, "focus_distance": 1}


def medium_probe(boundary, density, outer=None):
    """One medium (object 0, report id 0) with the given boundary JSON,
    optionally wrapped by outer(obj)."""
    med = {"kind": "ConstantMedium", "density": density, "boundary": boundary,
           "texture": {"kind": "SolidColor", "color": {"xyz_e": [0.5, 0.5, 0.5]}}}
    if outer:
        med = outer(med)
    return load_scene({
        "time_0": 0, "time_1": 1, "camera": MEDIUM_CAM, "background_color": [0, 0, 0],
        "objects": [med, {"kind": "Sphere", "priority": True, "center": [0, 50, 0], "radius": 1, "material": "lamp"}],
        "materials": [{"name": "lamp", "kind": "DiffuseLight", "emit": {"kind": "SolidColor", "color": {"xyz_e": [1, 1, 1]}}}]})


def free_flight_check(ids, ts, entry, chord, sigma, label):
    """Primary hits through a medium of optical depth sigma * chord: the medium
    (id 0) is hit with probability 1 - exp(-sigma chord), and the hit lies an
    exponential distance of rate sigma past the entry, truncated at the chord."""
    from scipy import stats
    n = ids.size
    hit = ids == 0
    p = 1.0 - np.exp(-sigma * chord)
    se = np.sqrt(p * (1 - p) / n)
    frac = hit.mean()
    print("%s: hit fraction %.4f want %.4f (se %.4f)" % (label, frac, p, se))
    assert abs(frac - p) < 4 * se + 1e-6, (label, frac, p)
    assert np.all((ids == 0) | (ids == 0xFFFFFFFF)), label
    x = ts[hit].astype(np.float64) - entry
    assert x.min() > -1e-3 and x.max() < chord + 1e-3, (label, x.min(), x.max())
    norm = 1.0 - np.exp(-sigma * chord)
    ks = stats.kstest(np.clip(x, 0, chord), lambda v: (1.0 - np.exp(-sigma * np.clip(v, 0, chord))) / norm)
    print("%s: KS p = %.3g over %d hits" % (label, ks.pvalue, hit.sum()))
    assert ks.pvalue > 1e-3, (label, ks)


def medium_furnace(density=1.0, background=0.5):
    """A unit fog sphere of white Isotropic albedo (Xyz<E>(1,1,1): a flat
    spectrum) under a uniform grey sky: nothing absorbs, every path ends in
    the sky, so every pixel's expectation is the sky's (with a sky-radiance
    lamp sphere as the light-sampling target, as in furnace())."""
    from clovers_amd.color import srgb_eotf
    g = float(srgb_eotf(background))
    return load_scene({
        "time_0": 0, "time_1": 1, "camera": CAM, "background_color": [background] * 3,
        "objects": [{"kind": "ConstantMedium", "density": density,
                     "texture": {"kind": "SolidColor", "color": {"xyz_e": [1, 1, 1]}},
                     "boundary": {"kind": "Sphere", "center": [0, 0, 0], "radius": 1.0}},
                    {"kind": "Sphere", "priority": True, "center": [0, 3, 0], "radius": 0.5, "material": "sky"}],
        "materials": [{"name": "sky", "kind": "DiffuseLight", "emit": {"kind": "SolidColor", "color": {"xyz_e": [g] * 3}}}]})
