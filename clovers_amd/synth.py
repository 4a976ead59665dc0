"""Seeded synthetic scenes of the measured workloads.

Every scene is generated from a seeded PCG64 stream (numpy), so the GPU run and
the CPU oracle see the same inputs.  Scenes are expressed in the reference's
scene-file schema (clovers-cli/src/scenefile.rs:21-29) and go through the same
loader as a scene file; large meshes are added in bulk as the reference
flattens mesh lists into one hitable per triangle (scenefile.rs:84-94).

  cornell      Cornell box: 5 wall quads (0.73 grey, red/green sides), two
               Translate{RotateY{Boxy}} boxes, one 130x105 D65 emitter of
               radiance ~15 (D65 is 100 at 560 nm, intensity 0.15).
               Reference analogue: scenes/cornell.json.
  spheres      22x22 grid of r=0.2 spheres (80% Lambertian albedo U*U, 15% metal
               fuzz U[0,0.5], 5% glass n=1.5) + 3 r=1 spheres + ground r=1000.
               Reference analogue: scenes/one_weekend.json (static spheres; with
               moving=True the Lambertian ones rise as MovingSpheres, as there).
  prism        Cauchy-glass triangular prism (12 triangles), glass sphere r=0.5,
               grey Lambertian floor, two 0.1x0.1 emitters of radiance 500.
               Reference analogue: scenes/dispersive.json.
  features     checker textures, cone light and moving spheres in one room.
  mesh         Frequency-224 geodesic icosphere (1,003,520 triangles) displaced
               by 4-octave value noise (amplitude 0.15), thin-film dielectric,
               quad light, 4 camera views.
"""
import math

import numpy as np

from . import abi
from .scenefile import load_scene


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


# ----------------------------------------------------------------- cornell
def cornell_json():
    grey = {"kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": [0.73, 0.73, 0.73]}}
    return {
        "time_0": 0, "time_1": 1,
        "camera": {"look_from": [278, 278, -800], "look_at": [278, 278, 0], "up": [0, 1, 0],
                   "vertical_fov": 40, "aperture": 0, "focus_distance": 10},
        "background_color": [0, 0, 0],
        "objects": [
            {"kind": "Quad", "q": [555, 0, 0], "u": [0, 0, 555], "v": [0, 555, 0], "material": "green wall"},
            {"kind": "Quad", "q": [0, 0, 555], "u": [0, 0, -555], "v": [0, 555, 0], "material": "red wall"},
            {"kind": "Quad", "q": [0, 0, 0], "u": [555, 0, 0], "v": [0, 0, 555], "material": "grey wall"},
            {"kind": "Quad", "q": [0, 555, 0], "u": [555, 0, 0], "v": [0, 0, 555], "material": "grey wall"},
            {"kind": "Quad", "q": [0, 0, 555], "u": [555, 0, 0], "v": [0, 555, 0], "material": "grey wall"},
            {"kind": "Quad", "priority": True, "q": [213, 554, 227], "u": [130, 0, 0], "v": [0, 0, 105],
             "material": "lamp"},
            {"kind": "Translate", "offset": [265, 0, 295],
             "object": {"kind": "RotateY", "angle": 15,
                        "object": {"kind": "Boxy", "corner_0": [0, 0, 0], "corner_1": [165, 330, 165],
                                   "material": "grey wall"}}},
            {"kind": "Translate", "offset": [130, 0, 65],
             "object": {"kind": "RotateY", "angle": -18,
                        "object": {"kind": "Boxy", "corner_0": [0, 0, 0], "corner_1": [165, 165, 165],
                                   "material": "grey wall"}}},
        ],
        "materials": [
            {"name": "lamp", "kind": "DiffuseLight", "emit": {"kind": "IlluminantD65", "intensity": 0.15}},
            {"name": "green wall", "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": [0.12, 0.45, 0.15]}},
            {"name": "red wall", "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": [0.65, 0.05, 0.05]}},
            dict(grey, name="grey wall"),
        ],
    }


def cornell():
    return load_scene(cornell_json())


# ----------------------------------------------------------------- spheres
def spheres_json(seed=1, moving=False):
    """moving=True: the Lambertian small spheres rise by U[0, 0.5] over the
    shutter as MovingSpheres, as in the reference's scenes/one_weekend.json;
    the rises come from a second stream, so the layout is the static one."""
    r = _rng(seed)
    U = lambda: float(r.random())
    rise = _rng(seed + 7919)
    mats = [
        {"name": "ground material", "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": [0.5, 0.5, 0.5]}},
        {"name": "glass", "kind": "Dielectric", "refractive_index": 1.5},
        {"name": "lambertian sphere material", "kind": "Lambertian",
         "albedo": {"kind": "SolidColor", "color": [0.4, 0.2, 0.1]}},
        {"name": "metal sphere material", "kind": "Metal", "fuzz": 0.0,
         "albedo": {"kind": "SolidColor", "color": [0.7, 0.6, 0.5]}},
    ]
    objs = [
        {"kind": "Sphere", "center": [0.0, -1000.0, 0.0], "radius": 1000.0, "material": "ground material"},
        {"kind": "Sphere", "priority": True, "center": [0.0, 1.0, 0.0], "radius": 1.0, "material": "glass"},
        {"kind": "Sphere", "center": [-4.0, 1.0, 0.0], "radius": 1.0, "material": "lambertian sphere material"},
        {"kind": "Sphere", "center": [4.0, 1.0, 0.0], "radius": 1.0, "material": "metal sphere material"},
    ]
    for a in range(-11, 11):
        for b in range(-11, 11):
            name = "sphere material %d_%d" % (a, b)
            choose = U()
            center = [a + 0.9 * U(), 0.2, b + 0.9 * U()]
            if choose < 0.8:
                col = [U() * U(), U() * U(), U() * U()]
                mats.append({"name": name, "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": col}})
                if moving:
                    c1 = [center[0], center[1] + 0.5 * float(rise.random()), center[2]]
                    objs.append({"kind": "MovingSphere", "center_0": center, "center_1": c1, "radius": 0.2,
                                 "material": name})
                else:
                    objs.append({"kind": "Sphere", "center": center, "radius": 0.2, "material": name})
            elif choose < 0.95:
                col = [0.5 + 0.5 * U(), 0.5 + 0.5 * U(), 0.5 + 0.5 * U()]
                fuzz = 0.5 * U()
                mats.append({"name": name, "kind": "Metal", "fuzz": fuzz,
                             "albedo": {"kind": "SolidColor", "color": col}})
                objs.append({"kind": "Sphere", "center": center, "radius": 0.2, "material": name})
            else:
                objs.append({"kind": "Sphere", "center": center, "radius": 0.2, "material": "glass"})
    return {
        "time_0": 0, "time_1": 1,
        "camera": {"look_from": [13.0, 2.0, 3.0], "look_at": [0.0, 0.0, 0.0], "up": [0.0, 1.0, 0.0],
                   "vertical_fov": 25.0, "aperture": 0.0, "focus_distance": 10.0},
        "background_color": [0.7, 0.7, 0.9],
        "objects": objs,
        "materials": mats,
    }


def spheres(seed=1, moving=False):
    return load_scene(spheres_json(seed, moving))


# -------------------------------------------------------------- features
def features_json():
    """Every texture, light and motion kind of the scene schema in one box:
    a spatial-checker floor, a surface-checker back wall and triangle, moving
    spheres (one inside Translate), a surface-checker metal sphere, a
    spatial-checker box under Translate{RotateY}, a ceiling lamp and a cone
    light aimed into the room (both light-sampling targets)."""
    chk = lambda kind, e, o, d: {"kind": kind, "even": {"xyz_e": e}, "odd": {"xyz_e": o}, "density": d}
    mats = [
        {"name": "floor", "kind": "Lambertian", "albedo": chk("SpatialChecker", [0.8, 0.8, 0.8], [0.2, 0.25, 0.3], 1.5)},
        {"name": "wall", "kind": "Lambertian", "albedo": chk("SurfaceChecker", [0.7, 0.3, 0.2], [0.3, 0.6, 0.3], 12.0)},
        {"name": "ball", "kind": "Metal", "fuzz": 0.2, "albedo": chk("SurfaceChecker", [0.9, 0.9, 0.9], [0.4, 0.3, 0.8], 10.0)},
        {"name": "blue", "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": [0.2, 0.3, 0.8]}},
        {"name": "boxm", "kind": "Metal", "fuzz": 0.05, "albedo": chk("SpatialChecker", [0.9, 0.8, 0.5], [0.3, 0.3, 0.3], 4.0)},
        {"name": "lamp", "kind": "DiffuseLight", "emit": {"kind": "IlluminantD65", "intensity": 0.12}},
        {"name": "cone", "kind": "ConeLight", "spread": 25.0, "emit": chk("SurfaceChecker", [40, 40, 40], [5, 5, 5], 3.0)},
        {"name": "glass", "kind": "Dielectric", "refractive_index": 1.5},
    ]
    objs = [
        {"kind": "Quad", "q": [-4, 0, -4], "u": [8, 0, 0], "v": [0, 0, 8], "material": "floor"},
        {"kind": "Quad", "q": [-4, 0, -4], "u": [0, 4, 0], "v": [8, 0, 0], "material": "wall"},
        {"kind": "Quad", "priority": True, "q": [-1, 3.99, -1], "u": [2, 0, 0], "v": [0, 0, 2], "material": "lamp"},
        {"kind": "Quad", "priority": True, "q": [-3.9, 0.5, 2.5], "u": [0, 0, -1], "v": [0, 1, 0], "material": "cone"},
        {"kind": "Triangle", "q": [1.5, 0.01, -3.5], "u": [2, 0, 0], "v": [0, 2.5, 0], "material": "wall"},
        {"kind": "MovingSphere", "center_0": [-1.5, 0.5, 0.5], "center_1": [-1.0, 1.2, 0.5], "radius": 0.5, "material": "blue"},
        {"kind": "Translate", "offset": [1.6, 0, 1.0], "object":
            {"kind": "MovingSphere", "center_0": [0, 0.4, 0], "center_1": [0.6, 0.4, 0], "radius": 0.4, "material": "glass"}},
        {"kind": "Sphere", "center": [0.2, 0.8, -1.2], "radius": 0.8, "material": "ball"},
        {"kind": "Translate", "offset": [-2.3, 0, -2.0], "object":
            {"kind": "RotateY", "angle": 25, "object": {"kind": "Boxy", "corner_0": [0, 0, 0], "corner_1": [1, 1.6, 1], "material": "boxm"}}},
    ]
    return {
        "time_0": 0, "time_1": 1,
        "camera": {"look_from": [0.0, 2.2, 7.5], "look_at": [0.0, 1.0, 0.0], "up": [0.0, 1.0, 0.0],
                   "vertical_fov": 45.0, "aperture": 0.02, "focus_distance": 7.0},
        "background_color": [0.05, 0.05, 0.08],
        "objects": objs,
        "materials": mats,
    }


def features():
    return load_scene(features_json())


def media_json():
    """Constant media in a lit room (objects/constant_medium.rs, the smoke
    boxes of the reference's cornell_smoke / the_next_week scenes): a fog
    sphere of grey Isotropic albedo, a smoke box with a spatial-checker
    albedo under Translate{RotateY} with its own wrapper inside the boundary,
    a thin blue haze around everything, a Lambertian floor and a lamp."""
    mats = [
        {"name": "floor", "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": [0.7, 0.7, 0.7]}},
        {"name": "lamp", "kind": "DiffuseLight", "emit": {"kind": "IlluminantD65", "intensity": 0.15}},
        {"name": "ball", "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": [0.8, 0.3, 0.2]}},
    ]
    objs = [
        {"kind": "Quad", "q": [-5, 0, -5], "u": [10, 0, 0], "v": [0, 0, 10], "material": "floor"},
        {"kind": "Quad", "priority": True, "q": [-1, 3.99, -1], "u": [2, 0, 0], "v": [0, 0, 2], "material": "lamp"},
        {"kind": "Sphere", "center": [1.6, 0.5, 0.8], "radius": 0.5, "material": "ball"},
        {"kind": "ConstantMedium", "density": 1.2,
         "texture": {"kind": "SolidColor", "color": [0.9, 0.9, 0.9]},
         "boundary": {"kind": "Sphere", "center": [-1.2, 1.0, 0.0], "radius": 0.9}},
        {"kind": "Translate", "offset": [0.6, 0.0, -1.2], "object":
            {"kind": "ConstantMedium", "density": 2.5,
             "texture": {"kind": "SpatialChecker", "even": {"xyz_e": [0.9, 0.8, 0.6]}, "odd": {"xyz_e": [0.2, 0.2, 0.3]},
                         "density": 2.0},
             "boundary": {"kind": "RotateY", "angle": 30, "object":
                          {"kind": "Boxy", "corner_0": [0, 0, 0], "corner_1": [1.0, 1.8, 1.0]}}}},
        {"kind": "ConstantMedium", "density": 0.02,
         "texture": {"kind": "SolidColor", "color": [0.6, 0.7, 1.0]},
         "boundary": {"kind": "Sphere", "center": [0.0, 0.0, 0.0], "radius": 20.0}},
    ]
    return {
        "time_0": 0, "time_1": 1,
        "camera": {"look_from": [0.0, 2.0, 7.0], "look_at": [0.0, 1.0, 0.0], "up": [0.0, 1.0, 0.0],
                   "vertical_fov": 45.0, "aperture": 0.0, "focus_distance": 7.0},
        "background_color": [0.1, 0.1, 0.12],
        "objects": objs,
        "materials": mats,
    }


def media():
    return load_scene(media_json())


# ------------------------------------------------------------------- prism
def prism_triangles(center=(0.0, 0.55, 0.0), side=0.8, length=1.2, yaw_deg=30.0):
    """12 triangles: an equilateral prism along z whose two caps are
    fan-triangulated around their centroid (3 + 3) and whose three sides are
    two triangles each; outward winding."""
    h = side * math.sqrt(3.0) / 2.0
    tri2d = [(-side / 2, -h / 3), (side / 2, -h / 3), (0.0, 2 * h / 3)]
    c, s = math.cos(math.radians(yaw_deg)), math.sin(math.radians(yaw_deg))

    def P(x, y, z):
        # rotate about y by yaw, then translate
        return (c * x + s * z + center[0], y + center[1], -s * x + c * z + center[2])
    z0, z1 = -length / 2, length / 2
    f = [P(x, y, z0) for x, y in tri2d]
    b = [P(x, y, z1) for x, y in tri2d]
    cf = P(0.0, 0.0, z0)
    cb = P(0.0, 0.0, z1)
    tris = []
    for i in range(3):
        j = (i + 1) % 3
        tris.append((cf, f[j], f[i]))     # front cap faces -z
        tris.append((cb, b[i], b[j]))     # back cap faces +z
        tris.append((f[i], f[j], b[j]))   # side quads
        tris.append((f[i], b[j], b[i]))
    return np.array([a + bb + cc for a, bb, cc in tris], np.float32)


def prism_json():
    mats = [
        {"name": "prism glass", "kind": "Dispersive", "cauchy_a": 1.5046, "cauchy_b": 0.00420},
        {"name": "sphere glass", "kind": "Dielectric", "refractive_index": 1.5},
        {"name": "floor", "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": {"lin_srgb": [0.5, 0.5, 0.5]}}},
        {"name": "lamp", "kind": "DiffuseLight", "emit": {"kind": "SolidColor", "color": {"xyz_e": [500, 500, 500]}}},
    ]
    objs = [
        {"kind": "Quad", "q": [-6.0, 0.0, -6.0], "u": [0.0, 0.0, 12.0], "v": [12.0, 0.0, 0.0], "material": "floor"},
        {"kind": "Sphere", "priority": True, "center": [1.3, 0.5, -0.6], "radius": 0.5, "material": "sphere glass"},
        {"kind": "Quad", "priority": True, "q": [-2.05, 3.0, -0.05], "u": [0.1, 0.0, 0.0], "v": [0.0, 0.0, 0.1],
         "material": "lamp"},
        {"kind": "Quad", "priority": True, "q": [1.95, 3.5, 1.45], "u": [0.1, 0.0, 0.0], "v": [0.0, 0.0, 0.1],
         "material": "lamp"},
    ]
    return {
        "time_0": 0, "time_1": 1,
        "camera": {"look_from": [0.0, 3.2, 5.5], "look_at": [0.0, 0.4, 0.0], "up": [0.0, 1.0, 0.0],
                   "vertical_fov": 40.0, "aperture": 0.0, "focus_distance": 10.0},
        "background_color": [0, 0, 0],
        "objects": objs,
        "materials": mats,
    }


def prism():
    s = load_scene(prism_json())
    glass = 0  # "prism glass" is the first shared material
    s.add_objects(abi.OBJ_TRIANGLE, prism_triangles(), glass, priority=True)
    return s


# -------------------------------------------------------------------- mesh
def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return v, f


def geodesic_icosphere(freq):
    """Class-I geodesic sphere: every icosahedron face split into freq^2
    triangles; vertices shared exactly between faces (edge points are computed
    from the edge's sorted endpoints).  Returns (vertices f64 [V,3] on the unit
    sphere, triangles int [20*freq^2, 3], outward winding)."""
    V0, F = _icosahedron()
    n = freq
    verts = [V0]
    nv = len(V0)
    edge_ids = {}
    # edge interior points
    for fa in F:
        for k in range(3):
            a, b = int(fa[k]), int(fa[(k + 1) % 3])
            key = (min(a, b), max(a, b))
            if key in edge_ids:
                continue
            lo, hi = key
            t = np.arange(1, n, dtype=np.float64)[:, None] / n
            pts = V0[lo] * (1.0 - t) + V0[hi] * t
            edge_ids[key] = nv
            verts.append(pts)
            nv += n - 1
    tris = []
    for fa in F:
        A, B, Cc = (int(x) for x in fa)
        idx = -np.ones((n + 1, n + 1), np.int64)  # idx[i, j]: point A + i/n (B-A) + j/n (C-A)

        def edge_index(p, q, k):  # k-th point from p towards q, 0<k<n
            lo, hi = min(p, q), max(p, q)
            base = edge_ids[(lo, hi)]
            kk = k if p == lo else n - k
            return base + kk - 1
        idx[0, 0], idx[n, 0], idx[0, n] = A, B, Cc
        for k in range(1, n):
            idx[k, 0] = edge_index(A, B, k)
            idx[0, k] = edge_index(A, Cc, k)
            idx[n - k, k] = edge_index(B, Cc, k)
        # interior
        ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
        mask = (ii > 0) & (jj > 0) & (ii + jj < n)
        cnt = int(mask.sum())
        if cnt:
            ni, nj = ii[mask], jj[mask]
            pts = (V0[A][None, :] * ((n - ni - nj) / n)[:, None] + V0[B][None, :] * (ni / n)[:, None]
                   + V0[Cc][None, :] * (nj / n)[:, None])
            idx[ni, nj] = nv + np.arange(cnt)
            verts.append(pts)
            nv += cnt
        # triangles (winding of the face A,B,C is outward)
        i_up, j_up = np.nonzero((ii + jj < n))
        a = idx[i_up, j_up]
        b = idx[i_up + 1, j_up]
        c = idx[i_up, j_up + 1]
        tris.append(np.stack([a, b, c], 1))
        i_dn, j_dn = np.nonzero((ii + jj < n - 1))
        a = idx[i_dn + 1, j_dn]
        b = idx[i_dn + 1, j_dn + 1]
        c = idx[i_dn, j_dn + 1]
        tris.append(np.stack([a, b, c], 1))
    V = np.concatenate(verts)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    T = np.concatenate(tris)
    return V, T


def value_noise(p, seed=1, octaves=4, period=256):
    """Fractal 3-D value noise in [-1, 1]: random lattice values (seeded PCG64),
    trilinear interpolation with a smoothstep fade; each octave doubles the
    frequency and halves the amplitude; normalised by the amplitude sum."""
    lat = _rng(seed).uniform(-1.0, 1.0, size=(period, period, period))
    out = np.zeros(len(p))
    amp, tot, freq = 1.0, 0.0, 2.0
    for _ in range(octaves):
        q = p * freq + 17.0
        i0 = np.floor(q).astype(np.int64)
        f = q - i0
        f = f * f * (3.0 - 2.0 * f)
        acc = 0.0
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    w = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * \
                        (f[:, 2] if dz else 1 - f[:, 2])
                    acc = acc + w * lat[(i0[:, 0] + dx) % period, (i0[:, 1] + dy) % period, (i0[:, 2] + dz) % period]
        out += amp * acc
        tot += amp
        amp *= 0.5
        freq *= 2.0
    return out / tot


def displaced_icosphere(freq=224, amplitude=0.15, seed=1, radius=1.0, center=(0.0, 1.0, 0.0)):
    V, T = geodesic_icosphere(freq)
    d = value_noise(V, seed=seed)
    P = V * (radius * (1.0 + amplitude * d))[:, None] + np.asarray(center)[None, :]
    P32 = P.astype(np.float32)
    return P32, T


MESH_VIEWS = [(0.0, 1.0, 4.2), (4.2, 1.0, 0.0), (0.0, 1.0, -4.2), (-4.2, 1.0, 0.0)]


def mesh(freq=224, seed=1, view=0):
    """Thin-film dielectric blob (dielectric.rs base n=1.5, ThinFilm n=1.33,
    thickness U[250,400] nm, one value per mesh) under a quad light, over a
    grey floor, against a sky background.  `view` selects one of the four
    camera azimuths 0/90/180/270 degrees."""
    r = _rng(seed + 1000)
    thickness = float(r.uniform(250.0, 400.0))
    lf = MESH_VIEWS[view]
    sf = {
        "time_0": 0, "time_1": 1,
        "camera": {"look_from": list(lf), "look_at": [0.0, 1.0, 0.0], "up": [0.0, 1.0, 0.0],
                   "vertical_fov": 40.0, "aperture": 0.0, "focus_distance": 10.0},
        "background_color": [0.7, 0.7, 0.9],
        "objects": [
            {"kind": "Quad", "q": [-20.0, -0.2, -20.0], "u": [0.0, 0.0, 40.0], "v": [40.0, 0.0, 0.0],
             "material": "floor"},
            {"kind": "Quad", "priority": True, "q": [-1.0, 4.0, -1.0], "u": [2.0, 0.0, 0.0], "v": [0.0, 0.0, 2.0],
             "material": "lamp"},
        ],
        "materials": [
            {"name": "film", "kind": "Dielectric", "refractive_index": 1.5,
             "thin_film": {"refractive_index": 1.33, "thickness": thickness}},
            {"name": "floor", "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": [0.6, 0.6, 0.6]}},
            {"name": "lamp", "kind": "DiffuseLight", "emit": {"kind": "IlluminantD65", "intensity": 0.1}},
        ],
    }
    s = load_scene(sf)
    P, T = displaced_icosphere(freq=freq, seed=seed)
    tri = np.concatenate([P[T[:, 0]], P[T[:, 1]], P[T[:, 2]]], axis=1)
    s.add_objects(abi.OBJ_TRIANGLE, tri, 0)
    return s


# ------------------------------------------------------------ uneven batch
def batch_scene(seed, n_tris=None, n_spheres=None):
    """One scene of the uneven batch: a noise-displaced geodesic icosphere of
    about `n_tris` triangles (20 f^2 for the nearest frequency f >= 7) with a
    random material (Lambertian, metal, dielectric or thin-film dielectric),
    `n_spheres` random spheres with the sphere field's material mix, a grey
    floor, a quad light (the MIS target) and a sky.  Seeded; returns
    (scene, triangles)."""
    r = _rng(10_000 + seed)
    U = lambda: float(r.random())
    if n_tris is None:
        n_tris = int(round(math.exp(r.uniform(math.log(1e3), math.log(5e5)))))
    if n_spheres is None:
        n_spheres = int(r.integers(0, 201))
    freq = max(7, int(round(math.sqrt(n_tris / 20.0))))
    kind = int(r.integers(0, 4))
    col = [0.2 + 0.7 * U(), 0.2 + 0.7 * U(), 0.2 + 0.7 * U()]
    blob = [{"name": "blob", "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": col}},
            {"name": "blob", "kind": "Metal", "fuzz": 0.3 * U(), "albedo": {"kind": "SolidColor", "color": col}},
            {"name": "blob", "kind": "Dielectric", "refractive_index": 1.5},
            {"name": "blob", "kind": "Dielectric", "refractive_index": 1.5,
             "thin_film": {"refractive_index": 1.33, "thickness": float(r.uniform(250.0, 400.0))}}][kind]
    mats = [blob,
            {"name": "floor", "kind": "Lambertian", "albedo": {"kind": "SolidColor", "color": [0.5, 0.5, 0.5]}},
            {"name": "lamp", "kind": "DiffuseLight", "emit": {"kind": "IlluminantD65", "intensity": 0.1}},
            {"name": "glass", "kind": "Dielectric", "refractive_index": 1.5}]
    objs = [{"kind": "Quad", "q": [-20.0, -0.2, -20.0], "u": [0.0, 0.0, 40.0], "v": [40.0, 0.0, 0.0],
             "material": "floor"},
            {"kind": "Quad", "priority": True, "q": [-1.0, 4.0, -1.0], "u": [2.0, 0.0, 0.0], "v": [0.0, 0.0, 2.0],
             "material": "lamp"}]
    for k in range(n_spheres):
        ang, rad = 2 * math.pi * U(), 1.4 + 3.0 * U()
        c = [rad * math.cos(ang), 0.0, rad * math.sin(ang)]
        rr = 0.05 + 0.15 * U()
        c[1] = -0.2 + rr
        choose = U()
        name = "s%d" % k
        if choose < 0.8:
            mats.append({"name": name, "kind": "Lambertian",
                         "albedo": {"kind": "SolidColor", "color": [U() * U(), U() * U(), U() * U()]}})
        elif choose < 0.95:
            mats.append({"name": name, "kind": "Metal", "fuzz": 0.5 * U(),
                         "albedo": {"kind": "SolidColor", "color": [0.5 + 0.5 * U(), 0.5 + 0.5 * U(), 0.5 + 0.5 * U()]}})
        else:
            name = "glass"
        objs.append({"kind": "Sphere", "center": c, "radius": rr, "material": name})
    az = 2 * math.pi * U()
    sf = {
        "time_0": 0, "time_1": 1,
        "camera": {"look_from": [4.2 * math.sin(az), 1.4, 4.2 * math.cos(az)], "look_at": [0.0, 0.9, 0.0],
                   "up": [0.0, 1.0, 0.0], "vertical_fov": 45.0, "aperture": 0.0, "focus_distance": 10.0},
        "background_color": [0.7, 0.7, 0.9],
        "objects": objs,
        "materials": mats,
    }
    s = load_scene(sf)
    P, T = displaced_icosphere(freq=freq, seed=seed)
    tri = np.concatenate([P[T[:, 0]], P[T[:, 1]], P[T[:, 2]]], axis=1)
    s.add_objects(abi.OBJ_TRIANGLE, tri, 0)
    return s, int(T.shape[0])


def uneven_batch(count=64, seed=1):
    """The uneven batch: `count` seeded scenes (batch_scene(seed + k)) with
    square resolutions log-uniform in 256..1024 pixels a side and samples
    per pixel log-uniform in 32..512.  Returns [(scene, triangles, side, spp)]."""
    r = _rng(20_000 + seed)
    out = []
    for k in range(count):
        side = int(round(math.exp(r.uniform(math.log(256), math.log(1024)))))
        spp = int(round(math.exp(r.uniform(math.log(32), math.log(512)))))
        scene, tris = batch_scene(seed + k)
        out.append((scene, tris, side, spp))
    return out


SCENES = {"cornell": cornell, "spheres": spheres, "prism": prism, "mesh": mesh}

# The configurations the measurements are quoted on: (scene, W, H, spp, max_depth)
CONFIGS = {
    "cornell": dict(scene="cornell", width=256, height=256, samples=16, max_depth=8),
    "spheres": dict(scene="spheres", width=1920, height=1080, samples=64, max_depth=16),
    "mesh": dict(scene="mesh", width=1920, height=1080, samples=256, max_depth=12, views=4),
    "prism": dict(scene="prism", width=1024, height=1024, samples=0, max_depth=24),
    "batch64": dict(scene="uneven_batch", count=64, max_depth=12),
}
