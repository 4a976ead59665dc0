"""Scene files -> the flat scene descriptor of the C ABI.

Restates the reference's scene assembly, clovers-cli/src/scenefile.rs:21-117
(SceneFile schema, flattening of top-level lists, the priority list, the
default material appended last, background conversion) and
clovers/src/objects.rs:59-188 (Object enum -> hitable, shared material lookup
by name with fallback to the default).  The JSON schema is accepted unchanged:
internally tagged ``kind`` for objects, materials and textures
(objects.rs:58, materials.rs:139, textures.rs:21), untagged MaterialInit
(a shared-material name or an inline material, materials.rs:33-39) and
untagged ColorInit (colorinit.rs:17-43).

Out of scope (raise NotImplementedError): GLTF files, constant media with a
boundary other than a Sphere or a Boxy, and the illuminants other than
D65/D50.  A list nested inside
RotateY/Translate is rejected too: the reference's HitableList::hit returns
None as soon as any child misses (hitable.rs:158), which no scene relies on.
STL and PLY meshes are read by clovers_amd.meshfile.
"""
import json

import numpy as np

from . import abi
from .color import color_init_to_xyz_e, srgb_to_xyz_e
from .meshfile import mesh_triangles


class SceneDescription:
    """Host-side scene: numpy arrays laid out exactly as the C ABI structs."""

    def __init__(self):
        self._objects = []          # list of np arrays (OBJECT_DTYPE) to concatenate
        self.materials = []         # list of MATERIAL_DTYPE scalars
        self.xforms = []            # list of XFORM_DTYPE scalars
        self.camera = None
        self.background = np.zeros(3, np.float32)
        self.mis_bvh_algorithm = abi.BVH_SAH
        self._obj_cache = None

    # -- building -------------------------------------------------------
    def add_material(self, rec):
        self.materials.append(rec)
        return len(self.materials) - 1

    def add_xform(self, ops):
        if not ops:
            return abi.NO_XFORM
        if len(ops) > abi.MAX_XFORM_OPS:
            raise NotImplementedError("more than %d nested RotateY/Translate wrappers" % abi.MAX_XFORM_OPS)
        x = np.zeros((), abi.XFORM_DTYPE)
        x["n_ops"] = len(ops)
        for i, (kind, v) in enumerate(ops):
            x["ops"][i]["kind"] = kind
            x["ops"][i]["v"] = v
        self.xforms.append(x)
        return len(self.xforms) - 1

    def add_object(self, kind, params, material, priority=False, xform=abi.NO_XFORM):
        o = np.zeros(1, abi.OBJECT_DTYPE)
        o["kind"] = kind
        o["material"] = material
        o["priority"] = 1 if priority else 0
        o["xform"] = xform
        p = np.zeros(9, np.float32)
        p[:len(params)] = params
        o["p"] = p
        self._objects.append(o)
        self._obj_cache = None

    def add_objects(self, kind, params, material, priority=False):
        """Bulk add (e.g. a triangle mesh, one object per element as the
        reference flattens mesh HitableLists, scenefile.rs:84-94)."""
        params = np.asarray(params, np.float32)
        n = params.shape[0]
        o = np.zeros(n, abi.OBJECT_DTYPE)
        o["kind"] = kind
        o["material"] = material
        o["priority"] = 1 if priority else 0
        o["xform"] = abi.NO_XFORM
        o["p"][:, :params.shape[1]] = params
        self._objects.append(o)
        self._obj_cache = None

    @property
    def objects(self):
        if self._obj_cache is None:
            self._obj_cache = (np.concatenate(self._objects) if self._objects
                               else np.zeros(0, abi.OBJECT_DTYPE))
        return self._obj_cache

    def set_camera(self, look_from, look_at, up, vertical_fov, aperture, focus_distance, time_0=0.0, time_1=1.0):
        self.camera = dict(look_from=list(map(float, look_from)), look_at=list(map(float, look_at)),
                           up=list(map(float, up)), vertical_fov=float(vertical_fov), aperture=float(aperture),
                           focus_distance=float(focus_distance), time_0=float(time_0), time_1=float(time_1))

    # -- C ABI ---------------------------------------------------------------
    def as_ctypes(self):
        """Return (clv_scene_desc, keepalive) -- keep `keepalive` referenced
        while the descriptor is in use."""
        objs = np.ascontiguousarray(self.objects)
        mats = np.ascontiguousarray(np.array(self.materials, dtype=abi.MATERIAL_DTYPE).reshape(-1))
        xfs = np.ascontiguousarray(np.array(self.xforms, dtype=abi.XFORM_DTYPE).reshape(-1)) \
            if self.xforms else np.zeros(1, abi.XFORM_DTYPE)
        d = abi.SceneDesc()
        d.abi_version = abi.ABI_VERSION
        d.n_objects = len(objs)
        d.objects = objs.ctypes.data
        d.n_materials = len(mats)
        d.materials = mats.ctypes.data
        d.n_xforms = len(self.xforms)
        d.xforms = xfs.ctypes.data
        c = self.camera
        cam = abi.CameraInit()
        cam.look_from[:] = c["look_from"]
        cam.look_at[:] = c["look_at"]
        cam.up[:] = c["up"]
        cam.vertical_fov = c["vertical_fov"]
        cam.aperture = c["aperture"]
        cam.focus_distance = c["focus_distance"]
        cam.time_0 = c["time_0"]
        cam.time_1 = c["time_1"]
        d.camera = cam
        d.background[:] = [float(v) for v in self.background]
        d.mis_bvh_algorithm = self.mis_bvh_algorithm
        return d, (objs, mats, xfs)

    @property
    def n_priority(self):
        return int(self.objects["priority"].sum())


# ----------------------------------------------------------------- materials
def _texture(t, default_xyz):
    """Texture JSON -> MATERIAL_DTYPE fields (texture kind, Xyz<E> colours,
    intensity, checker density). textures.rs:22-109."""
    if t is None:  # Texture::default() = SolidColor::default() = Xyz<E>(0.5,0.5,0.5) (solid_color.rs:50-56)
        return dict(texture=abi.TEX_SOLID, color=np.array(default_xyz, np.float32))
    kind = t.get("kind")
    if kind == "SolidColor":
        return dict(texture=abi.TEX_SOLID, color=color_init_to_xyz_e(t["color"]))
    if kind == "IlluminantD65":
        return dict(texture=abi.TEX_ILLUMINANT_D65, color=np.zeros(3, np.float32), intensity=float(t.get("intensity", 1.0)))
    if kind == "IlluminantD50":
        return dict(texture=abi.TEX_ILLUMINANT_D50, color=np.zeros(3, np.float32), intensity=float(t.get("intensity", 1.0)))
    if kind in ("SpatialChecker", "SurfaceChecker"):
        # defaults: even Xyz<E>(0.8, 0.8, 0.8), odd Xyz<E>(0.5, 0.5, 0.5), density 1 (spatial) / 10 (surface)
        # (spatial_checker.rs:46-61, surface_checker.rs:44-59)
        even = color_init_to_xyz_e(t["even"]) if "even" in t else np.full(3, 0.8, np.float32)
        odd = color_init_to_xyz_e(t["odd"]) if "odd" in t else np.full(3, 0.5, np.float32)
        spatial = kind == "SpatialChecker"
        return dict(texture=abi.TEX_SPATIAL_CHECKER if spatial else abi.TEX_SURFACE_CHECKER, color=even,
                    odd_color=odd, density=float(t.get("density", 1.0 if spatial else 10.0)))
    raise NotImplementedError("texture kind %r is not supported" % kind)


def _set(r, fields):
    for k, v in fields.items():
        r[k] = v


def material_record(m):
    """Material JSON (kind + fields, optional thin_film) -> MATERIAL_DTYPE. materials.rs:63-103."""
    r = np.zeros((), abi.MATERIAL_DTYPE)
    r["intensity"] = 1.0
    r["refractive_index"] = 1.5     # dielectric.rs:27-29
    r["cauchy_a"] = 1.5046          # dispersive.rs:39-45
    r["cauchy_b"] = 0.00420
    r["film_index"] = 1.5           # thin_film.rs:57-63
    r["film_thickness"] = 500.0
    r["color"] = (1.0, 1.0, 1.0)
    kind = m.get("kind", "Lambertian") if m else "Lambertian"
    if kind == "Lambertian":
        r["kind"] = abi.MAT_LAMBERTIAN
        _set(r, _texture(m.get("albedo") if m else None, (0.5, 0.5, 0.5)))
    elif kind == "Metal":
        r["kind"] = abi.MAT_METAL
        _set(r, _texture(m.get("albedo"), (0.5, 0.5, 0.5)))
        r["fuzz"] = float(m.get("fuzz", 0.0))   # not clamped on the serde path (metal.rs:60)
    elif kind == "Dielectric":
        r["kind"] = abi.MAT_DIELECTRIC
        r["refractive_index"] = float(m.get("refractive_index", 1.5))
        c = m.get("color", [1.0, 1.0, 1.0])    # Xyz<E> directly (dielectric.rs:19-23)
        r["color"] = [float(v) for v in (c if isinstance(c, list) else (c["x"], c["y"], c["z"]))]
    elif kind == "Dispersive":
        r["kind"] = abi.MAT_DISPERSIVE
        r["cauchy_a"] = float(m.get("cauchy_a", 1.5046))
        r["cauchy_b"] = float(m.get("cauchy_b", 0.00420))
    elif kind == "DiffuseLight":
        r["kind"] = abi.MAT_DIFFUSE_LIGHT
        # DiffuseLight::default(): SolidColor Xyz(100,100,100) (diffuse_light.rs:23-30)
        _set(r, _texture(m.get("emit"), (100.0, 100.0, 100.0)))
    elif kind == "Isotropic":   # isotropic.rs:13-19: albedo texture, default grey
        r["kind"] = abi.MAT_ISOTROPIC
        _set(r, _texture(m.get("albedo"), (0.5, 0.5, 0.5)))
        if r["texture"] == abi.TEX_SURFACE_CHECKER:
            raise NotImplementedError("a SurfaceChecker texture on an Isotropic material")
    elif kind == "ConeLight":
        # both fields are required by the scene-file schema (cone_light.rs:13-18: no serde defaults)
        if "spread" not in m or "emit" not in m:
            raise ValueError("ConeLight needs 'spread' and 'emit'")
        r["kind"] = abi.MAT_CONE_LIGHT
        r["spread"] = float(m["spread"])
        _set(r, _texture(m["emit"], (100.0, 100.0, 100.0)))
    else:
        raise NotImplementedError("material kind %r is not supported" % kind)
    tf = m.get("thin_film") if m else None
    if tf is not None:
        r["thin_film"] = 1
        r["film_index"] = float(tf.get("refractive_index", 1.5))
        r["film_thickness"] = float(tf.get("thickness", 500.0))
    return r


# ------------------------------------------------------------------- objects
_PRIM_KINDS = ("Sphere", "Quad", "Triangle", "Boxy", "MovingSphere")


def _leaf_params(o):
    k = o["kind"]
    if k == "Sphere":
        return abi.OBJ_SPHERE, list(o["center"]) + [o["radius"]]
    if k == "Quad":
        return abi.OBJ_QUAD, list(o["q"]) + list(o["u"]) + list(o["v"])
    if k == "Triangle":
        return abi.OBJ_TRIANGLE_QUV, list(o["q"]) + list(o["u"]) + list(o["v"])
    if k == "Boxy":
        return abi.OBJ_BOXY, list(o["corner_0"]) + list(o["corner_1"])
    if k == "MovingSphere":  # time_0/time_1 of the file are not read: the object moves over [0, 1] (objects.rs:104-109)
        return abi.OBJ_MOVING_SPHERE, list(o["center_0"]) + list(o["center_1"]) + [o["radius"]]
    raise NotImplementedError("object kind %r is not supported" % k)


class _Loader:
    def __init__(self, scene, shared, mesh_root=None):
        self.scene = scene
        self.shared = shared   # list of (name, material index); the default (name "") last
        self.mesh_root = mesh_root

    def material(self, init):
        """MaterialInit -> material index (objects.rs:159-188)."""
        if init is None:
            init = ""
        if isinstance(init, str):
            for name, idx in self.shared:
                if name == init:
                    return idx
            for name, idx in self.shared:
                if name == "":
                    return idx
            raise ValueError("no default material")
        return self.scene.add_material(material_record(init))

    def wrapped(self, o, ops):
        """Unwrap RotateY/Translate chains (outermost first) down to a primitive."""
        k = o["kind"]
        if k == "Translate":
            return self.wrapped(o["object"], ops + [(abi.XF_TRANSLATE, list(o["offset"]))])
        if k == "RotateY":
            return self.wrapped(o["object"], ops + [(abi.XF_ROTATE_Y, [float(o["angle"]), 0.0, 0.0])])
        if k in ("ObjectList", "STL", "PLY"):
            if ops:
                raise NotImplementedError("a list inside RotateY/Translate (hitable.rs:158 returns None on any miss)")
            return [("list" if k == "ObjectList" else "mesh", o)]
        if k == "ConstantMedium":
            return [("medium", o, ops)]
        if k not in _PRIM_KINDS:
            raise NotImplementedError("object kind %r is not supported" % k)
        return [("prim", o, ops)]

    def medium(self, o, priority, ops):
        """ConstantMedium (objects.rs:99-103, constant_medium.rs:19-46): a
        Sphere or Boxy boundary, whose own RotateY/Translate wrappers move onto
        the medium (the same hit along the same ray); density (default 0.1);
        texture (default SolidColor grey) as the Isotropic phase function.  The
        boundary's material is never read."""
        if priority:
            raise ValueError("a ConstantMedium cannot be a priority object (its random() is unimplemented: "
                             "the reference panics)")
        b = o["boundary"]
        ops = list(ops)
        while b["kind"] in ("Translate", "RotateY"):
            if b["kind"] == "Translate":
                ops.append((abi.XF_TRANSLATE, list(b["offset"])))
            else:
                ops.append((abi.XF_ROTATE_Y, [float(b["angle"]), 0.0, 0.0]))
            b = b["object"]
        density = float(o.get("density", 0.1))
        if b["kind"] == "Sphere":
            kind, params = abi.OBJ_MEDIUM_SPHERE, list(b["center"]) + [b["radius"], density]
        elif b["kind"] == "Boxy":
            kind, params = abi.OBJ_MEDIUM_BOXY, list(b["corner_0"]) + list(b["corner_1"]) + [density]
        else:
            raise NotImplementedError("a ConstantMedium boundary of kind %r is not supported" % b["kind"])
        r = material_record({"kind": "Isotropic", "albedo": o.get("texture")})
        mat = self.scene.add_material(r)
        self.scene.add_object(kind, params, mat, False, self.scene.add_xform(ops))

    def mesh_material(self, init):
        """STL/PLY material: a shared name must exist (stl.rs:50, ply.rs:43
        unwrap the lookup: no fallback to the default); the default
        MaterialInit is the name "" of the appended default material."""
        if init is None:
            init = ""
        if isinstance(init, str):
            for name, idx in self.shared:
                if name == init:
                    return idx
            raise ValueError("mesh material %r not found (the reference panics)" % init)
        return self.scene.add_material(material_record(init))

    def add(self, o, priority):
        for entry in self.wrapped(o, []):
            if entry[0] == "list":
                # top-level list: flatten one level (scenefile.rs:84-94)
                for child in entry[1]["objects"]:
                    if child["kind"] in ("ObjectList", "STL", "PLY"):
                        raise ValueError("Recursing too hard in scene initialization")
                    self.add(child, priority)
                continue
            if entry[0] == "mesh":
                # one Triangle::from_coordinates per face, flattened (objects.rs:133-141)
                tris = mesh_triangles(entry[1], self.mesh_root)
                mat = self.mesh_material(entry[1].get("material"))
                self.scene.add_objects(abi.OBJ_TRIANGLE, tris.reshape(-1, 9), mat, priority)
                continue
            if entry[0] == "medium":
                self.medium(entry[1], priority, entry[2])
                continue
            _, prim, ops = entry
            kind, params = _leaf_params(prim)
            mat = self.material(prim.get("material"))
            xf = self.scene.add_xform(ops)
            self.scene.add_object(kind, params, mat, priority, xf)


def load_scene(src, bvh_algorithm=abi.BVH_SAH, mesh_root=None):
    """Scene file (path, JSON text or parsed dict) -> SceneDescription.
    Mirrors SceneFile::initialize (scenefile.rs:34-117).  Relative STL/PLY
    paths are read from `mesh_root` (default: the working directory, as the
    reference does)."""
    if isinstance(src, dict):
        sf = src
    elif isinstance(src, str) and src.lstrip().startswith("{"):
        sf = json.loads(src)
    else:
        with open(src) as f:
            sf = json.load(f)
    s = SceneDescription()
    s.mis_bvh_algorithm = bvh_algorithm
    bg = sf.get("background_color", [0.0, 0.0, 0.0])
    if isinstance(bg, dict):
        bg = [bg["red"], bg["green"], bg["blue"]]
    s.background = np.asarray(srgb_to_xyz_e(bg), np.float32)    # scenefile.rs:43-44
    cam = sf["camera"]
    s.set_camera(cam["look_from"], cam["look_at"], cam["up"], cam["vertical_fov"], cam["aperture"],
                 cam["focus_distance"], sf.get("time_0", 0.0), sf.get("time_1", 1.0))
    shared = []
    for m in sf.get("materials", []):
        shared.append((m.get("name", ""), s.add_material(material_record(m))))
    shared.append(("", s.add_material(material_record(None))))   # SharedMaterial::default() (scenefile.rs:58-59)
    ld = _Loader(s, shared, mesh_root)
    for o in sf.get("objects", []):
        ld.add(o, bool(o.get("priority", False)))
    return s
