"""The reference renderer's scene files for the Python API: Wavefront OBJ
meshes with their MTL materials and the .obejota camera / point-light files,
read into a SceneDesc.  The same rules as the C++ loaders of include/rt_io.hpp:

- loadMTL      cuda_libraries/FileImporting.cu:27-246
- loadObj      cuda_libraries/FileImporting.cu:248-372
- loadObejota  cuda_libraries/FileImporting.cu:375-598 (storeObejotaObject)

OBJ `v` / `vn` lines are read as (x, z, y) (the file's second and third
coordinates swapped); .obejota `p`, `lookat`, `cts` and `up` have their y
negated, and `rotation ax ay az` (degrees) sets lookat = p + Rz(-az) Ry(-ay)
Rx(-ax) (0, 0, -1) and rotates the camera's up by the same matrix.  `newmtl`
does not reset Ke, sigmaS, sigmaA or g; an mtllib's materials never replace
a name already known; every `o` line advances a mesh counter that starts at
-1.  Malformed input raises ValueError with the file and line; a material
that the GPU path does not render (Phong, Beckmann, subsurface, textures,
anisotropic roughness, shadow-transparent Lambertian or Plain) raises
UnsupportedMaterial when a scene is made from it."""
from __future__ import annotations

import math
import os
import sys
from dataclasses import dataclass, field

import numpy as np

from .scenes import SceneDesc

ILLUM_CLASSES = ("PlainMaterial", "LambertianMaterial", "PhongMaterial", "DielectricMaterial", "BeckmannMaterial",
                 "TrowbridgeMaterial", "TrowbridgeBasicMaterial", "SSMaterial")
# rt_material_type values of include/rt.h
_LAMBERTIAN, _DIELECTRIC, _TROWBRIDGE, _TROWBRIDGE_BASIC, _PLAIN = 0, 1, 2, 3, 6


class UnsupportedMaterial(ValueError):
    """An MTL material whose class the GPU path does not render."""


def _f32(x):
    return float(np.float32(x))


@dataclass
class MtlMaterial:
    name: str = ""
    illum: int = 0
    ka: tuple = (0.0, 0.0, 0.0)
    kd: tuple = (1.0, 1.0, 1.0)
    ks: tuple = (0.0, 0.0, 0.0)
    ke: tuple = (0.0, 0.0, 0.0)
    ni: float = 1.0
    ns: float = 0.0
    rough_x: float = 0.0
    rough_y: float = 0.0
    eta: tuple = (1.0, 1.0, 1.0)
    k: tuple = (0.0, 0.0, 0.0)
    transparent: bool = False     # makesShadow 0
    has_indirect: bool = True     # hasIndirect
    sigma_s: tuple = (0.0, 0.0, 0.0)
    sigma_a: tuple = (0.0, 0.0, 0.0)
    g: float = 0.0
    map_kd: str = ""              # texture path (the MTL's directory, else its textures/ subdirectory)

    def copy(self):
        return MtlMaterial(**self.__dict__)


@dataclass
class ObjData:
    vertices: np.ndarray          # (nv, 3) float32, (x, z, y) of the file
    normals: np.ndarray           # (nn, 3) float32, (x, z, y) of the file
    texcoords: np.ndarray         # (nt, 2) float32
    v: np.ndarray                 # (ntri, 3) int32, 0-based
    vt: np.ndarray                # (ntri, 3) int32, -1 where absent
    vn: np.ndarray                # (ntri, 3) int32, -1 where absent
    mesh_id: np.ndarray           # (ntri,) int32
    material: list                # material name per triangle
    materials: dict               # name -> MtlMaterial ("default" and every mtllib's)
    next_mesh_id: int = -1


@dataclass
class ObejotaScene:
    camera: dict | None = None    # pos, lookat, up, cts, frame_y (the file's last camera)
    lights: list = field(default_factory=list)     # (pos, color, intensity) in file order
    warnings: list = field(default_factory=list)   # also printed to stderr, as the reference does


def _where(path, ln):
    return f"{path}:{ln}: "


def _nums(tok, n, path, ln, key):
    if len(tok) < n:
        raise ValueError(_where(path, ln) + f"expected a number after '{key}'")
    try:
        return tuple(_f32(t) for t in tok[:n])
    except ValueError:
        raise ValueError(_where(path, ln) + f"expected a number after '{key}'") from None


def _flag01(tok, path, ln, key):
    t = tok[0] if tok else ""
    if t == "1":
        return True
    if t == "0":
        return False
    raise ValueError(_where(path, ln) + f"{key} must be 0 or 1")


def load_mtl(path: str) -> dict:
    """loadMTL: name -> MtlMaterial.  A file that cannot be opened gives an
    empty dict and a message on stderr, as in the reference."""
    if not os.path.isfile(path):
        print(f"MTL file not found: {path}", file=sys.stderr)
        return {}
    out = {}
    cur = MtlMaterial()
    opened = False
    rgb = {"Ka": "ka", "Kd": "kd", "Ks": "ks", "Ke": "ke", "eta": "eta", "k": "k", "sigmaS": "sigma_s",
           "sigmaA": "sigma_a"}
    scalar = {"Ns": "ns", "Ni": "ni", "PrX": "rough_x", "PrY": "rough_y", "g": "g"}
    with open(path) as f:
        for ln, line in enumerate(f, 1):
            tok = line.split()
            if not tok:
                continue
            key, rest = tok[0], tok[1:]
            if key == "newmtl":
                if opened and cur.name:
                    out[cur.name] = cur.copy()
                cur.ka, cur.kd, cur.ks = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
                cur.ni, cur.ns, cur.illum, cur.map_kd = 1.0, 0.0, 0, ""
                cur.transparent, cur.has_indirect = False, True
                cur.rough_x = cur.rough_y = 0.0
                cur.eta, cur.k = (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
                if rest:  # a missing name keeps the previous one, as in the reference
                    cur.name = rest[0]
                opened = True
            elif key in rgb:
                setattr(cur, rgb[key], _nums(rest, 3, path, ln, key))
            elif key in scalar:
                setattr(cur, scalar[key], _nums(rest, 1, path, ln, key)[0])
            elif key == "Pr":
                cur.rough_x = cur.rough_y = _nums(rest, 1, path, ln, key)[0]
            elif key == "illum":
                try:
                    v = int(rest[0])
                except (IndexError, ValueError):
                    raise ValueError(_where(path, ln) + "expected an integer after 'illum'") from None
                if not 0 <= v <= 7:
                    raise ValueError(_where(path, ln) + "illum must be 0..7")
                cur.illum = v
            elif key == "makesShadow":
                cur.transparent = not _flag01(rest, path, ln, key)
            elif key == "hasIndirect":
                cur.has_indirect = _flag01(rest, path, ln, key)
            elif key == "map_Kd":
                # Texture(dir + name), else Texture(dir + "textures/" + name)
                # (FileImporting.cu:213-235): the file tried is texture_file_name's
                name = rest[0] if rest else ""
                d = os.path.dirname(path)
                p = os.path.join(d, name)
                cur.map_kd = p if os.path.isfile(texture_file_name(p, quiet=True)) else os.path.join(d, "textures", name)
    if opened and cur.name:
        out[cur.name] = cur.copy()
    return out


def _face_ref(ref, path, ln):
    parts = ref.split("/")
    if len(parts) > 3 or parts[0] == "":
        raise ValueError(_where(path, ln) + f"bad face vertex '{ref}'")
    out = []
    for p in parts + [""] * (3 - len(parts)):
        if p == "":
            out.append(-1)
            continue
        if not p.isdigit() or int(p) < 1 or int(p) > 2 ** 31 - 1:
            raise ValueError(_where(path, ln) + f"bad face index '{ref}'")
        out.append(int(p) - 1)
    return out


def load_obj(path: str, first_mesh_id: int = -1) -> ObjData:
    """loadObj.  first_mesh_id is the mesh counter before this file (-1 for
    the first file, then the previous file's next_mesh_id)."""
    if not os.path.isfile(path):
        raise ValueError(f"failed to open the OBJ file {path}")
    verts, norms, tex = [], [], []
    v, vt, vn, mids, mats = [], [], [], [], []
    materials = {"default": MtlMaterial(name="default", illum=1)}
    mesh_id = first_mesh_id
    material = "default"
    with open(path) as f:
        for ln, line in enumerate(f, 1):
            tok = line.split()
            if not tok:
                continue
            key, rest = tok[0], tok[1:]
            if key == "o":
                mesh_id += 1
            elif key in ("v", "vn"):
                x, z, y = _nums(rest, 3, path, ln, key)
                (verts if key == "v" else norms).append((x, y, z))
            elif key == "vt":
                tex.append(_nums(rest, 2, path, ln, key))
            elif key == "f":
                if len(rest) < 3:
                    raise ValueError(_where(path, ln) + "a face needs 3 vertices")
                refs = [_face_ref(r, path, ln) for r in rest[:3]]
                if material not in materials:
                    raise ValueError(_where(path, ln) + f"the OBJ file has a face with an unknown material {material}")
                if len(rest) > 3:
                    raise ValueError(_where(path, ln) + "the OBJ file has polygons, it should only have triangles")
                v.append([r[0] for r in refs])
                vt.append([r[1] for r in refs])
                vn.append([r[2] for r in refs])
                mids.append(mesh_id)
                mats.append(material)
            elif key == "usemtl":
                material = rest[0] if rest else ""
            elif key == "mtllib":
                if len(rest) != 1:
                    raise ValueError(_where(path, ln) + "the OBJ file has a material file name with spaces")
                for name, m in load_mtl(os.path.join(os.path.dirname(path), rest[0])).items():
                    materials.setdefault(name, m)  # known names are kept
    if not v:
        raise ValueError(f"{path}: the OBJ file has no triangles")
    o = ObjData(np.array(verts, np.float32).reshape(-1, 3), np.array(norms, np.float32).reshape(-1, 3),
                np.array(tex, np.float32).reshape(-1, 2), np.array(v, np.int32), np.array(vt, np.int32),
                np.array(vn, np.int32), np.array(mids, np.int32), mats, materials, mesh_id)
    if (o.v >= len(o.vertices)).any() or (o.vn >= len(o.normals)).any() or (o.vt >= len(o.texcoords)).any():
        raise ValueError(f"{path}: a face refers to a vertex, normal or texture coordinate the file does not have")
    return o


def _rotation(theta_deg, axis):
    """Rodrigues rotation matrix (transformations/Transform.cu:580-610) in
    float32, applied as m @ v."""
    ax, ay, az = (np.float32(a) for a in axis)
    t = np.float32(theta_deg) * (np.float32(math.pi) / np.float32(180.0))
    s, c = np.float32(math.sin(t)), np.float32(math.cos(t))
    one = np.float32(1)
    return np.array([[ax * ax + (one - ax * ax) * c, ax * ay * (one - c) - az * s, ax * az * (one - c) + ay * s],
                     [ax * ay * (one - c) + az * s, ay * ay + (one - ay * ay) * c, ay * az * (one - c) - ax * s],
                     [ax * az * (one - c) - ay * s, ay * az * (one - c) + ax * s, az * az + (one - az * az) * c]],
                    np.float32)


def load_obejota(path: str) -> ObejotaScene:
    """loadObejota with storeObejotaObject: the (last) camera and the point
    lights of the file."""
    if not os.path.isfile(path):
        raise ValueError(f"failed to open the obejota file {path}")
    sc = ObejotaScene()
    zero = (0.0, 0.0, 0.0)
    st = dict(type="", name="", p=zero, lookat=zero, cts=zero, up=zero, i=0.0, c=zero, frame_y=0.0)
    rot = [None]  # kept for the rest of the file once set, as in the reference

    def warn(w):
        print(f"Warning: {w}", file=sys.stderr)
        sc.warnings.append(w)

    def store():
        if st["type"] == "camera":
            up = st["up"]
            if rot[0] is not None:
                up = tuple(float(x) for x in rot[0] @ np.array(up, np.float32))
            if st["cts"] == zero:
                raise ValueError(f"{path}: camera {st['name']} has no camera to screen vector")
            if up == zero:
                raise ValueError(f"{path}: camera {st['name']} has no up vector")
            if st["frame_y"] == 0.0:
                raise ValueError(f"{path}: camera {st['name']} has no frame size")
            sc.camera = dict(pos=st["p"], lookat=st["lookat"], up=up, cts=st["cts"], frame_y=st["frame_y"])
        elif st["type"] == "lightPoint":
            if st["c"] == zero:
                warn(f"Light {st['name']} has no color.")
            if st["i"] == 0.0:
                warn(f"Light {st['name']} has no intensity.")
            sc.lights.append((st["p"], st["c"], st["i"]))
        st.update(type="", name="", p=zero, lookat=zero, cts=zero, up=zero, i=0.0, c=zero, frame_y=0.0)

    camera_found = False
    with open(path) as f:
        for ln, line in enumerate(f, 1):
            tok = line.split()
            if not tok:
                continue
            key, rest = tok[0], tok[1:]
            if key == "lightPoint":
                if st["name"]:
                    store()
                if rest:
                    st["name"] = rest[0]
                st["type"] = "lightPoint"
            elif key == "camera":
                if st["name"]:
                    store()
                st["name"], st["type"] = "camera", "camera"
                camera_found = True
            elif key in ("p", "lookat", "cts", "up"):
                x, y, z = _nums(rest, 3, path, ln, key)
                st[key] = (x, -y, z)  # Blender's y axis is inverted
            elif key == "i":
                st["i"] = _nums(rest, 1, path, ln, key)[0]
            elif key == "c":
                st["c"] = _nums(rest, 3, path, ln, key)
            elif key == "rotation":
                ax, ay, az = _nums(rest, 3, path, ln, key)
                m = _rotation(-az, (0, 0, 1)) @ (_rotation(-ay, (0, 1, 0)) @ _rotation(-ax, (1, 0, 0)))
                rot[0] = m
                local = m @ np.array([0, 0, -1], np.float32)
                st["lookat"] = tuple(float(np.float32(st["p"][j]) + local[j]) for j in range(3))
            elif key == "frameY":
                st["frame_y"] = _nums(rest, 1, path, ln, key)[0]
            elif key != "#":
                warn(f"Unknown token in obejota file: {key}")
    if st["name"]:
        store()
    if not camera_found:
        warn("No camera found in obejota file.")
    return sc


def texture_file_name(path: str, quiet: bool = False) -> str:
    """Texture::Texture's file name (textures/Texture.cu:127-149): a .jpg or
    .png name becomes .bmp (with the reference's warning on stdout), then
    every '.' after the first character except the last one becomes '_'."""
    dot = path.rfind(".")
    ext = path[dot + 1:] if dot >= 0 else path
    if ext in ("jpg", "png"):
        if not quiet:
            print(f"Warning: {path} is not a bmp file. Changing extension to bmp.")
        path = path[:dot] + ".bmp"
    last = path.rfind(".")
    return "".join("_" if (c == "." and 0 < i != last) else c for i, c in enumerate(path))


def load_bmp(name: str) -> np.ndarray:
    """Texture::loadBmpImage (textures/Texture.cu:40-119): (height, width, 3)
    uint8 RGB, row y = the file's row y (pixelMatrix[y]); 24- or 32-bit
    pixels read as consecutive B, G, R(, A) groups from the header's offset,
    rows not skipped to 4-byte boundaries, as in the reference."""
    import struct
    path = texture_file_name(name)
    if not os.path.isfile(path):
        raise ValueError(f"Error loading {path}: Failed to open the texture file")
    data = open(path, "rb").read()
    if len(data) < 54:
        raise ValueError(f"Error loading {path}: truncated header")
    if data[:2] != b"BM":
        raise ValueError(f"Error loading {path}: Not a valid BMP file.")
    offset, = struct.unpack_from("<I", data, 10)
    width, height = struct.unpack_from("<ii", data, 18)
    bpp, = struct.unpack_from("<H", data, 28)
    if bpp not in (24, 32):
        raise ValueError(f"Error loading {path}: Only 24 or 32 bit BMP files are supported, file is {bpp} bit")
    if width <= 0 or height <= 0:
        raise ValueError(f"Error loading {path}: only bottom-up BMP files with a positive size")
    step = bpp // 8
    n = width * height * step
    px = np.frombuffer(data, np.uint8, count=min(n, max(0, len(data) - offset)), offset=min(offset, len(data)))
    if px.size < n:
        raise ValueError(f"Error loading {path}: truncated pixel data")
    px = px.reshape(height, width, step)
    return np.ascontiguousarray(px[:, :, 2::-1])


def material_desc(m: MtlMaterial, texture: int = 0) -> dict:
    """The SceneDesc material an MTL block stands for, as the GPU path renders
    it: illum 0 Plain(Kd, Ke, map_Kd), 1 Lambertian(Kd, map_Kd),
    3 Dielectric(1, Ni), 5 Trowbridge(Kd, Pr, eta, k), 6 TrowbridgeBasic(Kd,
    Pr).  `texture` is the scene's texture index + 1 for a map_Kd (SceneFiles
    loads the file).  Raises UnsupportedMaterial for the others."""
    def refuse(why):
        raise UnsupportedMaterial(f"material {m.name} (illum {m.illum}, {ILLUM_CLASSES[m.illum]}): {why}")
    if m.map_kd and m.illum not in (0, 1, 3):
        refuse("textures are not rendered on the GPU path")
    if m.map_kd and m.illum in (0, 1) and not texture:
        refuse("its texture must be loaded into the scene (SceneFiles does that)")
    if m.illum == 0:
        if m.transparent:
            refuse("a Plain material without shadows is not rendered on the GPU path")
        return dict(type=_PLAIN, kd=tuple(m.kd), ior=1.5, roughness=0.0, eta=(1, 1, 1), k=(0, 0, 0), intensity=0.0,
                    ke=tuple(m.ke), texture=texture if m.map_kd else 0)
    if m.illum == 1:
        if m.transparent:
            refuse("a Lambertian material without shadows is not rendered on the GPU path")
        return dict(type=_LAMBERTIAN, kd=tuple(m.kd), ior=1.5, roughness=0.0, eta=(1, 1, 1), k=(0, 0, 0),
                    intensity=0.0, texture=texture if m.map_kd else 0)
    if m.illum == 3:
        return dict(type=_DIELECTRIC, kd=(1.0, 1.0, 1.0), ior=m.ni, roughness=0.0, eta=(1, 1, 1), k=(0, 0, 0),
                    intensity=0.0)
    if m.illum in (5, 6):
        if m.rough_x != m.rough_y:
            refuse("anisotropic roughness is not rendered on the GPU path")
        if m.illum == 5:
            return dict(type=_TROWBRIDGE, kd=tuple(m.kd), ior=1.5, roughness=m.rough_x, eta=tuple(m.eta),
                        k=tuple(m.k), intensity=0.0)
        return dict(type=_TROWBRIDGE_BASIC, kd=tuple(m.kd), ior=1.5, roughness=m.rough_x, eta=(1, 1, 1),
                    k=(0, 0, 0), intensity=0.0)
    refuse("this material class is not rendered on the GPU path")


class SceneFiles:
    """A scene read from OBJ files (meshes in file order, triangles in face
    order: the reference's triangle ids) and an .obejota file (camera and
    point lights), as cornellBoxGuion.cu and box.cu assemble theirs.  Spheres
    are added by code, as in those drivers; `material(name)` gives a
    material id for them."""

    def __init__(self, obj_paths, obejota_path=None):
        self.objs = []
        mesh_id = -1
        for p in obj_paths:
            o = load_obj(p, mesh_id)
            mesh_id = o.next_mesh_id
            self.objs.append(o)
        self.obejota = load_obejota(obejota_path) if obejota_path else ObejotaScene()
        self._mats, self._ids = [], {}
        self.textures, self._tex_ids = [], {}
        xyz, idx, tri_mat, uv, nrm = [], [], [], [], []
        base = 0
        default = np.array([0, 0, 1, 0, 1, 1], np.float32)
        for o in self.objs:
            xyz.append(o.vertices)
            idx.append(o.v.astype(np.int64) + base)
            tri_mat += [self._material_id(o, name) for name in o.material]
            base += len(o.vertices)
            # corner texture coordinates where the file has vt lines (faces
            # without them, and files without any, keep the default corners)
            t = np.tile(default, (len(o.v), 1))
            if len(o.texcoords):
                has = (o.vt >= 0).all(axis=1)
                t[has] = o.texcoords[o.vt[has]].reshape(-1, 6)
            uv.append(t)
            # corner normals where the file has vn lines (TriangleMesh::normals
            # is per file, Triangle.cu:379-418); a face that names no normals
            # gets none (the reference would read its unset normal indices)
            nr = np.full((len(o.v), 9), np.nan, np.float32)
            if len(o.normals):
                has = (o.vn >= 0).all(axis=1)
                nr[has] = o.normals[o.vn[has]].reshape(-1, 9)
            nrm.append(nr)
        self.xyz = np.concatenate(xyz).astype(np.float32) if xyz else np.zeros((0, 3), np.float32)
        self.idx = np.concatenate(idx).astype(np.uint32) if idx else np.zeros((0, 3), np.uint32)
        self.tri_mat = np.array(tri_mat, np.uint32)
        self.tri_uv = np.concatenate(uv) if uv and any(len(o.texcoords) for o in self.objs) else None
        self.tri_nrm = np.concatenate(nrm) if nrm and any(len(o.normals) for o in self.objs) else None
        self.spheres = []

    def _material_id(self, o, name):
        # materials are keyed by (file, name): two OBJ files may define one name differently
        key = (id(o), name)
        if key not in self._ids:
            m = o.materials[name]
            tex = 0
            if m.map_kd and m.illum in (0, 1):
                if m.map_kd not in self._tex_ids:
                    self.textures.append(load_bmp(m.map_kd))
                    self._tex_ids[m.map_kd] = len(self.textures)
                tex = self._tex_ids[m.map_kd]
            self._ids[key] = len(self._mats)
            self._mats.append(material_desc(m, tex))
        return self._ids[key]

    def material(self, name, obj=0):
        """Material id of `name` as the obj-th file's MTL files define it."""
        o = self.objs[obj]
        if name not in o.materials:
            raise KeyError(f"unknown material {name}")
        return self._material_id(o, name)

    def add_material(self, desc: dict) -> int:
        self._mats.append(dict(desc))
        return len(self._mats) - 1

    def add_sphere(self, center, radius, material, is_light=False):
        self.spheres.append((tuple(center), float(radius), int(material), int(bool(is_light))))

    def scene(self, width, height, spp=16, max_depth=8, rr_depth=3, seed=1) -> SceneDesc:
        if self.obejota.camera is None:
            raise ValueError("the scene files have no camera")
        cam = dict(self.obejota.camera, width=int(width), height=int(height))
        lights = [(tuple(p), tuple(c), float(i)) for p, c, i in self.obejota.lights]
        d = SceneDesc(self.xyz, self.idx, self.tri_mat, list(self._mats), list(self.spheres), lights, cam,
                      spp, max_depth, rr_depth, seed)
        d.textures = list(self.textures)
        d.tri_uv = self.tri_uv
        d.tri_nrm = self.tri_nrm
        return d
