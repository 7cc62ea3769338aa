"""The scene-file loaders (include/rt_io.hpp from C++, scene_files.py from
Python) against the reference's reading rules (cuda_libraries/
FileImporting.cu:27-598), without a GPU.

Pinned by:
- the Cornell box written as OBJ / MTL / .obejota loads into exactly the
  arrays of the built-in Cornell workload;
- hand-derived answers for every keyword, index form and quirk (axis swap,
  y negation, Ke / sigma carried across newmtl, mtllib precedence, the mesh
  counter), including a camera at Euler angles (90, 0, 0) whose lookat and
  up follow by hand;
- the camera rotation against scipy's extrinsic-XYZ Euler matrices at random
  angles (an independent construction of Rz Ry Rx);
- the C++ and Python loaders printing the same scene for the same files;
- the reference's own .obejota and .mtl files, where /root/reference is
  present (skipped otherwise)."""
import importlib
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "scene_files")
EXE = os.path.join(ROOT, "2023_render_engine_amd", "host", "scene_files")
REF = "/root/reference"

sf = importlib.import_module("2023_render_engine_amd.scene_files")
scenes = importlib.import_module("2023_render_engine_amd.scenes")


def fx(name):
    return os.path.join(FIX, name)


@pytest.fixture(scope="module")
def exe(rt):
    if not os.path.exists(EXE):
        subprocess.run(["make", "-C", os.path.join(ROOT, "2023_render_engine_amd", "csrc")], check=True)
    return EXE


# ---------------------------------------------------------------------------
# the Cornell box from files == the built-in Cornell workload
# ---------------------------------------------------------------------------
def cornell_from_files(width=256, height=256, spp=16):
    s = sf.SceneFiles([fx("cornell.obj")], fx("cornell.obejota"))
    s.add_sphere((0.2, -0.2, -0.3), 0.2, s.material("white"))
    s.add_sphere((-0.1, 0.2, -0.3), 0.2, s.material("glass"))
    return s.scene(width, height, spp=spp)


def test_cornell_files_load_to_the_cornell_scene():
    a, b = cornell_from_files(), scenes.cornell()
    assert np.array_equal(a.xyz, b.xyz) and a.xyz.dtype == np.float32
    assert np.array_equal(a.idx, b.idx)
    assert np.array_equal(a.tri_mat, b.tri_mat)
    assert len(a.materials) == len(b.materials)
    for ma, mb in zip(a.material_structs(), b.material_structs()):
        assert bytes(ma) == bytes(mb)
    assert a.spheres == [tuple((tuple(c), r, m, l)) for c, r, m, l in b.spheres]
    assert [(tuple(np.float32(p)), tuple(np.float32(c)), np.float32(i)) for p, c, i in a.lights] == \
        [(tuple(np.float32(p)), tuple(np.float32(c)), np.float32(i)) for p, c, i in b.lights]
    # equal values (the y negation turns 0 into -0, which compares and renders the same)
    ca, cb = a.camera_struct(), b.camera_struct()
    for k in ("pos", "lookat", "up", "cts"):
        assert list(getattr(ca, k)) == list(getattr(cb, k))
    assert (ca.frame_y, ca.width, ca.height) == (cb.frame_y, cb.width, cb.height)


# ---------------------------------------------------------------------------
# known answers
# ---------------------------------------------------------------------------
def test_obj_index_forms_objects_and_axis_swap():
    o = sf.load_obj(fx("features.obj"))
    # file "x y z" -> (x, z, y)
    assert o.vertices.tolist() == [[0, 0, 0], [1, 0, 0], [0, 0, 1], [0, 1, 0], [1, 1, 1]]
    assert o.normals.tolist() == [[0, 1, 0], [0, 0, 1]]
    assert o.texcoords.tolist() == [[0, 0], [1, 0], [0, 1]]
    assert o.v.tolist() == [[0, 1, 2], [0, 1, 2], [0, 1, 3], [0, 1, 4], [1, 2, 3], [2, 3, 4], [0, 3, 4]]
    assert o.vt.tolist() == [[-1] * 3, [0, 1, 2], [-1] * 3, [0, 1, 2], [-1] * 3, [-1] * 3, [-1] * 3]
    assert o.vn.tolist() == [[-1] * 3, [-1] * 3, [0, 1, 0], [0, 1, 1], [-1] * 3, [-1] * 3, [-1] * 3]
    # the mesh counter starts at -1 and every `o` advances it
    assert o.mesh_id.tolist() == [-1, 0, 0, 0, 1, 1, 1] and o.next_mesh_id == 1
    assert o.material == ["default", "lamb", "lamb", "ggx", "extra", "ggx_basic", "glass2"]
    # a second file continues the counter
    assert sf.load_obj(fx("cornell.obj"), o.next_mesh_id).mesh_id.tolist() == [2, 2, 3, 3, 4, 4, 5, 5, 6, 6]


def test_mtl_keywords_resets_and_precedence():
    m = sf.load_obj(fx("features.obj")).materials
    assert sorted(m) == ["aniso", "brick", "default", "extra", "ggx", "ggx_basic", "glass2", "lamb", "plain",
                         "shadowless", "skin"]
    f = lambda *v: tuple(float(np.float32(x)) for x in v)  # noqa: E731
    lamb = m["lamb"]  # the first mtllib's definition wins over extra.mtl's
    assert (lamb.illum, lamb.ka, lamb.kd, lamb.ks, lamb.ke) == (1, f(.1, .2, .3), f(.5, .6, .7), f(.01, .02, .03),
                                                                 f(2, 3, 4))
    assert (lamb.ns, lamb.ni, lamb.transparent, lamb.has_indirect) == (12.0, f(1.3)[0], False, False)
    assert m["default"].kd == (1, 1, 1) and m["default"].illum == 1  # built in, never replaced
    assert m["extra"].kd == f(.3, .35, .4)
    ggx = m["ggx"]
    # newmtl resets Ka/Kd/Ks/Ni/Ns/illum/flags/roughness/eta/k but not Ke
    assert (ggx.ka, ggx.ke, ggx.ni, ggx.ns, ggx.has_indirect) == ((0, 0, 0), f(2, 3, 4), 1.0, 0.0, True)
    assert (ggx.rough_x, ggx.rough_y, ggx.eta, ggx.k) == (0.25, 0.25, f(.2, .3, .4), f(3, 2.5, 2))
    assert (m["aniso"].rough_x, m["aniso"].rough_y) == f(.1, .3)
    assert m["glass2"].transparent and m["glass2"].ni == f(1.33)[0]
    # Ke, sigmaS, sigmaA and g carry over into later blocks
    assert m["skin"].ke == (5, 5, 5) and m["shadowless"].ke == (5, 5, 5)
    for name in ("skin", "brick", "shadowless"):
        assert (m[name].sigma_s, m[name].sigma_a, m[name].g) == (f(1, 2, 3), f(.1, .2, .3), 0.5)
    # a texture that is not next to the MTL is looked for in textures/
    assert m["brick"].map_kd == os.path.join(FIX, "textures", "bricks.png")


def test_material_classes_on_the_gpu_path():
    m = sf.load_obj(fx("features.obj")).materials
    assert sf.material_desc(m["lamb"])["type"] == 0
    g = sf.material_desc(m["ggx"])
    assert (g["type"], g["roughness"], g["eta"], g["k"]) == (2, 0.25, m["ggx"].eta, m["ggx"].k)
    assert sf.material_desc(m["ggx_basic"])["type"] == 3
    d = sf.material_desc(m["glass2"])  # DielectricMaterial(Spectrum(1), Ni)
    assert (d["type"], d["kd"], d["ior"]) == (1, (1.0, 1.0, 1.0), m["glass2"].ni)
    p = sf.material_desc(m["plain"])  # PlainMaterial(kd, ke); Ke 5 5 5 of its block
    assert (p["type"], p["kd"], p["ke"]) == (6, (1, 1, 1), (5, 5, 5))
    for name, why in [("aniso", "anisotropic"), ("skin", "SSMaterial"), ("brick", "texture must be loaded"),
                      ("shadowless", "without shadows")]:
        with pytest.raises(sf.UnsupportedMaterial, match=why):
            sf.material_desc(m[name])
    # a scene only converts the materials its shapes use
    s = sf.SceneFiles([fx("features.obj")])
    assert s.tri_mat.tolist() == [0, 1, 1, 2, 3, 4, 5]
    with pytest.raises(sf.UnsupportedMaterial):
        s.material("skin")


def test_obejota_known_answers():
    sc = sf.load_obejota(fx("rotated.obejota"))
    c = sc.camera
    assert c["pos"] == (1, -2, 3) and c["cts"] == (0, 0, -1) and c["frame_y"] == 0.5
    # Rx(-90) maps (0, 0, -1) to (0, -1, 0): lookat = p + that
    assert np.allclose(c["lookat"], (1, -3, 3), atol=1e-6)
    # up (0, 1, 0) -> y negated (0, -1, 0) -> Rx(-90) -> (0, 0, 1)
    assert np.allclose(c["up"], (0, 0, 1), atol=1e-6)
    assert sc.lights[0] == ((1, 2, 3), (1, 0.5, 0.25), 7)
    assert sc.lights[1] == ((0, 0, 1), (0, 0, 0), 0)
    assert sc.warnings == ["Unknown token in obejota file: weird", "Light dark has no color.",
                           "Light dark has no intensity."]
    cam = sf.load_obejota(fx("cornell.obejota")).camera
    assert (cam["pos"], cam["lookat"], cam["up"], cam["cts"], cam["frame_y"]) == \
        ((-1.75, 0, 0), (0, 0, 0), (0, 0, 1), (0, 0, float(np.float32(-0.8))), float(np.float32(0.6)))


def test_obejota_rotation_is_extrinsic_xyz_euler(tmp_path):
    """lookat - p = Rz(-az) Ry(-ay) Rx(-ax) (0, 0, -1): scipy's extrinsic
    'xyz' Euler matrix of the negated angles, an independent construction."""
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(11)
    for _ in range(20):
        ang = rng.uniform(-180, 180, 3)
        p = rng.uniform(-3, 3, 3)
        up = rng.normal(size=3)
        path = tmp_path / "r.obejota"
        path.write_text(f"camera\np {p[0]} {p[1]} {p[2]}\nrotation {ang[0]} {ang[1]} {ang[2]}\n"
                        f"up {up[0]} {up[1]} {up[2]}\ncts 0 0 -1\nframeY 1\n")
        c = sf.load_obejota(str(path)).camera
        R = Rotation.from_euler("xyz", -ang, degrees=True).as_matrix()
        pn = p * [1, -1, 1]
        assert np.allclose(np.array(c["lookat"]) - pn, R @ [0, 0, -1], atol=2e-6)
        assert np.allclose(c["up"], R @ (up * [1, -1, 1]), atol=2e-6 * (1 + np.abs(up).max()))


@pytest.mark.parametrize("text,match", [
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf 1 2 3 4\n", "polygons"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl nothing\nf 1 2 3\n", "unknown material nothing"),
    ("v 0 0 0\nv 1 0 0\n", "no triangles"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n", "does not have"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 -1\n", "bad face index"),
    ("v 0 0\n", "expected a number"),
    ("mtllib a b.mtl\n", "spaces"),
])
def test_obj_errors(exe, tmp_path, text, match):
    p = tmp_path / "bad.obj"
    p.write_text(text)
    with pytest.raises(ValueError, match=match):
        sf.load_obj(str(p))
    r = subprocess.run([exe, "--dump", "-", str(p)], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and match in r.stderr


@pytest.mark.parametrize("text,match", [
    ("camera\np 0 0 0\nup 0 0 1\nframeY 1\n", "camera to screen"),
    ("camera\np 0 0 0\ncts 0 0 -1\nframeY 1\n", "no up vector"),
    ("camera\np 0 0 0\ncts 0 0 -1\nup 0 0 1\n", "no frame size"),
    ("camera\np 0 0\n", "expected a number"),
])
def test_obejota_errors(exe, tmp_path, text, match):
    p = tmp_path / "bad.obejota"
    p.write_text(text)
    with pytest.raises(ValueError, match=match):
        sf.load_obejota(str(p))
    r = subprocess.run([exe, "--dump", str(p)], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and match in r.stderr


def test_mtl_errors_and_missing_file(tmp_path):
    assert sf.load_mtl(str(tmp_path / "none.mtl")) == {}
    p = tmp_path / "bad.mtl"
    p.write_text("newmtl a\nmakesShadow 2\n")
    with pytest.raises(ValueError, match="makesShadow must be 0 or 1"):
        sf.load_mtl(str(p))
    p.write_text("newmtl a\nillum 9\n")
    with pytest.raises(ValueError, match="illum must be 0..7"):
        sf.load_mtl(str(p))


# ---------------------------------------------------------------------------
# C++ and Python read the same scene
# ---------------------------------------------------------------------------
def _g(x):
    return "%.9g" % float(np.float32(x))


def py_dump(obejota, objs):
    out = []
    mesh = -1
    for path in objs:
        o = sf.load_obj(path, mesh)
        mesh = o.next_mesh_id
        out.append("obj")
        out.append(f"vertices {len(o.vertices)}")
        out += ["v " + " ".join(_g(x) for x in v) for v in o.vertices]
        out.append(f"normals {len(o.normals)}")
        out += ["n " + " ".join(_g(x) for x in v) for v in o.normals]
        out.append(f"texcoords {len(o.texcoords)}")
        out += ["t " + " ".join(_g(x) for x in v) for v in o.texcoords]
        out.append(f"triangles {len(o.v)}")
        for i in range(len(o.v)):
            out.append("f " + " ".join(str(int(x)) for x in [*o.v[i], *o.vt[i], *o.vn[i], o.mesh_id[i]])
                       + " " + o.material[i])
        out.append(f"materials {len(o.materials)}")
        for name in sorted(o.materials):
            m = o.materials[name]
            vals = [*m.ka, *m.kd, *m.ks, *m.ke, m.ni, m.ns, m.rough_x, m.rough_y, *m.eta, *m.k]
            tail = [*m.sigma_s, *m.sigma_a]
            out.append(f"m {m.name} {m.illum} " + " ".join(_g(x) for x in vals)
                       + f" {int(m.transparent)} {int(m.has_indirect)} " + " ".join(_g(x) for x in tail)
                       + f" {_g(m.g)} {m.map_kd or '-'}")
        out.append(f"next_mesh_id {o.next_mesh_id}")
    if obejota != "-":
        sc = sf.load_obejota(obejota)
        out.append(f"camera {int(sc.camera is not None)}")
        if sc.camera is not None:
            for k in ("pos", "lookat", "up", "cts"):
                out.append(k + " " + " ".join(_g(x) for x in sc.camera[k]))
            out.append(f"frame_y {_g(sc.camera['frame_y'])}")
        out.append(f"lights {len(sc.lights)}")
        for p, c, i in sc.lights:
            out.append("l " + " ".join(_g(x) for x in [*p, *c, i]))
        out.append(f"warnings {len(sc.warnings)}")
    return out


def same_dump(a, b):
    """Token-wise equality; numbers may differ by an ulp or two (sin and cos
    of the camera rotation)."""
    assert len(a) == len(b)
    for la, lb in zip(a, b):
        ta, tb = la.split(), lb.split()
        assert len(ta) == len(tb), (la, lb)
        for x, y in zip(ta, tb):
            if x == y:
                continue
            fx_, fy = float(x), float(y)
            assert abs(fx_ - fy) <= 1e-6 * max(1.0, abs(fx_)), (la, lb)


@pytest.mark.parametrize("obejota,objs", [
    (fx("cornell.obejota"), [fx("cornell.obj")]),
    (fx("rotated.obejota"), [fx("features.obj"), fx("cornell.obj")]),
])
def test_cpp_and_python_loaders_agree(exe, obejota, objs):
    r = subprocess.run([exe, "--dump", obejota, *objs], capture_output=True, text=True, timeout=30)
    assert r.returncode == 0, r.stderr
    same_dump(r.stdout.strip().splitlines(), py_dump(obejota, objs))


# ---------------------------------------------------------------------------
# the reference's own scene files (read as data where they are present)
# ---------------------------------------------------------------------------
REF_OBEJOTA = ["cornellBoxGuion.obejota", "box.obejota", "models/bath/bath.obejota",
               "models/kaerMorhen/kaerMorhen.obejota"]


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference's scene files are not present")
@pytest.mark.parametrize("name", REF_OBEJOTA)
def test_reference_obejota_files(exe, name):
    path = os.path.join(REF, name)
    sc = sf.load_obejota(path)
    assert sc.camera is not None and sc.lights
    r = subprocess.run([exe, "--dump", path], capture_output=True, text=True, timeout=30)
    assert r.returncode == 0, r.stderr
    same_dump(r.stdout.strip().splitlines(), py_dump(path, []))
    if name == "cornellBoxGuion.obejota":
        # the Cornell workload's camera; the file's light is at z = 0.25, intensity 5
        c = sc.camera
        assert (c["pos"], c["lookat"], c["up"], c["frame_y"]) == ((-1.75, 0, 0), (0, 0, 0), (0, 0, 1),
                                                                   float(np.float32(0.6)))
        assert sc.lights == [((0, 0, 0.25), (1, 1, 1), 5)]


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference's scene files are not present")
def test_reference_mtl_files():
    m = sf.load_mtl(os.path.join(REF, "models/bath/bath.mtl"))
    illum = sorted(x.illum for x in m.values())
    assert len(m) == 39
    # 40 illum lines over 39 blocks: "Hair" states illum 6 twice
    assert {k: illum.count(k) for k in set(illum)} == {0: 1, 1: 1, 2: 18, 3: 5, 4: 7, 6: 4, 7: 3}
    m2 = sf.load_mtl(os.path.join(REF, "models/bath/bathMaterials.mtl"))
    assert len(m2) == 13


# ---------------------------------------------------------------------------
# textures: BMP files, map_Kd and vt
# ---------------------------------------------------------------------------
def write_bmp(path, rgb, bpp=24, pad_rows=False):
    """A bottom-up BMP whose pixel rows are rgb[0], rgb[1], ... (file order),
    rows packed without padding (as the reference reads them) unless
    pad_rows."""
    import struct
    h, w, _ = rgb.shape
    step = bpp // 8
    rows = []
    for y in range(h):
        row = bytearray()
        for x in range(w):
            r, g, b = (int(c) for c in rgb[y, x])
            row += bytes([b, g, r] + ([255] if bpp == 32 else []))
        if pad_rows:
            row += bytes((-len(row)) % 4)
        rows.append(bytes(row))
    data = b"".join(rows)
    hdr = b"BM" + struct.pack("<IHHIIiiHHIIiiII", 54 + len(data), 0, 0, 54, 40, w, h, 1, bpp, 0, len(data), 2835,
                              2835, 0, 0)
    open(path, "wb").write(hdr + data)


def test_bmp_reading_matches_between_cpp_and_python(exe, tmp_path):
    rng = np.random.default_rng(5)
    for bpp in (24, 32):
        rgb = rng.integers(0, 256, (5, 3, 3), dtype=np.uint8)
        p = str(tmp_path / f"t{bpp}.bmp")
        write_bmp(p, rgb, bpp)
        got = sf.load_bmp(p)
        assert np.array_equal(got, rgb)
        r = subprocess.run([exe, "--bmp", p], capture_output=True, text=True, timeout=30)
        assert r.returncode == 0, r.stderr
        lines = r.stdout.split("\n")
        assert lines[0] == "3 5"
        assert np.array_equal(np.array([[int(v) for v in ln.split()] for ln in lines[1:16]], np.uint8).reshape(5, 3, 3),
                              rgb)


def test_bmp_rows_are_read_without_padding(tmp_path):
    """A 3-pixel-wide 24-bit row is 9 bytes; a padded file has 3 more per row,
    which the reference reads as pixel data (Texture.cu:84-116)."""
    rgb = np.arange(2 * 3 * 3, dtype=np.uint8).reshape(2, 3, 3) * 7
    p = str(tmp_path / "pad.bmp")
    write_bmp(p, rgb, 24, pad_rows=True)
    got = sf.load_bmp(p)
    flat = open(p, "rb").read()[54:54 + 18]
    want = np.frombuffer(flat, np.uint8).reshape(2, 3, 3)[:, :, ::-1]
    assert np.array_equal(got, want) and not np.array_equal(got, rgb)


def test_texture_file_names():
    assert sf.texture_file_name("textures/bricks.png") == "textures/bricks.bmp"
    assert sf.texture_file_name("a.b.c.jpg") == "a_b_c.bmp"
    assert sf.texture_file_name("./tex/x.y.bmp") == "./tex/x_y.bmp"
    assert sf.texture_file_name("x.bmp") == "x.bmp"


def textured_quad_files(tmp_path):
    """The textured quad probe as files: an OBJ with vt corners, its MTL
    (Plain, kd 1, map_Kd tex.png -> tex.bmp) and the 2 x 2 texture."""
    import material_scenes as ms
    write_bmp(str(tmp_path / "tex.bmp"), ms.TEX2)
    (tmp_path / "q.mtl").write_text("newmtl tex\nKd 1 1 1\nKe 0 0 0\nmap_Kd tex.png\nillum 0\n")
    # the file stores (x, z, y)
    (tmp_path / "q.obj").write_text("mtllib q.mtl\n"
                                    "v 0.5 -1 -1\nv 0.5 -1 1\nv 0.5 1 1\nv 0.5 1 -1\n"
                                    "vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
                                    "usemtl tex\nf 1/1 2/2 3/3\nf 1/1 3/3 4/4\n")
    (tmp_path / "q.obejota").write_text("camera\np -1.75 0 0\nlookat 0 0 0\nup 0 0 1\ncts 0 0 -0.8\nframeY 0.6\n")
    return str(tmp_path / "q.obj"), str(tmp_path / "q.obejota")


def test_map_kd_and_vt_load_to_the_textured_probe(tmp_path):
    import material_scenes as ms
    obj, obejota = textured_quad_files(tmp_path)
    d = sf.SceneFiles([obj], obejota).scene(32, 32, spp=4)
    want = ms.textured_probe("quad", 32, 32, 4)
    assert np.array_equal(d.xyz, want.xyz) and np.array_equal(d.idx, want.idx)
    assert np.array_equal(d.tri_uv, want.tri_uv)
    assert len(d.textures) == 1 and np.array_equal(d.textures[0], ms.TEX2)
    m = d.material_structs()[0]
    assert (m.type, m.texture, tuple(m.kd)) == (6, 1, (1.0, 1.0, 1.0))
