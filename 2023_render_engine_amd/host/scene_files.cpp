// scene_files.cpp -- the scene-file loaders of include/rt_io.hpp from C++:
// a scene assembled like cornellBoxGuion.cu / box.cu (loadObj meshes,
// loadObejota camera and point lights, spheres added by code).
//
//   scene_files --dump <obejota|-> <obj>...
//       no GPU: print everything the loaders read, one item per line
//   scene_files --bmp <file>
//       no GPU: print the texels loadBMP reads
//   scene_files --render <out> <width> <height> <spp> <obejota> <obj>...
//               [--sphere cx cy cz r <material> [<obj index>]]...
//       render on GPU 0 with PathTracerIntegrator, store <out>.hdrbmp
//   scene_files --tonemap <film> <width> <height> <mode> <out> [bmp|ppm]
//       load <film>.hdrbmp and write <out>.bmp (or .ppm) tone-mapped on
//       GPU 0, as ToneMapper.cu does; <mode> is a ToneMappingType name
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/rt_io.hpp"

using namespace rt_amd;

static void p3(const char *tag, const Point3f &p) { std::printf("%s %.9g %.9g %.9g\n", tag, p.x, p.y, p.z); }
static void s3(const Spectrum &s) { std::printf(" %.9g %.9g %.9g", s.r, s.g, s.b); }

static int dump(const std::string &obejota, const std::vector<std::string> &objs)
{
    int32_t mesh = -1;
    for (const auto &path : objs) {
        const ObjData o = loadObj(path, mesh);
        mesh = o.next_mesh_id;
        std::printf("obj\n");
        std::printf("vertices %zu\n", o.vertices.size());
        for (const auto &v : o.vertices) p3("v", v);
        std::printf("normals %zu\n", o.normals.size());
        for (const auto &v : o.normals) p3("n", v);
        std::printf("texcoords %zu\n", o.texcoords.size());
        for (const auto &t : o.texcoords) std::printf("t %.9g %.9g\n", t.first, t.second);
        std::printf("triangles %zu\n", o.triangles.size());
        for (const auto &t : o.triangles)
            std::printf("f %d %d %d %d %d %d %d %d %d %d %s\n", t.v[0], t.v[1], t.v[2], t.vt[0], t.vt[1], t.vt[2],
                        t.vn[0], t.vn[1], t.vn[2], t.mesh_id, t.material.c_str());
        std::printf("materials %zu\n", o.materials.size());
        for (const auto &kv : o.materials) {
            const MtlMaterial &m = kv.second;
            std::printf("m %s %d", m.name.c_str(), m.illum);
            s3(m.ka); s3(m.kd); s3(m.ks); s3(m.ke);
            std::printf(" %.9g %.9g %.9g %.9g", m.ni, m.ns, m.rough_x, m.rough_y);
            s3(m.eta); s3(m.k);
            std::printf(" %d %d", m.transparent ? 1 : 0, m.has_indirect ? 1 : 0);
            s3(m.sigma_s); s3(m.sigma_a);
            std::printf(" %.9g %s\n", m.g, m.map_kd.empty() ? "-" : m.map_kd.c_str());
        }
        std::printf("next_mesh_id %d\n", o.next_mesh_id);
    }
    if (obejota != "-") {
        const ObejotaScene sc = loadObejota(obejota);
        std::printf("camera %d\n", sc.has_camera ? 1 : 0);
        if (sc.has_camera) {
            p3("pos", sc.camera.pos);
            p3("lookat", sc.camera.lookat);
            p3("up", sc.camera.up);
            p3("cts", sc.camera.cts);
            std::printf("frame_y %.9g\n", sc.camera.frame_y);
        }
        std::printf("lights %zu\n", sc.lights.size());
        for (const auto &l : sc.lights) {
            std::printf("l %.9g %.9g %.9g", l.pos.x, l.pos.y, l.pos.z);
            s3(l.c);
            std::printf(" %.9g\n", l.I);
        }
        std::printf("warnings %zu\n", sc.warnings.size());
    }
    return 0;
}

static int render(int argc, char **argv)
{
    if (argc < 8) return 2;
    const std::string out = argv[2];
    const uint32_t w = (uint32_t)std::atoi(argv[3]), h = (uint32_t)std::atoi(argv[4]);
    const uint32_t spp = (uint32_t)std::atoi(argv[5]);
    const ObejotaScene sc = loadObejota(argv[6]);
    if (!sc.has_camera) throw Error(RT_ERR_INVALID, std::string(argv[6]) + ": no camera");
    std::vector<ObjData> objs;
    std::vector<std::map<std::string, std::shared_ptr<Material>>> mats;
    TextureCache textures;
    std::vector<std::shared_ptr<Shape>> shapes;
    int i = 7;
    int32_t mesh = -1;
    for (; i < argc && std::strcmp(argv[i], "--sphere") != 0; ++i) {
        objs.push_back(loadObj(argv[i], mesh));
        mesh = objs.back().next_mesh_id;
    }
    mats.resize(objs.size());  // one material cache per OBJ file, as loadObj keeps one map per call
    for (size_t k = 0; k < objs.size(); ++k)
        for (auto &s : obj_shapes(objs[k], mats[k], &textures)) shapes.push_back(s);
    while (i < argc && std::strcmp(argv[i], "--sphere") == 0) {
        if (i + 5 >= argc) return 2;
        const Point3f c{(float)std::atof(argv[i + 1]), (float)std::atof(argv[i + 2]), (float)std::atof(argv[i + 3])};
        const float r = (float)std::atof(argv[i + 4]);
        const std::string name = argv[i + 5];
        size_t k = 0;
        i += 6;
        if (i < argc && std::strcmp(argv[i], "--sphere") != 0) k = (size_t)std::atoi(argv[i++]);
        auto it = mats.at(k).find(name);
        if (it == mats[k].end())
            it = mats[k].emplace(name, make_material(objs.at(k).materials.at(name), &textures)).first;
        shapes.push_back(std::make_shared<Sphere>(c, r, it->second));
    }
    Context ctx(0);
    Scene scene(ctx);
    ProjectiveCamera cam = make_camera(sc.camera, w, h);
    scene.build(shapes, sc.lights, cam);
    RenderSettings s;
    s.spp = spp;
    PathTracerIntegrator(&cam).render(scene, s);
    cam.film.storeContributionsToFile(out);
    std::printf("rendered %ux%u at %u spp into %s.hdrbmp\n", w, h, spp, out.c_str());
    return 0;
}

static int tonemap(int argc, char **argv)
{
    if (argc < 7 || argc > 8) return 2;
    static const char *names[] = {"equalize", "reinhard", "reinhardJodie", "ACESfilmic", "uncharted2", "gammaClamp"};
    int mode = -1;
    for (int k = 0; k < 6; ++k)
        if (std::strcmp(argv[5], names[k]) == 0) mode = k;
    if (mode < 0) throw Error(RT_ERR_INVALID, std::string("unknown tone mapping ") + argv[5]);
    const std::string fmt = argc == 8 ? argv[7] : "bmp";
    if (fmt != "bmp" && fmt != "ppm") throw Error(RT_ERR_INVALID, "format must be bmp or ppm");
    Film film((uint32_t)std::atoi(argv[3]), (uint32_t)std::atoi(argv[4]));
    if (!film.loadContributionsFromFile(argv[2])) throw Error(RT_ERR_INVALID, std::string(argv[2]) + ".hdrbmp not found");
    if (fmt == "bmp")
        film.writeToBMP(argv[6], (ToneMappingType)mode);
    else
        film.writeToPPM(argv[6], (ToneMappingType)mode);
    std::printf("wrote %s.%s (%s, %llu samples)\n", argv[6], fmt.c_str(), names[mode],
                (unsigned long long)film.samples());
    return 0;
}

int main(int argc, char **argv)
{
    try {
        if (argc >= 3 && std::strcmp(argv[1], "--dump") == 0)
            return dump(argv[2], std::vector<std::string>(argv + 3, argv + argc));
        if (argc == 3 && std::strcmp(argv[1], "--bmp") == 0) {
            // the texels as loadBMP reads them: "w h" then one "r g b" line per texel
            const Texture t = loadBMP(argv[2]);
            std::printf("%u %u\n", t.width, t.height);
            for (size_t i = 0; i < (size_t)t.width * t.height; ++i)
                std::printf("%u %u %u\n", t.rgb[3 * i], t.rgb[3 * i + 1], t.rgb[3 * i + 2]);
            return 0;
        }
        if (argc >= 2 && std::strcmp(argv[1], "--render") == 0) {
            const int r = render(argc, argv);
            if (r == 2) std::fprintf(stderr, "usage: scene_files --render <out> <w> <h> <spp> <obejota> <obj>...\n");
            return r;
        }
        if (argc >= 2 && std::strcmp(argv[1], "--tonemap") == 0) {
            const int r = tonemap(argc, argv);
            if (r == 2) std::fprintf(stderr, "usage: scene_files --tonemap <film> <w> <h> <mode> <out> [bmp|ppm]\n");
            return r;
        }
        std::fprintf(stderr, "usage: scene_files --dump <obejota|-> <obj>... | --render ... | --tonemap ...\n");
        return 2;
    } catch (const Error &e) {
        std::fprintf(stderr, "error (status %d): %s\n", (int)e.status, e.what());
        return 1;
    }
}
