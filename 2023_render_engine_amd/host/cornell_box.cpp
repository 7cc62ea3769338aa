// cornell_box.cpp -- the Cornell-box driver (cornellBoxGuion.cu) written
// against the C++ host API (include/rt.hpp): builds the scene from shapes,
// materials, a point light and the .obejota camera, renders it with
// PathTracerIntegrator and stores the film as <out>.hdrbmp.
//
//   cornell_box <out> [width height spp [mode [photons]]]   render on GPU 0
//       mode: pt (path tracing, the default), photon (PHOTON_MAPPING) or
//       hybrid (HYBRID); photons: photon paths of the pre-pass (2^20)
//   cornell_box --host-check <dir>           no GPU: BVH build + film I/O
//
// Geometry, materials and camera are those of the Cornell-box workload: a
// box [-0.5, 0.5]^3 of 5 quads open towards -x, a Lambertian (0.73) and a
// glass sphere of radius 0.2 on the floor, a point light of intensity 10
// just below the ceiling centre, camera p = (-1.75, 0, 0) looking at the
// origin, up +z, cts (0, 0, -0.8), frameY 0.6.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "../../include/rt.hpp"

using namespace rt_amd;

static std::shared_ptr<TriangleMesh> quad(const float p[4][3], std::shared_ptr<Material> m)
{
    std::vector<float> v;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 3; ++k) v.push_back(p[i][k]);
    return std::make_shared<TriangleMesh>(v, std::vector<uint32_t>{0, 1, 2, 0, 2, 3}, m);
}

static std::vector<std::shared_ptr<Shape>> cornell_shapes()
{
    const float a = 0.5f;
    auto white = std::make_shared<LambertianMaterial>(Spectrum(0.73f));
    auto red = std::make_shared<LambertianMaterial>(Spectrum(0.63f, 0.06f, 0.05f));
    auto green = std::make_shared<LambertianMaterial>(Spectrum(0.14f, 0.45f, 0.09f));
    auto glass = std::make_shared<GlassMaterial>(Spectrum(1.0f));
    const float floor_[4][3] = {{-a, -a, -a}, {a, -a, -a}, {a, a, -a}, {-a, a, -a}};
    const float ceil_[4][3] = {{-a, -a, a}, {-a, a, a}, {a, a, a}, {a, -a, a}};
    const float back[4][3] = {{a, -a, -a}, {a, -a, a}, {a, a, a}, {a, a, -a}};
    const float left[4][3] = {{-a, -a, -a}, {-a, -a, a}, {a, -a, a}, {a, -a, -a}};
    const float right[4][3] = {{-a, a, -a}, {a, a, -a}, {a, a, a}, {-a, a, a}};
    return {quad(floor_, white), quad(ceil_, white), quad(back, white), quad(left, red), quad(right, green),
            std::make_shared<Sphere>(Point3f{0.2f, -0.2f, -0.3f}, 0.2f, white),
            std::make_shared<Sphere>(Point3f{-0.1f, 0.2f, -0.3f}, 0.2f, glass)};
}

static int host_check(const std::string &dir)
{
    // BVH over the box's triangles, built on the host only
    std::vector<float> xyz;
    std::vector<uint32_t> idx;
    for (auto &s : cornell_shapes())
        if (auto *t = dynamic_cast<TriangleMesh *>(s.get())) {
            const uint32_t base = (uint32_t)(xyz.size() / 3);
            xyz.insert(xyz.end(), t->vertices.begin(), t->vertices.end());
            for (uint32_t i : t->indices) idx.push_back(base + i);
        }
    BVH bvh(nullptr, xyz, idx);
    const rt_bvh_info inf = bvh.info();
    if (inf.num_tris != 10) { std::fprintf(stderr, "bvh: %u triangles\n", inf.num_tris); return 1; }
    // film round trip: store, load into a fresh film (adds), RMSE 0
    Film f(4, 3);
    for (uint32_t y = 0; y < 3; ++y)
        for (uint32_t x = 0; x < 4; ++x) {
            const double rgb[3] = {0.25 * x, 0.5 * y, 1.0 + x * y};
            f.addSample(x, y, rgb, 8);
        }
    f.storeContributionsToFile(dir + "/film");
    Film g(4, 3);
    if (!g.loadContributionsFromFile(dir + "/film")) { std::fprintf(stderr, "load failed\n"); return 1; }
    if (g.samples() != 8 || computeRMSE(f, g) != 0.0f) { std::fprintf(stderr, "round trip differs\n"); return 1; }
    Film wrong(3, 3);
    bool threw = false;
    try {
        wrong.loadContributionsFromFile(dir + "/film");
    } catch (const Error &) {
        threw = true;
    }
    if (!threw) { std::fprintf(stderr, "resolution mismatch accepted\n"); return 1; }
    if (computeRMSE(f, wrong) != -1.0f) return 1;
    std::printf("host check ok: %u nodes, film round trip exact\n", inf.num_nodes);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc >= 3 && std::string(argv[1]) == "--host-check") return host_check(argv[2]);
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <out> [width height spp [pt|photon|hybrid [photons]]] | --host-check <dir>\n",
                     argv[0]);
        return 2;
    }
    const uint32_t w = argc > 3 ? (uint32_t)std::atoi(argv[2]) : 256;
    const uint32_t h = argc > 3 ? (uint32_t)std::atoi(argv[3]) : 256;
    const uint32_t spp = argc > 4 ? (uint32_t)std::atoi(argv[4]) : 16;
    const std::string mode = argc > 5 ? argv[5] : "pt";
    RenderSettings s;
    s.spp = spp;
    if (mode == "photon") {
        s.mode = RenderMode::PHOTON_MAPPING;
    } else if (mode == "hybrid") {
        s.mode = RenderMode::HYBRID;
    } else if (mode != "pt") {
        std::fprintf(stderr, "unknown mode %s (pt, photon or hybrid)\n", mode.c_str());
        return 2;
    }
    if (argc > 6) s.photons = (uint32_t)std::atoi(argv[6]);
    try {
        Context ctx(0);
        ProjectiveCamera cam({-1.75f, 0, 0}, {0, 0, 0}, {0, 0, 1}, {0, 0, -0.8f}, 0.6f, w, h);
        Scene scene(ctx);
        scene.build(cornell_shapes(), {PointLight({0, 0, 0.5f - 1e-3f}, Spectrum(1.0f), 10.0f)}, cam);
        PathTracerIntegrator integrator(&cam);
        integrator.render(scene, s);
        cam.film.storeContributionsToFile(argv[1]);
        std::printf("rendered %ux%u at %u spp (%s) into %s.hdrbmp\n", w, h, spp, mode.c_str(), argv[1]);
    } catch (const Error &e) {
        std::fprintf(stderr, "error (%d): %s\n", (int)e.status, e.what());
        return 1;
    }
    return 0;
}
