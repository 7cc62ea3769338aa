// rt.hpp -- C++17 host API over the C ABI of rt.h, with the reference
// renderer's nouns: Scene, Shape (TriangleMesh, Sphere, Plane), Material
// (Lambertian, Dielectric/Glass, Trowbridge, TrowbridgeBasic, Light, Mirror,
// Plain),
// PointLight, ProjectiveCamera, Film, PathTracerIntegrator, BVH.
//
// Header-only; link with librt_amd.so.  Every failing C call throws
// rt_amd::Error carrying the status and rt_last_error().  Device buffers are
// owned by the objects that allocate them (RAII).
//
// Reference counterparts:
//   Scene::build                      integration/Scene.cu:56-95
//   PathTracerIntegrator::render      integration/PathTracerIntegrator.cu:146-210
//   Film::addSample / computeRMSE     camera/Film.cu:111-118, 611-632
//   Film::{store,load}ContributionsToFile  camera/Film.cu:366-449 (.hdrbmp text)
//   BVH::intersectRays                cuda_libraries/BVH.cu:541-560
//   materials and lights              materials/*.cu, light/PointLight.cu
#pragma once

#include <cmath>
#include <cstdint>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rt.h"

namespace rt_amd {

class Error : public std::runtime_error {
public:
    Error(rt_status s, const std::string &what) : std::runtime_error(what), status(s) {}
    rt_status status;
};

inline void check(rt_status s)
{
    if (s != RT_OK) {
        const char *m = rt_last_error();
        throw Error(s, m ? m : "rt error");
    }
}

struct Spectrum {
    float r = 0, g = 0, b = 0;
    Spectrum() = default;
    Spectrum(float v) : r(v), g(v), b(v) {}
    Spectrum(float r_, float g_, float b_) : r(r_), g(g_), b(b_) {}
};

struct Point3f {
    float x = 0, y = 0, z = 0;
};

// ---------------------------------------------------------------------
// context and device memory
// ---------------------------------------------------------------------
class Context {
public:
    explicit Context(int device = 0) { check(rt_ctx_create(device, &h_)); }
    ~Context() { if (h_) rt_ctx_destroy(h_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    rt_ctx *get() const { return h_; }
    void synchronize() const { check(rt_ctx_synchronize(h_)); }

private:
    rt_ctx *h_ = nullptr;
};

class DeviceBuffer {
public:
    DeviceBuffer(Context &ctx, size_t bytes) : ctx_(&ctx), bytes_(bytes) { check(rt_malloc(ctx.get(), bytes, &p_)); }
    ~DeviceBuffer() { if (p_) rt_free(ctx_->get(), p_); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    void *get() const { return p_; }
    size_t size() const { return bytes_; }
    void upload(const void *src, size_t n) { check(rt_memcpy_h2d(ctx_->get(), p_, src, n)); }
    void download(void *dst, size_t n) const { check(rt_memcpy_d2h(ctx_->get(), dst, p_, n)); }

private:
    Context *ctx_;
    size_t bytes_;
    void *p_ = nullptr;
};

// ---------------------------------------------------------------------
// materials (materials/*.cu)
// ---------------------------------------------------------------------
// A texture's texels (textures/Texture.cu): width x height, 3 bytes (R, G,
// B) each, row y = the reference's pixelMatrix[y] (BMP file row order).
// include/rt_io.hpp's loadBMP reads one from a file.
struct Texture {
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> rgb;
};

class Material {
public:
    virtual ~Material() = default;
    virtual rt_material desc() const = 0;
    // the texture multiplying kd (Lambertian, Plain), or null
    virtual std::shared_ptr<const Texture> texture() const { return nullptr; }
};

// LambertianMaterial(kd, texture) (materials/LambertianMaterial.cu:18;
// kd times the texel at the hit, LambertianBSDF.cu:22-74)
class LambertianMaterial : public Material {
public:
    explicit LambertianMaterial(const Spectrum &kd, std::shared_ptr<const Texture> tex = nullptr)
        : kd_(kd), tex_(std::move(tex)) {}
    rt_material desc() const override
    {
        rt_material m{};
        m.type = RT_MAT_LAMBERTIAN;
        m.kd[0] = kd_.r; m.kd[1] = kd_.g; m.kd[2] = kd_.b;
        return m;
    }
    std::shared_ptr<const Texture> texture() const override { return tex_; }

private:
    Spectrum kd_;
    std::shared_ptr<const Texture> tex_;
};

// DielectricMaterial(color, ior) (materials/DielectricMaterial.cu:21)
class DielectricMaterial : public Material {
public:
    DielectricMaterial(const Spectrum &color, float ior) : c_(color), ior_(ior) {}
    rt_material desc() const override
    {
        rt_material m{};
        m.type = RT_MAT_DIELECTRIC;
        m.kd[0] = c_.r; m.kd[1] = c_.g; m.kd[2] = c_.b;
        m.ior = ior_;
        return m;
    }

private:
    Spectrum c_;
    float ior_;
};

// GlassMaterial(color) = DielectricMaterial(color, 1.5) (materials/GlassMaterial.cu:12-13)
class GlassMaterial : public DielectricMaterial {
public:
    explicit GlassMaterial(const Spectrum &color) : DielectricMaterial(color, 1.5f) {}
};

// TrowbridgeMaterial(kd, roughness, eta, k): GGX with conductor Fresnel
// (materials/TrowbridgeMaterial.cu:15); alpha = sqrt(roughness)
class TrowbridgeMaterial : public Material {
public:
    TrowbridgeMaterial(const Spectrum &kd, float roughness, const Spectrum &eta, const Spectrum &k)
        : kd_(kd), rough_(roughness), eta_(eta), k_(k) {}
    rt_material desc() const override
    {
        rt_material m{};
        m.type = RT_MAT_TROWBRIDGE;
        m.kd[0] = kd_.r; m.kd[1] = kd_.g; m.kd[2] = kd_.b;
        m.roughness = rough_;
        m.eta[0] = eta_.r; m.eta[1] = eta_.g; m.eta[2] = eta_.b;
        m.k[0] = k_.r; m.k[1] = k_.g; m.k[2] = k_.b;
        return m;
    }

private:
    Spectrum kd_;
    float rough_;
    Spectrum eta_, k_;
};

// TrowbridgeBasicMaterial(kd, roughness): GGX with F = 1
// (materials/TrowbridgeBasicMaterial.cu:15)
class TrowbridgeBasicMaterial : public Material {
public:
    TrowbridgeBasicMaterial(const Spectrum &kd, float roughness) : kd_(kd), rough_(roughness) {}
    rt_material desc() const override
    {
        rt_material m{};
        m.type = RT_MAT_TROWBRIDGE_BASIC;
        m.kd[0] = kd_.r; m.kd[1] = kd_.g; m.kd[2] = kd_.b;
        m.roughness = rough_;
        return m;
    }

private:
    Spectrum kd_;
    float rough_;
};

// LightMaterial(color, intensity): emissive, NullBSDF (materials/LightMaterial.cu:19)
class LightMaterial : public Material {
public:
    LightMaterial(const Spectrum &color, float intensity) : c_(color), i_(intensity) {}
    rt_material desc() const override
    {
        rt_material m{};
        m.type = RT_MAT_LIGHT;
        m.kd[0] = c_.r; m.kd[1] = c_.g; m.kd[2] = c_.b;
        m.intensity = i_;
        return m;
    }

private:
    Spectrum c_;
    float i_;
};

// MirrorMaterial(color): perfect reflection, weight = color, no next-event
// estimation (materials/MirrorMaterial.cu, PerfectReflectionBSDF.cu)
class MirrorMaterial : public Material {
public:
    explicit MirrorMaterial(const Spectrum &color) : c_(color) {}
    rt_material desc() const override
    {
        rt_material m{};
        m.type = RT_MAT_MIRROR;
        m.kd[0] = c_.r; m.kd[1] = c_.g; m.kd[2] = c_.b;
        return m;
    }

private:
    Spectrum c_;
};

// PlainMaterial(kd, ke): emits kd towards camera rays (and rays that reached
// it through specular hits only) and ke towards the others; reflects nothing
// (materials/PlainMaterial.cu:10-80)
class PlainMaterial : public Material {
public:
    explicit PlainMaterial(const Spectrum &kd, const Spectrum &ke = Spectrum(0.0f),
                           std::shared_ptr<const Texture> tex = nullptr)
        : kd_(kd), ke_(ke), tex_(std::move(tex)) {}
    rt_material desc() const override
    {
        rt_material m{};
        m.type = RT_MAT_PLAIN;
        m.kd[0] = kd_.r; m.kd[1] = kd_.g; m.kd[2] = kd_.b;
        m.ke[0] = ke_.r; m.ke[1] = ke_.g; m.ke[2] = ke_.b;
        return m;
    }
    std::shared_ptr<const Texture> texture() const override { return tex_; }

private:
    Spectrum kd_, ke_;
    std::shared_ptr<const Texture> tex_;
};

// ---------------------------------------------------------------------
// shapes and lights
// ---------------------------------------------------------------------
class Shape {
public:
    explicit Shape(std::shared_ptr<Material> m) : material(std::move(m)) {}
    virtual ~Shape() = default;
    std::shared_ptr<Material> material;
};

// A triangle mesh: xyz (3 floats per vertex) and 3 indices per triangle.
// Triangle ids are assigned in mesh order, then face order (BVH.cu:288-291).
class TriangleMesh : public Shape {
public:
    TriangleMesh(std::vector<float> xyz, std::vector<uint32_t> idx, std::shared_ptr<Material> m,
                 std::vector<float> corner_uv = {}, std::vector<float> corner_normals = {})
        : Shape(std::move(m)), vertices(std::move(xyz)), indices(std::move(idx)), uv(std::move(corner_uv)),
          normals(std::move(corner_normals)) {}
    std::vector<float> vertices;
    std::vector<uint32_t> indices;
    std::vector<float> uv;       // 6 per triangle (corner texture coordinates), or empty: the default corners
    // 9 per triangle (corner normals; a NaN first component: none), or empty:
    // the mesh has no vertex normals (TriangleMeshData::normals,
    // shapes/Triangle.cu:379-418: interpolated shading normal, hit point moved
    // 0.15 along it)
    std::vector<float> normals;
};

// Sphere(center, radius) (shapes/Sphere.cu:233-245).  area_light: also
// sampled as a sphere light (needs a LightMaterial).
class Sphere : public Shape {
public:
    Sphere(Point3f c, float r, std::shared_ptr<Material> m, bool area_light = false)
        : Shape(std::move(m)), center(c), radius(r), is_light(area_light) {}
    Point3f center;
    float radius;
    bool is_light;
};

struct Vector3f {
    float x = 0, y = 0, z = 0;
};

// Plane(bottomLeftCorner, axisX, axisY, normal) (shapes/Plane.cu:27-48): the
// parallelogram corner + [0,1] axisX + [0,1] axisY after the reference's
// truncation of the axis lengths to integers; normal picks the side.
class Plane : public Shape {
public:
    Plane(Point3f corner, Vector3f axis_x, Vector3f axis_y, Vector3f normal, std::shared_ptr<Material> m)
        : Shape(std::move(m)), corner(corner), axis_x(axis_x), axis_y(axis_y), normal(normal) {}
    Point3f corner;
    Vector3f axis_x, axis_y, normal;
};

// PointLight(position, color, intensity) (light/PointLight.cu:21)
class PointLight {
public:
    PointLight(Point3f p, const Spectrum &color, float intensity) : pos(p), c(color), I(intensity) {}
    Point3f pos;
    Spectrum c;
    float I;
};

// ---------------------------------------------------------------------
// film (camera/Film.cu)
// ---------------------------------------------------------------------
// Per-pixel radiance sums in FP64 and one sample count, as the reference's
// Film keeps them (the count is tracked for the whole film, Film.cu:112-118).
// ToneMappingType (camera/Film.cu:13-21)
enum ToneMappingType {
    equalize = RT_TONE_EQUALIZE,
    reinhard = RT_TONE_REINHARD,
    reinhardJodie = RT_TONE_REINHARD_JODIE,
    ACESfilmic = RT_TONE_ACES_FILMIC,
    uncharted2 = RT_TONE_UNCHARTED2,
    gammaClamp = RT_TONE_GAMMA_CLAMP
};

class Film {
public:
    Film(uint32_t w, uint32_t h) : w_(w), h_(h), sums_((size_t)w * h * 3, 0.0) {}
    uint32_t width() const { return w_; }
    uint32_t height() const { return h_; }
    uint64_t samples() const { return samples_; }
    std::vector<double> &sums() { return sums_; }
    const std::vector<double> &sums() const { return sums_; }

    // Film::addSample: adds a pixel's radiance sum over nSamples samples
    void addSample(uint32_t x, uint32_t y, const double rgb[3], int nSamples)
    {
        const size_t i = (size_t)y * w_ + x;
        sums_[3 * i] += rgb[0]; sums_[3 * i + 1] += rgb[1]; sums_[3 * i + 2] += rgb[2];
        if (x == 0 && y == 0) samples_ += (uint64_t)nSamples;
    }

    // Adds a device render (FP64 valid-sample sums + valid counts for spp
    // samples per pixel) with the reference's pixel estimate: the sum of the
    // valid samples plus the mean of the valid ones for every invalid sample
    // (integration/pathTracing.cu:474-486).
    void addRender(const std::vector<double> &valid_sums, const std::vector<uint32_t> &nvalid, uint32_t spp)
    {
        for (size_t i = 0; i < (size_t)w_ * h_; ++i) {
            const double s = nvalid[i] ? (double)spp / (double)nvalid[i] : 0.0;
            for (int c = 0; c < 3; ++c) sums_[3 * i + c] += valid_sums[3 * i + c] * s;
        }
        samples_ += spp;
    }

    // storeContributionsToFile (Film.cu:366-394): "P6\n<w> <h>F\n<samples>\n"
    // then "r g b " per pixel, default stream precision.
    void storeContributionsToFile(const std::string &name) const
    {
        std::ofstream f(name + ".hdrbmp", std::ios::binary);
        if (!f) throw Error(RT_ERR_INVALID, "cannot write " + name + ".hdrbmp");
        f << "P6\n" << w_ << " " << h_ << "F\n" << samples_ << "\n";
        for (size_t i = 0; i < (size_t)w_ * h_; ++i)
            f << sums_[3 * i] << " " << sums_[3 * i + 1] << " " << sums_[3 * i + 2] << " ";
        if (!f) throw Error(RT_ERR_INVALID, "short write to " + name + ".hdrbmp");
    }

    // loadContributionsFromFile (Film.cu:396-449): ADDS the file's sums and
    // sample count.  Returns false if the file does not exist; throws on a
    // resolution mismatch or a truncated file.
    bool loadContributionsFromFile(const std::string &name)
    {
        std::ifstream f(name + ".hdrbmp", std::ios::binary);
        if (!f.is_open()) return false;
        std::string header;
        long long w = 0, h = 0, n = 0;
        char type = 0;
        f >> header >> w >> h >> type >> n;
        if (!f || header != "P6") throw Error(RT_ERR_INVALID, name + ".hdrbmp: bad header");
        if (w != (long long)w_ || h != (long long)h_)
            throw Error(RT_ERR_INVALID, name + ".hdrbmp: resolution does not match the film");
        std::vector<double> add((size_t)w_ * h_ * 3);
        for (double &v : add)
            if (!(f >> v)) throw Error(RT_ERR_INVALID, name + ".hdrbmp: truncated pixel data");
        for (size_t i = 0; i < add.size(); ++i) sums_[i] += add[i];
        samples_ += (uint64_t)n;
        return true;
    }

    // Film::writeToBMP / writeToPPM (camera/Film.cu:339-362): tone-mapped on
    // GPU `device` and written to filename + ".bmp" / ".ppm".
    void writeToBMP(const std::string &filename, ToneMappingType toneMode, int device = 0) const
    {
        writeImage(filename + ".bmp", RT_IMAGE_BMP, toneMode, device);
    }
    void writeToPPM(const std::string &filename, ToneMappingType toneMode, int device = 0) const
    {
        writeImage(filename + ".ppm", RT_IMAGE_PPM, toneMode, device);
    }

    // The image file's bytes (header, then the pixel array rt_film_tonemap
    // writes from these sums on the context's GPU).
    std::vector<uint8_t> toneMap(Context &ctx, rt_image_format format, ToneMappingType toneMode) const
    {
        const size_t hb = rt_image_header(format, w_, h_, nullptr, 0);
        const size_t pb = rt_image_pixel_bytes(format, w_, h_);
        if (!hb || !pb) throw Error(RT_ERR_INVALID, "film size outside 1..32768 per side");
        std::vector<uint8_t> out(hb + pb);
        rt_image_header(format, w_, h_, out.data(), hb);
        DeviceBuffer sums(ctx, sums_.size() * sizeof(double)), px(ctx, pb);
        sums.upload(sums_.data(), sums_.size() * sizeof(double));
        check(rt_film_tonemap(ctx.get(), static_cast<const double *>(sums.get()), nullptr, w_, h_, samples_,
                              static_cast<rt_tone_mapping>(toneMode), format, static_cast<uint8_t *>(px.get()), pb,
                              nullptr));
        px.download(out.data() + hb, pb);
        return out;
    }

private:
    void writeImage(const std::string &path, rt_image_format format, ToneMappingType toneMode, int device) const
    {
        Context ctx(device);
        const std::vector<uint8_t> bytes = toneMap(ctx, format, toneMode);
        std::ofstream f(path, std::ios::binary);
        if (!f) throw Error(RT_ERR_INVALID, "cannot write " + path);
        f.write(reinterpret_cast<const char *>(bytes.data()), (std::streamsize)bytes.size());
        if (!f) throw Error(RT_ERR_INVALID, "short write to " + path);
    }

    uint32_t w_, h_;
    std::vector<double> sums_;
    uint64_t samples_ = 0;
};

// computeRMSE (camera/Film.cu:611-632): sqrt of the mean over pixels of the
// squared norm (RGB channel sum) of the difference of mean radiances; -1 on a
// resolution mismatch.  Accumulated in float like the reference.
inline float computeRMSE(const Film &a, const Film &b)
{
    if (a.width() != b.width() || a.height() != b.height()) return -1;
    const size_t n = (size_t)a.width() * a.height();
    float acc = 0;
    for (size_t i = 0; i < n; ++i) {
        double d2 = 0;
        for (int c = 0; c < 3; ++c) {
            const double d = a.sums()[3 * i + c] / (double)a.samples() - b.sums()[3 * i + c] / (double)b.samples();
            d2 += d * d;
        }
        acc += (float)d2;
    }
    return std::sqrt(acc / (float)n);
}

// ProjectiveCamera as loaded from a .obejota file (FileImporting.cu:404-436):
// position, look-at point, up, camera-to-screen translation cts, screen
// window height frameY, resolution; owns its film.
class ProjectiveCamera {
public:
    ProjectiveCamera(Point3f pos, Point3f lookat, Point3f up, Point3f cts, float frame_y, uint32_t w, uint32_t h)
        : film(w, h)
    {
        c_.pos[0] = pos.x; c_.pos[1] = pos.y; c_.pos[2] = pos.z;
        c_.lookat[0] = lookat.x; c_.lookat[1] = lookat.y; c_.lookat[2] = lookat.z;
        c_.up[0] = up.x; c_.up[1] = up.y; c_.up[2] = up.z;
        c_.cts[0] = cts.x; c_.cts[1] = cts.y; c_.cts[2] = cts.z;
        c_.frame_y = frame_y;
        c_.width = w;
        c_.height = h;
    }
    const rt_camera &desc() const { return c_; }
    Film film;

private:
    rt_camera c_{};
};

// ---------------------------------------------------------------------
// scene (integration/Scene.cu)
// ---------------------------------------------------------------------
class Scene {
public:
    explicit Scene(Context &ctx) : ctx_(&ctx) { check(rt_scene_create(ctx.get(), &h_)); }
    ~Scene() { if (h_) rt_scene_destroy(h_); }
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;

    // The BVH the next build() makes: on the GPU from the scene's triangles
    // (rt_bvh_build_device, as the reference builds its tree) or, by
    // default, the host's binned SAH.
    void setBvh(bool on_device, uint32_t max_leaf = 4, uint32_t width = 4)
    {
        rt_bvh_opts o{max_leaf, 32, 1.2f, width};
        check(rt_scene_set_bvh(h_, &o, on_device ? 1u : 0u));
    }

    // Scene::build(shapes, lights, ...): materials are registered once per
    // distinct object; meshes first in the given order (triangle ids follow
    // it), then spheres, then planes, each in the given order (the analytic
    // shapes' intersection order, Scene.cu:221-227).
    void build(const std::vector<std::shared_ptr<Shape>> &shapes, const std::vector<PointLight> &lights,
               const ProjectiveCamera &camera)
    {
        std::vector<std::pair<const Material *, uint32_t>> ids;
        std::vector<std::pair<const Texture *, uint32_t>> tex_ids;
        auto mat_id = [&](const std::shared_ptr<Material> &m) {
            for (auto &p : ids)
                if (p.first == m.get()) return p.second;
            rt_material d = m->desc();
            if (auto t = m->texture()) {
                uint32_t tid = 0;
                bool found = false;
                for (auto &p : tex_ids)
                    if (p.first == t.get()) { tid = p.second; found = true; }
                if (!found) {
                    check(rt_scene_add_texture(h_, t->rgb.data(), t->width, t->height, &tid));
                    tex_ids.emplace_back(t.get(), tid);
                }
                d.texture = tid + 1;
            }
            uint32_t id = 0;
            check(rt_scene_add_material(h_, &d, &id));
            ids.emplace_back(m.get(), id);
            return id;
        };
        for (const auto &s : shapes)
            if (auto *t = dynamic_cast<const TriangleMesh *>(s.get())) {
                const uint32_t m = mat_id(s->material);
                const uint32_t nt = (uint32_t)(t->indices.size() / 3);
                if (!t->uv.empty() && t->uv.size() != 6 * (size_t)nt)
                    throw Error(RT_ERR_INVALID, "TriangleMesh: uv needs 6 floats per triangle");
                if (!t->normals.empty() && t->normals.size() != 9 * (size_t)nt)
                    throw Error(RT_ERR_INVALID, "TriangleMesh: normals need 9 floats per triangle");
                check(rt_scene_add_mesh_normals(h_, t->vertices.data(), (uint32_t)(t->vertices.size() / 3),
                                                t->indices.data(), nt, t->uv.empty() ? nullptr : t->uv.data(),
                                                t->normals.empty() ? nullptr : t->normals.data(), m, nullptr));
            }
        for (const auto &s : shapes)
            if (auto *sp = dynamic_cast<const Sphere *>(s.get())) {
                const uint32_t m = mat_id(s->material);
                const float c[3] = {sp->center.x, sp->center.y, sp->center.z};
                check(rt_scene_add_sphere(h_, c, sp->radius, m, sp->is_light ? 1 : 0));
            }
        for (const auto &s : shapes)
            if (auto *pl = dynamic_cast<const Plane *>(s.get())) {
                const uint32_t m = mat_id(s->material);
                const float c[3] = {pl->corner.x, pl->corner.y, pl->corner.z};
                const float ax[3] = {pl->axis_x.x, pl->axis_x.y, pl->axis_x.z};
                const float ay[3] = {pl->axis_y.x, pl->axis_y.y, pl->axis_y.z};
                const float n[3] = {pl->normal.x, pl->normal.y, pl->normal.z};
                check(rt_scene_add_plane(h_, c, ax, ay, n, m));
            }
        for (const auto &l : lights) {
            rt_point_light pl{};
            pl.pos[0] = l.pos.x; pl.pos[1] = l.pos.y; pl.pos[2] = l.pos.z;
            pl.color[0] = l.c.r; pl.color[1] = l.c.g; pl.color[2] = l.c.b;
            pl.intensity = l.I;
            check(rt_scene_add_point_light(h_, &pl));
        }
        check(rt_scene_set_camera(h_, &camera.desc()));
        check(rt_scene_commit(h_));
    }

    rt_scene *get() const { return h_; }
    Context &context() const { return *ctx_; }

private:
    Context *ctx_;
    rt_scene *h_ = nullptr;
};

// Render settings of the reference's compile-time macros (types.h:41-42,73-74).
// RENDER_MODE (cuda_libraries/types.h:51) as a run-time setting; the
// reference's MACHINE_LEARNING mode is not offered.
enum class RenderMode { PATH_TRACING, PHOTON_MAPPING, HYBRID };

struct RenderSettings {
    uint32_t spp = 16, max_depth = 8, rr_depth = 3, seed = 1, spp_offset = 0, path_capacity = 0, grid = 0;
    uint32_t chunk_paths = 0;  // paths per streamed chunk, 0 = the library's default
    RenderMode mode = RenderMode::PATH_TRACING;
    // photon-mapping modes (types.h:57-60): photon paths traced before the
    // render, their depth and roulette, and the search's k and radius
    uint32_t photons = 1u << 20, photon_max_depth = 30, photon_rr_depth = 3, photon_seed = 1;
    uint32_t neighbours = 100;
    float search_radius = 100.0f;
};

// PathTracerIntegrator(camera).render(scene): renders every pixel at
// settings.spp samples on the GPU and adds the result to camera->film.
class PathTracerIntegrator {
public:
    explicit PathTracerIntegrator(ProjectiveCamera *camera) : cam_(camera) {}

    void render(Scene &scene, const RenderSettings &s = RenderSettings())
    {
        Context &ctx = scene.context();
        const size_t npix = (size_t)cam_->film.width() * cam_->film.height();
        DeviceBuffer accum(ctx, npix * 3 * sizeof(double)), nvalid(ctx, npix * sizeof(uint32_t));
        std::vector<double> zeros(npix * 3, 0.0);
        std::vector<uint32_t> zc(npix, 0);
        accum.upload(zeros.data(), zeros.size() * sizeof(double));
        nvalid.upload(zc.data(), zc.size() * sizeof(uint32_t));
        rt_render_opts o{};
        o.spp = s.spp;
        o.spp_offset = s.spp_offset;
        o.max_depth = s.max_depth;
        o.rr_depth = s.rr_depth;
        o.seed = s.seed;
        o.path_capacity = s.path_capacity;
        o.grid = s.grid;
        o.chunk_paths = s.chunk_paths;
        if (s.mode == RenderMode::PATH_TRACING) {
            check(rt_render(scene.get(), &o, (double *)accum.get(), (uint32_t *)nvalid.get(), 0, nullptr));
        } else {
            renderPhotonMapping(scene, o, s, accum, nvalid);
        }
        ctx.synchronize();
        std::vector<double> sums(npix * 3);
        std::vector<uint32_t> nv(npix);
        accum.download(sums.data(), sums.size() * sizeof(double));
        nvalid.download(nv.data(), nv.size() * sizeof(uint32_t));
        cam_->film.addRender(sums, nv, s.spp);
    }

private:
    // PathTracerIntegrator::render with RENDER_MODE PHOTON_MAPPING or HYBRID
    // (PathTracerIntegrator.cu:157-161): generatePhotons on the GPU
    // (rt_scene_trace_photons, growing the arrays once if they were too
    // small), buildPhotonStructures (rt_photon_map_build over the photons'
    // locations and shapes), then the bounce loop with the map.
    void renderPhotonMapping(Scene &scene, const rt_render_opts &o, const RenderSettings &s, DeviceBuffer &accum,
                             DeviceBuffer &nvalid)
    {
        Context &ctx = scene.context();
        const rt_photon_opts po{s.photons, s.photon_max_depth, s.photon_rr_depth, s.photon_seed};
        uint32_t cap = s.photons * 4 + 1024, n = 0;
        for (int attempt = 0;; ++attempt) {
            DeviceBuffer loc(ctx, (size_t)cap * 12), dir(ctx, (size_t)cap * 12), nrm(ctx, (size_t)cap * 12),
                rgb(ctx, (size_t)cap * 24), shp(ctx, (size_t)cap * 4);
            const rt_status st = rt_scene_trace_photons(scene.get(), &po, cap, (float *)loc.get(), (float *)dir.get(),
                                                        (float *)nrm.get(), (double *)rgb.get(), (int32_t *)shp.get(),
                                                        nullptr, &n, nullptr);
            if (st == RT_ERR_INVALID && n > cap && attempt == 0) {
                cap = n;
                continue;
            }
            check(st);
            std::vector<float> hloc((size_t)n * 3);
            std::vector<int32_t> hshp(n);
            if (n) {
                loc.download(hloc.data(), hloc.size() * sizeof(float));
                shp.download(hshp.data(), hshp.size() * sizeof(int32_t));
            }
            rt_photon_map *map = nullptr;
            check(rt_photon_map_build(ctx.get(), hloc.data(), hshp.data(), n, &map));
            const rt_photon_render_opts ro{s.mode == RenderMode::HYBRID ? (uint32_t)RT_RENDER_HYBRID
                                                                        : (uint32_t)RT_RENDER_PHOTON_MAPPING,
                                           s.neighbours, s.search_radius};
            const rt_status rs = rt_render_photon(scene.get(), &o, map, (const double *)rgb.get(),
                                                  (const float *)dir.get(), &ro, (double *)accum.get(),
                                                  (uint32_t *)nvalid.get(), 0, nullptr);
            if (rs == RT_OK) ctx.synchronize();
            rt_photon_map_destroy(map);
            check(rs);
            return;
        }
    }

public:

private:
    ProjectiveCamera *cam_;
};

// ---------------------------------------------------------------------
// BVH (cuda_libraries/BVH.cu): a tree over a triangle soup and the batch
// closest-hit call on host arrays (BVH::intersectRays)
// ---------------------------------------------------------------------
class BVH {
public:
    // ctx may be null: host-only tree (info and export only)
    BVH(Context *ctx, const std::vector<float> &xyz, const std::vector<uint32_t> &idx, uint32_t max_leaf = 4)
    {
        rt_bvh_opts o{max_leaf, 32, 1.2f, 4};
        check(rt_bvh_build(ctx ? ctx->get() : nullptr, xyz.data(), (uint32_t)(xyz.size() / 3), idx.data(),
                           (uint32_t)(idx.size() / 3), &o, &h_));
    }
    // built on the GPU from device arrays (rt_bvh_build_device): xyz holds
    // nv x 3 floats and idx nt x 3 indices on ctx's device
    struct OnDevice {};
    BVH(OnDevice, Context &ctx, const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt,
        uint32_t max_leaf = 4, uint32_t width = 4)
    {
        rt_bvh_opts o{max_leaf, 32, 1.2f, width};
        check(rt_bvh_build_device(ctx.get(), xyz, nv, idx, nt, &o, &h_));
    }
    ~BVH() { if (h_) rt_bvh_destroy(h_); }
    BVH(const BVH &) = delete;
    BVH &operator=(const BVH &) = delete;

    rt_bvh_info info() const
    {
        rt_bvh_info i{};
        check(rt_bvh_get_info(h_, &i));
        return i;
    }

    // SoA host rays (o, d, tMax) -> hit ids (-1 = miss) and distances.
    void intersectRays(const std::vector<float> &ox, const std::vector<float> &oy, const std::vector<float> &oz,
                       const std::vector<float> &dx, const std::vector<float> &dy, const std::vector<float> &dz,
                       const std::vector<float> &tmax, std::vector<int32_t> &hitIds, std::vector<float> &t) const
    {
        const size_t n = ox.size();
        hitIds.resize(n);
        t.resize(n);
        rt_rays r{ox.data(), oy.data(), oz.data(), dx.data(), dy.data(), dz.data(), tmax.data()};
        check(rt_trace_closest(h_, &r, (uint32_t)n, hitIds.data(), t.data(), RT_FLAG_HOST_PTRS, nullptr));
    }

    rt_bvh *get() const { return h_; }

private:
    rt_bvh *h_ = nullptr;
};

}  // namespace rt_amd
