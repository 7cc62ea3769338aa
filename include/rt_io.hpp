// rt_io.hpp -- the reference renderer's scene files for the C++ host API
// (rt.hpp): Wavefront OBJ meshes with their MTL materials, and the .obejota
// camera / point-light files exported from Blender.  Header-only, host only
// (no GPU call until the shapes are given to Scene::build).
//
// Reference counterparts:
//   loadMTL                         cuda_libraries/FileImporting.cu:27-246
//   loadObj                         cuda_libraries/FileImporting.cu:248-372
//   storeObejotaObject, loadObejota cuda_libraries/FileImporting.cu:375-598
//
// Conventions kept from the reference:
//   - OBJ `v` and `vn` lines are read as (x, z, y): the file's second and
//     third coordinates are swapped (Blender's y-up export to the renderer's
//     z-up frame, FileImporting.cu:287-297);
//   - .obejota `p`, `lookat`, `cts` and `up` have their y negated
//     (FileImporting.cu:490-556); `rotation ax ay az` (degrees, Blender's
//     XYZ Euler angles) sets lookat = p + Rz(-az) Ry(-ay) Rx(-ax) (0, 0, -1)
//     and rotates the camera's up vector by the same matrix when the camera
//     is stored (:520-534, :393-394);
//   - faces are triangles only, indices 1-based; a face's material is the
//     last `usemtl` (the built-in "default" is Lambertian(1)); every `o`
//     line starts a new mesh id from a counter that starts at -1 (types.h:20);
//   - `newmtl` resets Ka, Kd, Ks, Ni, Ns, illum, the texture, makesShadow,
//     hasIndirect, the roughnesses, eta and k, but not Ke, sigmaS, sigmaA
//     or g (FileImporting.cu:84-98); an `mtllib` file's materials are merged
//     without replacing names already known (std::unordered_map::insert);
//   - material classes by `illum`: 0 Plain, 1 Lambertian, 2 Phong,
//     3 Dielectric (colour 1, IOR Ni), 4 Beckmann, 5 Trowbridge,
//     6 TrowbridgeBasic, 7 subsurface (FileImporting.cu:65-81).
// Where the reference reads garbage or crashes (a missing number, an index
// out of range, an illum it does not know) these loaders throw rt_amd::Error
// with RT_ERR_INVALID and the file and line.  Materials the GPU path does not
// render (Phong, Beckmann, subsurface, textures on other than Lambertian
// and Plain, anisotropic roughness, shadow-transparent Lambertian or Plain)
// load as data and are refused with RT_ERR_UNSUPPORTED only when a shape is
// made from them.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "rt.hpp"

namespace rt_amd {

// One `newmtl` block of an MTL file.
struct MtlMaterial {
    std::string name;
    int illum = 0;
    Spectrum ka{0.0f}, kd{1.0f}, ks{0.0f}, ke{0.0f};
    float ni = 1.0f, ns = 0.0f;
    float rough_x = 0.0f, rough_y = 0.0f;  // Pr sets both, PrX / PrY one each
    Spectrum eta{1.0f}, k{0.0f};
    bool transparent = false;   // makesShadow 0
    bool has_indirect = true;   // hasIndirect
    Spectrum sigma_s{0.0f}, sigma_a{0.0f};
    float g = 0.0f;
    std::string map_kd;         // texture path (the MTL's directory, else its textures/ subdirectory)
};

// One OBJ face: 0-based indices, -1 where the face gives none.
struct ObjTriangle {
    int32_t v[3] = {-1, -1, -1};
    int32_t vt[3] = {-1, -1, -1};
    int32_t vn[3] = {-1, -1, -1};
    int32_t mesh_id = -1;
    std::string material;
};

struct ObjData {
    std::vector<Point3f> vertices;                  // (x, z, y) of the file
    std::vector<Point3f> normals;                   // (x, z, y) of the file
    std::vector<std::pair<float, float>> texcoords;
    std::vector<ObjTriangle> triangles;             // face order = triangle id order
    std::map<std::string, MtlMaterial> materials;   // "default" and every mtllib's
    int32_t next_mesh_id = -1;                      // the mesh counter after this file
};

struct ObejotaCamera {
    Point3f pos, lookat, up, cts;
    float frame_y = 0.0f;
};

struct ObejotaScene {
    bool has_camera = false;
    ObejotaCamera camera;               // the last camera of the file
    std::vector<PointLight> lights;     // in file order
    std::vector<std::string> warnings;  // also printed to stderr, as the reference does
};

namespace io_detail {

inline std::string where(const std::string &file, size_t line)
{
    return file + ":" + std::to_string(line) + ": ";
}

inline float read_float(std::istringstream &ss, const std::string &file, size_t line, const std::string &key)
{
    float v = 0.0f;
    if (!(ss >> v)) throw Error(RT_ERR_INVALID, where(file, line) + "expected a number after '" + key + "'");
    return v;
}

inline Spectrum read_rgb(std::istringstream &ss, const std::string &file, size_t line, const std::string &key)
{
    const float r = read_float(ss, file, line, key);
    const float g = read_float(ss, file, line, key);
    const float b = read_float(ss, file, line, key);
    return Spectrum(r, g, b);
}

inline std::string dir_of(const std::string &path)
{
    const size_t p = path.find_last_of("/\\");
    return p == std::string::npos ? std::string() : path.substr(0, p + 1);
}

inline bool readable(const std::string &path)
{
    std::ifstream f(path);
    return f.good();
}

inline bool flag01(std::istringstream &ss, const std::string &file, size_t line, const std::string &key)
{
    std::string t;
    ss >> t;
    if (t == "1") return true;
    if (t == "0") return false;
    throw Error(RT_ERR_INVALID, where(file, line) + key + " must be 0 or 1");
}

// 3x3 rotation about a unit axis by theta degrees (the Rodrigues matrix of
// transformations/Transform.cu:580-610), row-major, applied as m * v.
struct Mat3 {
    float m[3][3];
};

inline Mat3 rotation(float theta_deg, float ax, float ay, float az)
{
    const float pi = 3.14159265358979323846f;
    const float s = std::sin(theta_deg * (pi / 180.0f)), c = std::cos(theta_deg * (pi / 180.0f));
    Mat3 r{};
    r.m[0][0] = ax * ax + (1 - ax * ax) * c;
    r.m[0][1] = ax * ay * (1 - c) - az * s;
    r.m[0][2] = ax * az * (1 - c) + ay * s;
    r.m[1][0] = ax * ay * (1 - c) + az * s;
    r.m[1][1] = ay * ay + (1 - ay * ay) * c;
    r.m[1][2] = ay * az * (1 - c) - ax * s;
    r.m[2][0] = ax * az * (1 - c) - ay * s;
    r.m[2][1] = ay * az * (1 - c) + ax * s;
    r.m[2][2] = az * az + (1 - az * az) * c;
    return r;
}

inline Mat3 mul(const Mat3 &a, const Mat3 &b)
{
    Mat3 r{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}

inline Point3f apply(const Mat3 &a, const Point3f &v)
{
    return Point3f{a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z,
                   a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
                   a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}

inline bool is_zero(const Point3f &p) { return p.x == 0.0f && p.y == 0.0f && p.z == 0.0f; }

inline void warn(ObejotaScene &sc, const std::string &w)
{
    std::cerr << "Warning: " << w << std::endl;
    sc.warnings.push_back(w);
}

}  // namespace io_detail

// Texture::Texture's file name (textures/Texture.cu:127-149): a .jpg or .png
// name becomes .bmp (with the reference's warning on stdout), then every '.'
// after the first character except the last one becomes '_'.
inline std::string texture_file_name(std::string path, bool quiet = false)
{
    const size_t dot = path.find_last_of('.');
    const std::string ext = path.substr(dot == std::string::npos ? 0 : dot + 1);
    if (ext == "jpg" || ext == "png") {
        if (!quiet) std::cout << "Warning: " << path << " is not a bmp file. Changing extension to bmp." << std::endl;
        path = path.substr(0, dot) + ".bmp";
    }
    const size_t last = path.find_last_of('.');
    for (size_t i = 1; i < path.size(); ++i)
        if (path[i] == '.' && i != last) path[i] = '_';
    return path;
}

// loadMTL (FileImporting.cu:27-246).  A file that cannot be opened gives an
// empty map and a message on stderr, as in the reference.
inline std::map<std::string, MtlMaterial> loadMTL(const std::string &path)
{
    using namespace io_detail;
    std::map<std::string, MtlMaterial> out;
    std::ifstream f(path);
    if (!f) {
        std::cerr << "MTL file not found: " << path << std::endl;
        return out;
    }
    MtlMaterial cur;  // Ke, sigmaS, sigmaA and g carry over between blocks
    bool open = false;
    std::string line;
    size_t ln = 0;
    auto store = [&]() {
        if (open && !cur.name.empty()) out[cur.name] = cur;
    };
    while (std::getline(f, line)) {
        ++ln;
        std::istringstream ss(line);
        std::string key;
        ss >> key;
        if (key == "newmtl") {
            store();
            cur.ka = Spectrum(0.0f);
            cur.kd = Spectrum(1.0f);
            cur.ks = Spectrum(0.0f);
            cur.ni = 1.0f;
            cur.ns = 0.0f;
            cur.illum = 0;
            cur.map_kd.clear();
            cur.transparent = false;
            cur.has_indirect = true;
            cur.rough_x = cur.rough_y = 0.0f;
            cur.eta = Spectrum(1.0f);
            cur.k = Spectrum(0.0f);
            ss >> cur.name;  // a missing name keeps the previous one, as in the reference
            open = true;
        } else if (key == "Ka") {
            cur.ka = read_rgb(ss, path, ln, key);
        } else if (key == "Kd") {
            cur.kd = read_rgb(ss, path, ln, key);
        } else if (key == "Ks") {
            cur.ks = read_rgb(ss, path, ln, key);
        } else if (key == "Ke") {
            cur.ke = read_rgb(ss, path, ln, key);
        } else if (key == "Ns") {
            cur.ns = read_float(ss, path, ln, key);
        } else if (key == "Ni") {
            cur.ni = read_float(ss, path, ln, key);
        } else if (key == "Pr") {
            cur.rough_x = cur.rough_y = read_float(ss, path, ln, key);
        } else if (key == "PrX") {
            cur.rough_x = read_float(ss, path, ln, key);
        } else if (key == "PrY") {
            cur.rough_y = read_float(ss, path, ln, key);
        } else if (key == "eta") {
            cur.eta = read_rgb(ss, path, ln, key);
        } else if (key == "k") {
            cur.k = read_rgb(ss, path, ln, key);
        } else if (key == "illum") {
            int v = 0;
            if (!(ss >> v)) throw Error(RT_ERR_INVALID, where(path, ln) + "expected an integer after 'illum'");
            if (v < 0 || v > 7) throw Error(RT_ERR_INVALID, where(path, ln) + "illum must be 0..7");
            cur.illum = v;
        } else if (key == "makesShadow") {
            cur.transparent = !flag01(ss, path, ln, key);
        } else if (key == "hasIndirect") {
            cur.has_indirect = flag01(ss, path, ln, key);
        } else if (key == "sigmaS") {
            cur.sigma_s = read_rgb(ss, path, ln, key);
        } else if (key == "sigmaA") {
            cur.sigma_a = read_rgb(ss, path, ln, key);
        } else if (key == "g") {
            cur.g = read_float(ss, path, ln, key);
        } else if (key == "map_Kd") {
            // the reference tries Texture(dir + name) and, if that file cannot
            // be opened, Texture(dir + "textures/" + name) (FileImporting.cu:
            // 213-235); the file tried is the one texture_file_name gives
            std::string name;
            ss >> name;
            const std::string dir = dir_of(path);
            cur.map_kd = readable(texture_file_name(dir + name, true)) ? dir + name : dir + "textures/" + name;
        }
    }
    store();
    return out;
}

// loadObj (FileImporting.cu:248-372).  first_mesh_id is the mesh counter
// before this file (-1 for the first file, then the previous file's
// next_mesh_id).
inline ObjData loadObj(const std::string &path, int32_t first_mesh_id = -1)
{
    using namespace io_detail;
    std::ifstream f(path);
    if (!f.is_open()) throw Error(RT_ERR_INVALID, "failed to open the OBJ file " + path);
    ObjData o;
    MtlMaterial def;
    def.name = "default";
    def.illum = 1;
    def.kd = Spectrum(1.0f);
    o.materials["default"] = def;
    int32_t mesh_id = first_mesh_id;
    std::string material = "default";
    std::string line;
    size_t ln = 0;
    while (std::getline(f, line)) {
        ++ln;
        std::istringstream ss(line);
        std::string key;
        ss >> key;
        if (key == "o") {
            ++mesh_id;
        } else if (key == "v" || key == "vn") {
            const float x = read_float(ss, path, ln, key);
            const float z = read_float(ss, path, ln, key);
            const float y = read_float(ss, path, ln, key);
            (key == "v" ? o.vertices : o.normals).push_back(Point3f{x, y, z});
        } else if (key == "vt") {
            const float u = read_float(ss, path, ln, key);
            const float v = read_float(ss, path, ln, key);
            o.texcoords.emplace_back(u, v);
        } else if (key == "f") {
            ObjTriangle t;
            t.mesh_id = mesh_id;
            for (int i = 0; i < 3; ++i) {
                std::string ref;
                if (!(ss >> ref)) throw Error(RT_ERR_INVALID, where(path, ln) + "a face needs 3 vertices");
                // v, v/t, v//n or v/t/n
                int32_t part[3] = {0, 0, 0};
                int slot = 0;
                std::string num;
                for (size_t c = 0; c <= ref.size(); ++c) {
                    if (c == ref.size() || ref[c] == '/') {
                        if (slot > 2) throw Error(RT_ERR_INVALID, where(path, ln) + "bad face vertex '" + ref + "'");
                        if (!num.empty()) {
                            char *end = nullptr;
                            const long v = std::strtol(num.c_str(), &end, 10);
                            if (*end != '\0' || v < 1 || v > INT32_MAX)
                                throw Error(RT_ERR_INVALID, where(path, ln) + "bad face index '" + ref + "'");
                            part[slot] = (int32_t)v;
                        } else if (slot == 0) {
                            throw Error(RT_ERR_INVALID, where(path, ln) + "bad face vertex '" + ref + "'");
                        }
                        num.clear();
                        ++slot;
                    } else {
                        num += ref[c];
                    }
                }
                t.v[i] = part[0] - 1;
                t.vt[i] = part[1] - 1;
                t.vn[i] = part[2] - 1;
            }
            if (o.materials.find(material) == o.materials.end())
                throw Error(RT_ERR_INVALID,
                            where(path, ln) + "the OBJ file has a face with an unknown material " + material);
            std::string extra;
            if (ss >> extra)
                throw Error(RT_ERR_INVALID, where(path, ln) + "the OBJ file has polygons, it should only have triangles");
            t.material = material;
            o.triangles.push_back(t);
        } else if (key == "usemtl") {
            std::string m;
            ss >> m;
            material = m;  // a bare usemtl selects "", which no face may use
        } else if (key == "mtllib") {
            std::string name, rest;
            ss >> name;
            if (ss >> rest) throw Error(RT_ERR_INVALID, where(path, ln) + "the OBJ file has a material file name with spaces");
            for (auto &kv : loadMTL(dir_of(path) + name)) o.materials.insert(kv);  // known names are kept
        }
    }
    if (o.triangles.empty()) throw Error(RT_ERR_INVALID, path + ": the OBJ file has no triangles");
    const int32_t nv = (int32_t)o.vertices.size(), nn = (int32_t)o.normals.size(), nt = (int32_t)o.texcoords.size();
    for (const auto &t : o.triangles)
        for (int i = 0; i < 3; ++i)
            if (t.v[i] >= nv || t.vn[i] >= nn || t.vt[i] >= nt)
                throw Error(RT_ERR_INVALID, path + ": a face refers to a vertex, normal or texture coordinate the file does not have");
    o.next_mesh_id = mesh_id;
    return o;
}

// loadObejota with storeObejotaObject (FileImporting.cu:375-598).
inline ObejotaScene loadObejota(const std::string &path)
{
    using namespace io_detail;
    std::ifstream f(path);
    if (!f.is_open()) throw Error(RT_ERR_INVALID, "failed to open the obejota file " + path);
    ObejotaScene sc;
    std::string type, name;
    Point3f pos, lookat, cts, up;
    float intensity = 0.0f, frame_y = 0.0f;
    Spectrum color(0.0f);
    Mat3 rot{};
    bool rotated = false;  // like the reference, kept for the rest of the file once set
    auto store = [&]() {
        if (type == "camera") {
            if (rotated) up = apply(rot, up);
            if (is_zero(cts)) throw Error(RT_ERR_INVALID, path + ": camera " + name + " has no camera to screen vector");
            if (is_zero(up)) throw Error(RT_ERR_INVALID, path + ": camera " + name + " has no up vector");
            if (frame_y == 0.0f) throw Error(RT_ERR_INVALID, path + ": camera " + name + " has no frame size");
            sc.has_camera = true;
            sc.camera = ObejotaCamera{pos, lookat, up, cts, frame_y};
        } else if (type == "lightPoint") {
            if (color.r == 0.0f && color.g == 0.0f && color.b == 0.0f) warn(sc, "Light " + name + " has no color.");
            if (intensity == 0.0f) warn(sc, "Light " + name + " has no intensity.");
            sc.lights.emplace_back(pos, color, intensity);
        }
        type.clear();
        name.clear();
        pos = lookat = cts = up = Point3f{};
        intensity = 0.0f;
        color = Spectrum(0.0f);
        frame_y = 0.0f;
    };
    bool camera_found = false;
    std::string line;
    size_t ln = 0;
    while (std::getline(f, line)) {
        ++ln;
        std::istringstream ss(line);
        std::string key;
        ss >> key;
        if (key == "lightPoint") {
            if (!name.empty()) store();
            ss >> name;
            type = "lightPoint";
        } else if (key == "camera") {
            if (!name.empty()) store();
            name = "camera";
            type = "camera";
            camera_found = true;
        } else if (key == "p" || key == "lookat" || key == "cts" || key == "up") {
            Point3f v;
            v.x = read_float(ss, path, ln, key);
            v.y = -read_float(ss, path, ln, key);  // Blender's y axis is inverted
            v.z = read_float(ss, path, ln, key);
            (key == "p" ? pos : key == "lookat" ? lookat : key == "cts" ? cts : up) = v;
        } else if (key == "i") {
            intensity = read_float(ss, path, ln, key);
        } else if (key == "c") {
            color = read_rgb(ss, path, ln, key);
        } else if (key == "rotation") {
            const float ax = read_float(ss, path, ln, key);
            const float ay = read_float(ss, path, ln, key);
            const float az = read_float(ss, path, ln, key);
            rot = mul(rotation(-az, 0, 0, 1), mul(rotation(-ay, 0, 1, 0), rotation(-ax, 1, 0, 0)));
            const Point3f local = apply(rot, Point3f{0.0f, 0.0f, -1.0f});
            lookat = Point3f{pos.x + local.x, pos.y + local.y, pos.z + local.z};
            rotated = true;
        } else if (key == "frameY") {
            frame_y = read_float(ss, path, ln, key);
        } else if (!key.empty() && key != "#") {
            warn(sc, "Unknown token in obejota file: " + key);
        }
    }
    if (!name.empty()) store();
    if (!camera_found) warn(sc, "No camera found in obejota file.");
    return sc;
}

// Texture::loadBmpImage (textures/Texture.cu:40-119): the 54-byte header,
// 24- or 32-bit pixels read from the header's offset as width x height
// consecutive B, G, R(, A) groups -- rows are not skipped to 4-byte
// boundaries, as in the reference -- row y of the file is pixelMatrix[y].
// Where the reference would read past the end or size a negative height,
// this throws.
inline Texture loadBMP(const std::string &name)
{
    const std::string path = texture_file_name(name);
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(RT_ERR_INVALID, "Error loading " + path + ": Failed to open the texture file");
    unsigned char h[54];
    if (!f.read(reinterpret_cast<char *>(h), 54)) throw Error(RT_ERR_INVALID, "Error loading " + path + ": truncated header");
    if (h[0] != 'B' || h[1] != 'M') throw Error(RT_ERR_INVALID, "Error loading " + path + ": Not a valid BMP file.");
    auto u32 = [&](int o) {
        return (uint32_t)h[o] | ((uint32_t)h[o + 1] << 8) | ((uint32_t)h[o + 2] << 16) | ((uint32_t)h[o + 3] << 24);
    };
    const uint32_t offset = u32(10);
    const int32_t width = (int32_t)u32(18), height = (int32_t)u32(22);
    const uint32_t bpp = (uint32_t)h[28] | ((uint32_t)h[29] << 8);
    if (bpp != 24 && bpp != 32)
        throw Error(RT_ERR_INVALID, "Error loading " + path + ": Only 24 or 32 bit BMP files are supported, file is " +
                                        std::to_string(bpp) + " bit");
    if (width <= 0 || height <= 0)
        throw Error(RT_ERR_UNSUPPORTED, "Error loading " + path + ": only bottom-up BMP files with a positive size");
    f.seekg(offset, std::ios::beg);
    Texture t;
    t.width = (uint32_t)width;
    t.height = (uint32_t)height;
    t.rgb.resize((size_t)width * height * 3);
    const std::streamsize bytes = bpp / 8;
    for (size_t i = 0; i < (size_t)width * height; ++i) {
        unsigned char px[4];
        if (!f.read(reinterpret_cast<char *>(px), bytes))
            throw Error(RT_ERR_INVALID, "Error loading " + path + ": truncated pixel data");
        t.rgb[3 * i] = px[2];
        t.rgb[3 * i + 1] = px[1];
        t.rgb[3 * i + 2] = px[0];
    }
    return t;
}

// Textures by file name, so a file used by several materials loads once.
using TextureCache = std::map<std::string, std::shared_ptr<const Texture>>;

// The .obejota camera with a film of w x h pixels (storeObejotaObject,
// FileImporting.cu:387-422).
inline ProjectiveCamera make_camera(const ObejotaCamera &c, uint32_t w, uint32_t h)
{
    return ProjectiveCamera(c.pos, c.lookat, c.up, c.cts, c.frame_y, w, h);
}

// The material class an MTL block stands for, as the GPU path renders it:
// illum 0 Plain(Kd, Ke, map_Kd), 1 Lambertian(Kd, map_Kd), 3 Dielectric(1,
// Ni), 5 Trowbridge(Kd, Pr, eta, k), 6 TrowbridgeBasic(Kd, Pr).  A map_Kd
// texture is loaded with loadBMP (through `textures` when given).  Throws
// RT_ERR_UNSUPPORTED for the others.
inline std::shared_ptr<Material> make_material(const MtlMaterial &m, TextureCache *textures = nullptr)
{
    static const char *kNames[8] = {"PlainMaterial", "LambertianMaterial", "PhongMaterial", "DielectricMaterial",
                                    "BeckmannMaterial", "TrowbridgeMaterial", "TrowbridgeBasicMaterial",
                                    "SSMaterial"};
    auto refuse = [&](const std::string &why) -> std::shared_ptr<Material> {
        throw Error(RT_ERR_UNSUPPORTED, "material " + m.name + " (illum " + std::to_string(m.illum) + ", " +
                                            kNames[m.illum & 7] + "): " + why);
    };
    if (!m.map_kd.empty() && m.illum != 0 && m.illum != 1 && m.illum != 3)
        return refuse("textures are not rendered on the GPU path");
    std::shared_ptr<const Texture> tex;
    if (!m.map_kd.empty() && m.illum != 3) {
        if (textures) {
            auto it = textures->find(m.map_kd);
            if (it == textures->end())
                it = textures->emplace(m.map_kd, std::make_shared<const Texture>(loadBMP(m.map_kd))).first;
            tex = it->second;
        } else {
            tex = std::make_shared<const Texture>(loadBMP(m.map_kd));
        }
    }
    switch (m.illum) {
    case 0:
        if (m.transparent) return refuse("a Plain material without shadows is not rendered on the GPU path");
        return std::make_shared<PlainMaterial>(m.kd, m.ke, tex);
    case 1:
        if (m.transparent) return refuse("a Lambertian material without shadows is not rendered on the GPU path");
        return std::make_shared<LambertianMaterial>(m.kd, tex);
    case 3:
        return std::make_shared<DielectricMaterial>(Spectrum(1.0f), m.ni);
    case 5:
    case 6:
        if (m.rough_x != m.rough_y) return refuse("anisotropic roughness is not rendered on the GPU path");
        if (m.illum == 5) return std::make_shared<TrowbridgeMaterial>(m.kd, m.rough_x, m.eta, m.k);
        return std::make_shared<TrowbridgeBasicMaterial>(m.kd, m.rough_x);
    default:
        return refuse("this material class is not rendered on the GPU path");
    }
}

// The OBJ's triangles as TriangleMeshes for Scene::build: one mesh per run
// of consecutive faces with one material, so triangle ids keep face order.
// `materials` caches the converted materials by name (one object per
// name, so Scene::build registers each once).  A file with `vt` lines gives
// every face its corners' texture coordinates (a face without them gets the
// default corners (0,0), (1,0), (1,1), where the reference would read
// uninitialised indices); a file without them leaves the default corners to
// every face.  A file with `vn` lines gives every face its corners' normals
// (TriangleMeshData::normals: the shading normal is interpolated and the hit
// point moved 0.15 along it, Triangle.cu:379-418); a face that names no
// normals gets none (NaN), where the reference would read unset indices.
inline std::vector<std::shared_ptr<Shape>> obj_shapes(const ObjData &o,
                                                      std::map<std::string, std::shared_ptr<Material>> &materials,
                                                      TextureCache *textures = nullptr)
{
    std::vector<std::shared_ptr<Shape>> out;
    size_t i = 0;
    while (i < o.triangles.size()) {
        const std::string &mname = o.triangles[i].material;
        auto it = materials.find(mname);
        if (it == materials.end())
            it = materials.emplace(mname, make_material(o.materials.at(mname), textures)).first;
        std::vector<float> xyz, uv, nrm;
        std::vector<uint32_t> idx;
        std::map<int32_t, uint32_t> remap;
        for (; i < o.triangles.size() && o.triangles[i].material == mname; ++i) {
            if (!o.texcoords.empty()) {
                const ObjTriangle &t = o.triangles[i];
                const bool has = t.vt[0] >= 0 && t.vt[1] >= 0 && t.vt[2] >= 0;
                static const float kDefault[6] = {0, 0, 1, 0, 1, 1};
                for (int k = 0; k < 3; ++k) {
                    uv.push_back(has ? o.texcoords[(size_t)t.vt[k]].first : kDefault[2 * k]);
                    uv.push_back(has ? o.texcoords[(size_t)t.vt[k]].second : kDefault[2 * k + 1]);
                }
            }
            if (!o.normals.empty()) {
                const ObjTriangle &t = o.triangles[i];
                const bool has = t.vn[0] >= 0 && t.vn[1] >= 0 && t.vn[2] >= 0;
                for (int k = 0; k < 3; ++k) {
                    if (has) {
                        const Point3f &n = o.normals[(size_t)t.vn[k]];
                        nrm.insert(nrm.end(), {n.x, n.y, n.z});
                    } else {
                        const float q = std::numeric_limits<float>::quiet_NaN();
                        nrm.insert(nrm.end(), {q, q, q});
                    }
                }
            }
            for (int k = 0; k < 3; ++k) {
                const int32_t v = o.triangles[i].v[k];
                auto r = remap.find(v);
                if (r == remap.end()) {
                    r = remap.emplace(v, (uint32_t)(xyz.size() / 3)).first;
                    const Point3f &p = o.vertices[(size_t)v];
                    xyz.insert(xyz.end(), {p.x, p.y, p.z});
                }
                idx.push_back(r->second);
            }
        }
        out.push_back(
            std::make_shared<TriangleMesh>(std::move(xyz), std::move(idx), it->second, std::move(uv), std::move(nrm)));
    }
    return out;
}

}  // namespace rt_amd
