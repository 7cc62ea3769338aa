// rt_scene_types.h -- POD records shared by the host scene code and the
// render kernels.
#pragma once

#include <cstdint>

#include "../../include/rt.h"

namespace rt {

// Camera matrices composed exactly as the reference composes them
// (camera/ProjectiveCamera.cu:26-57, transformations/Transform.cu:611-640).
struct CameraDev {
    float r2c[16];  // rasterToCamera
    float c2w[16];  // cameraToWorld (= inverse(inverse(lookAt matrix)))
    uint32_t width, height;
};

// Sphere(center, radius): objectToWorld = translate(c) * rotateZ(90)
// (shapes/Sphere.cu:233-245), worldToObject its cofactor inverse.
struct SphereDev {
    float o2w[16], w2o[16];
    float center[3];
    float radius;
    uint32_t material;
    uint32_t is_light;
    uint32_t transparent;  // dielectric: skipped by shadow rays (Sphere.cu:362-363)
    uint32_t pad;
};

// Plane(corner, axisX, axisY, normal) (shapes/Plane.cu:27-48):
// objectToWorld = baseChange(corner, axisX, axisY, normal) with the axis
// lengths truncated to integers, worldToObject its cofactor inverse; the
// shape is the unit square of the object's z = 0 plane.
struct PlaneDev {
    float o2w[16], w2o[16];
    uint32_t material;
    uint32_t transparent;  // dielectric: skipped by shadow rays (Plane.cu:128-129)
    uint32_t pad[2];
};

// Texture (textures/Texture.cu): width x height texels from offset in the
// scene's texel array, 0x00BBGGRR, row y = pixelMatrix[y].
struct TexDev {
    uint32_t offset, width, height, pad;
};

constexpr int kMaxSpheres = 16;
constexpr int kMaxPlanes = 64;
constexpr int kMaxPointLights = 32;

}  // namespace rt
