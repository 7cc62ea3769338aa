// rt_internal.h -- shared declarations between the host runtime (C++) and the
// HIP kernels.  Plain data only; no HIP types, so host-only translation units
// can include it.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/rt.h"

namespace rt {

// Two-wide BVH node: both child boxes live in the parent, so one visit reads
// one 64-byte record (a single cache line) and tests two boxes.
//   f[0..3]   c0.lo.x c0.hi.x c0.lo.y c0.hi.y
//   f[4..7]   c1.lo.x c1.hi.x c1.lo.y c1.hi.y
//   f[8..11]  c0.lo.z c0.hi.z c1.lo.z c1.hi.z
//   i[12..13] child0 child1  (>= 0: node index; < 0: leaf, ~c = first<<3 | count-1)
//   i[14..15] unused
struct alignas(64) Node2 {
    float f[12];
    int32_t child[2];
    int32_t pad[2];
};
static_assert(sizeof(Node2) == 64, "node record must be one cache line");

// Triangle record in leaf order (48 B): three float4 with the raw vertices so
// the watertight test sees exactly the caller's FP32 coordinates.
//   v0.xyz, id (int bits) | v1.xyz, flags (bit0 = transparent) | v2.xyz, 0
struct alignas(16) Tri3 {
    float v0[3]; int32_t id;
    float v1[3]; uint32_t flags;
    float v2[3]; uint32_t pad;
};
static_assert(sizeof(Tri3) == 48, "triangle record is 48 bytes");

constexpr uint32_t kLeafCountBits = 3;
constexpr uint32_t kMaxLeafCount = 1u << kLeafCountBits;

inline int32_t encode_leaf(uint32_t first, uint32_t count)
{
    return (int32_t)~((first << kLeafCountBits) | (count - 1));
}

// Four-wide node whose child boxes are quantized to 8 bits per bound against
// the node's origin and a power-of-two scale per axis, rounded outward, so
// one 64-byte record (one cache line) tests four boxes conservatively.
//   origin[3]  float; meta = ex | ey << 8 | ez << 16 | valid << 24, where
//              e? is the float exponent field of the scale (scale = 2^(e - 127))
//              and bit i of valid marks child slot i as used
//   child[4]   node index, or leaf ~(first << 3 | count - 1)
//   q[0..3]    lo_x, hi_x, lo_y, hi_y: byte i belongs to child i
//   q[4..5]    lo_z, hi_z; q[6..7] zero
// An unused slot holds an empty box (lo byte 255, hi byte 0) and a leaf
// reference to the tree's degenerate sentinel triangle.
// Child box i along x: [origin.x + lo_x[i] * sx, origin.x + hi_x[i] * sx].
struct alignas(64) Node4 {
    float origin[3];
    uint32_t meta;
    int32_t child[4];
    uint32_t q[8];
};
static_assert(sizeof(Node4) == 64, "wide node must be one cache line");

struct HostBvh {
    std::vector<Node2> nodes;
    std::vector<Node4> nodes4;   // the same tree collapsed to four-wide nodes
    std::vector<Tri3> tris;     // leaf order, plus one degenerate sentinel at the end
    uint32_t num_tris = 0;
    uint32_t num_leaves = 0;
    uint32_t max_depth = 0;
    uint32_t max_stack4 = 0;     // pending-entry bound of the four-wide traversal
    uint32_t depth4 = 0;
    double sah_cost = 0.0;
    double build_ms = 0.0;
};

// Binned-SAH build on the host (bvh_build.cpp).  Returns false and fills err
// on invalid input.
bool build_bvh2(const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt,
                const rt_bvh_opts &opts, HostBvh &out, std::string &err);

// Collapses out.nodes (two-wide) into out.nodes4: each wide node takes its
// two children and repeatedly opens the internal child of largest surface
// area until it has four; boxes are quantized outward.  Sets max_stack4 and
// depth4.
void build_bvh4(HostBvh &out);

// Thread-local error text (api.cpp).
void set_error(const std::string &msg);

}  // namespace rt
