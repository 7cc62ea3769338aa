// rt_photon.h -- the photon map's kd-tree shared by the host builder
// (photon_build.cpp) and the search kernel (photon.hip).  Plain data only.
#pragma once

#include <cstdint>
#include <vector>

#include "rt_pcg.h"  // RT_HD

namespace rt {

// One tree position (16 B on the device): the photon's location and
// meta = (shape + 1) << 2 | split axis, so shapes -1 .. 2^30 - 2 fit; the
// axis of a one-element range is 0 and never read.
struct alignas(16) PhotonNode {
    float x, y, z;
    uint32_t meta;
};
static_assert(sizeof(PhotonNode) == 16, "photon node is 16 bytes");

inline uint32_t photon_meta(int32_t shape, uint32_t axis) { return (uint32_t)(shape + 1) << 2 | axis; }

// Shape signature bit of a shape id + 1 (the node's meta >> 2); 0 (no
// shape) has none.  A subtree's signature is the OR of its photons' bits:
// a search restricted to shape s can pass over a subtree whose signature
// lacks s's bit, since no photon in it can be taken.
RT_HD uint32_t photon_shape_bit(uint32_t shape1) { return shape1 ? 1u << (shape1 & 31u) : 0u; }

struct HostPhotonTree {
    std::vector<PhotonNode> nodes;   // tree order
    std::vector<uint32_t> order;     // tree position -> input index
    std::vector<uint32_t> sig;       // tree position -> shape signature of its subtree
    uint32_t depth = 0;              // levels of the implicit tree
};

// The reference's build (kdtree.h build_tree) over n photon locations
// (xyz: 3n floats) with their shape ids (null: all -1).
void build_photon_tree(const float *xyz, const int32_t *shape, uint32_t n, HostPhotonTree &out);

}  // namespace rt
