// rt_photon.h -- the photon map's kd-tree shared by the host builder
// (photon_build.cpp) and the search kernel (photon.hip).  Plain data only.
#pragma once

#include <cstdint>
#include <vector>

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

struct HostPhotonTree {
    std::vector<PhotonNode> nodes;   // tree order
    std::vector<uint32_t> order;     // tree position -> input index
    uint32_t depth = 0;              // levels of the implicit tree
};

// The reference's build (kdtree.h build_tree) over n photon locations
// (xyz: 3n floats) with their shape ids (null: all -1).
void build_photon_tree(const float *xyz, const int32_t *shape, uint32_t n, HostPhotonTree &out);

}  // namespace rt
