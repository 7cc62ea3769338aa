// photon_build.cpp -- the photon map's kd-tree, built on the host as
// nn::KDTree<Photon, 3, PhotonPosition> builds it (cuda_libraries/kdtree.h:
// build_tree, used through Scene::buildPhotonStructures,
// integration/Scene.cu:1102-1114).
//
// The tree is implicit in the element order: the node of an index range
// [l, r) is element (l + r) / 2, its children are [l, median) and
// [median + 1, r), and a range of two or more elements splits on the axis of
// its bounding box's largest extent (the first such axis on a tie) at the
// median of that coordinate.  Only that recursion fixes the order; which of
// several elements with equal coordinates lands on which side of a median is
// left to std::nth_element, as in the reference.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <vector>

#include "rt_photon.h"

namespace rt {

namespace {

struct Builder {
    const float *xyz;
    std::vector<uint32_t> &order;  // tree position -> input index
    std::vector<uint8_t> &axis;
    uint32_t depth = 0;

    float c(uint32_t e, int a) const { return xyz[3 * (size_t)e + a]; }

    void build(size_t left, size_t right, uint32_t level)
    {
        depth = std::max(depth, level + 1);
        if (right - left <= 1) return;
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) lo[a] = hi[a] = c(order[left], a);
        for (size_t i = left + 1; i < right; ++i)
            for (int a = 0; a < 3; ++a) {
                const float v = c(order[i], a);
                if (lo[a] > v) lo[a] = v;
                if (hi[a] < v) hi[a] = v;
            }
        const size_t median = (left + right) / 2;
        int ax = 0;
        float ext = hi[0] - lo[0];
        for (int a = 1; a < 3; ++a)
            if (hi[a] - lo[a] > ext) { ax = a; ext = hi[a] - lo[a]; }
        std::nth_element(order.begin() + (std::ptrdiff_t)left, order.begin() + (std::ptrdiff_t)median,
                         order.begin() + (std::ptrdiff_t)right,
                         [&](uint32_t a, uint32_t b) { return c(a, ax) < c(b, ax); });
        axis[median] = (uint8_t)ax;
        build(left, median, level + 1);
        build(median + 1, right, level + 1);
    }
};

// Shape signatures of the subtree of every range [l, r), bottom-up.
uint32_t fill_signatures(const std::vector<PhotonNode> &nodes, std::vector<uint32_t> &sig, size_t l, size_t r)
{
    if (r <= l) return 0u;
    const size_t med = (l + r) / 2;
    const uint32_t s = photon_shape_bit(nodes[med].meta >> 2) | fill_signatures(nodes, sig, l, med) |
                       fill_signatures(nodes, sig, med + 1, r);
    sig[med] = s;
    return s;
}

}  // namespace

void build_photon_tree(const float *xyz, const int32_t *shape, uint32_t n, HostPhotonTree &out)
{
    out.order.resize(n);
    std::iota(out.order.begin(), out.order.end(), 0u);
    std::vector<uint8_t> axis(n, 0);
    Builder b{xyz, out.order, axis};
    if (n) b.build(0, n, 0);
    out.depth = b.depth;
    out.nodes.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t e = out.order[i];
        PhotonNode &p = out.nodes[i];
        p.x = xyz[3 * (size_t)e];
        p.y = xyz[3 * (size_t)e + 1];
        p.z = xyz[3 * (size_t)e + 2];
        p.meta = photon_meta(shape ? shape[e] : -1, axis[i]);
    }
    out.sig.assign(n, 0u);
    fill_signatures(out.nodes, out.sig, 0, n);
}

}  // namespace rt
