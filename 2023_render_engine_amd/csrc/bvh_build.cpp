// bvh_build.cpp -- binned-SAH construction of the two-wide BVH on the host.
//
// Replaces the reference's Morton-pairing build (cuda_libraries/BVH.cu:
// 153-341 with shapes/Triangle.cu:737-776), which has one triangle per leaf
// and a topology fixed by the sort order.  Here: 32 centroid bins per axis,
// SAH with a leaf-size cap, subtrees built on a thread pool, then flattened
// depth-first so a parent and its near child are usually adjacent in HBM.
//
// Child boxes are padded outward so the FP32 slab test in the kernel is
// conservative with respect to the FP32 watertight triangle test: a
// triangle the test accepts always sits in boxes the ray enters.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <future>
#include <limits>
#include <thread>

#include "rt_internal.h"

namespace rt {
namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box &b)
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const float *p)
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    double area() const
    {
        double d[3];
        for (int k = 0; k < 3; ++k) d[k] = std::max(0.0, (double)hi[k] - (double)lo[k]);
        return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
    }
    bool empty() const { return lo[0] > hi[0]; }
};

struct Ref {
    Box box;
    float c[3];
    uint32_t id;
};

struct BNode {
    Box box;
    int32_t left = -1, right = -1;  // build-node indices; -1 for a leaf
    uint32_t first = 0, count = 0;  // leaf range in refs
};

struct Builder {
    std::vector<Ref> refs;
    std::vector<BNode> nodes;
    std::atomic<uint32_t> next{0};
    uint32_t bins = 32, max_leaf = 4;
    float trav_cost = 1.2f;

    uint32_t alloc() { return next.fetch_add(1, std::memory_order_relaxed); }

    // Builds the subtree over refs[b, e) into node slot `slot`.
    void build(uint32_t slot, uint32_t b, uint32_t e, int par_depth)
    {
        BNode &node = nodes[slot];
        Box box, cbox;
        for (uint32_t i = b; i < e; ++i) {
            box.grow(refs[i].box);
            cbox.grow(refs[i].c);
        }
        node.box = box;
        const uint32_t n = e - b;
        if (n <= 1) {
            node.first = b;
            node.count = n;
            return;
        }

        // Binned SAH over the centroid bounds.
        int best_axis = -1;
        uint32_t best_split = 0;
        double best_cost = INFINITY;
        std::vector<Box> bbox(bins);
        std::vector<uint32_t> bcnt(bins);
        std::vector<double> rarea(bins);
        std::vector<uint32_t> rcnt(bins);
        for (int axis = 0; axis < 3; ++axis) {
            float extent = cbox.hi[axis] - cbox.lo[axis];
            if (!(extent > 0.0f)) continue;
            float scale = (float)bins / extent;
            std::fill(bbox.begin(), bbox.end(), Box());
            std::fill(bcnt.begin(), bcnt.end(), 0u);
            for (uint32_t i = b; i < e; ++i) {
                int bi = (int)((refs[i].c[axis] - cbox.lo[axis]) * scale);
                bi = std::min(std::max(bi, 0), (int)bins - 1);
                bbox[bi].grow(refs[i].box);
                ++bcnt[bi];
            }
            Box acc;
            uint32_t cnt = 0;
            for (int k = (int)bins - 1; k > 0; --k) {
                acc.grow(bbox[k]);
                cnt += bcnt[k];
                rarea[k] = acc.empty() ? 0.0 : acc.area();
                rcnt[k] = cnt;
            }
            acc = Box();
            cnt = 0;
            for (uint32_t k = 0; k + 1 < bins; ++k) {
                acc.grow(bbox[k]);
                cnt += bcnt[k];
                if (cnt == 0 || rcnt[k + 1] == 0) continue;
                double cost = acc.area() * cnt + rarea[k + 1] * rcnt[k + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_split = k + 1;
                }
            }
        }
        const double parea = box.area();
        const double split_cost = best_axis >= 0 && parea > 0.0 ? trav_cost + best_cost / parea : INFINITY;
        if (n <= max_leaf && (double)n <= split_cost) {
            node.first = b;
            node.count = n;
            return;
        }

        uint32_t mid;
        if (best_axis >= 0) {
            float scale = (float)bins / (cbox.hi[best_axis] - cbox.lo[best_axis]);
            const float lo = cbox.lo[best_axis];
            auto it = std::partition(refs.begin() + b, refs.begin() + e, [&](const Ref &r) {
                int bi = (int)((r.c[best_axis] - lo) * scale);
                bi = std::min(std::max(bi, 0), (int)bins - 1);
                return (uint32_t)bi < best_split;
            });
            mid = (uint32_t)(it - refs.begin());
            if (mid == b || mid == e) mid = b + n / 2;
        } else {
            // All centroids coincide: split by count.
            mid = b + n / 2;
        }

        uint32_t l = alloc(), r = alloc();
        node.left = (int32_t)l;
        node.right = (int32_t)r;
        if (par_depth > 0 && n > 65536) {
            auto fut = std::async(std::launch::async, [&, l, b, mid, par_depth] { build(l, b, mid, par_depth - 1); });
            build(r, mid, e, par_depth - 1);
            fut.get();
        } else {
            build(l, b, mid, 0);
            build(r, mid, e, 0);
        }
    }
};

void pad_box(const Box &in, float pad, float *lo, float *hi)
{
    for (int k = 0; k < 3; ++k) {
        lo[k] = in.lo[k] - pad;
        hi[k] = in.hi[k] + pad;
    }
}

}  // namespace

bool build_bvh2(const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt,
                const rt_bvh_opts &opts, HostBvh &out, std::string &err)
{
    auto t0 = std::chrono::steady_clock::now();
    if (nt > 0 && (!xyz || !idx)) {
        err = "rt_bvh_build: null vertex or index array";
        return false;
    }
    if (nt >= (1u << 28)) {
        err = "rt_bvh_build: more than 2^28 triangles";
        return false;
    }
    Builder B;
    B.max_leaf = opts.max_leaf ? opts.max_leaf : 4;
    B.bins = opts.bins ? opts.bins : 32;
    B.trav_cost = opts.trav_cost > 0 ? opts.trav_cost : 1.2f;
    if (B.max_leaf > kMaxLeafCount || B.bins < 4 || B.bins > 64) {
        err = "rt_bvh_build: max_leaf must be 1..8 and bins 4..64";
        return false;
    }
    B.refs.resize(nt);
    Box scene;
    for (uint32_t i = 0; i < nt; ++i) {
        Ref &r = B.refs[i];
        r.id = i;
        for (int v = 0; v < 3; ++v) {
            uint32_t vi = idx[3 * i + v];
            if (vi >= nv) {
                err = "rt_bvh_build: vertex index out of range at triangle " + std::to_string(i);
                return false;
            }
            const float *p = xyz + 3 * (size_t)vi;
            if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) {
                err = "rt_bvh_build: non-finite vertex at triangle " + std::to_string(i);
                return false;
            }
            r.box.grow(p);
        }
        for (int k = 0; k < 3; ++k) r.c[k] = 0.5f * (r.box.lo[k] + r.box.hi[k]);
        scene.grow(r.box);
    }

    // Outward padding: 2^-19 of (scene diagonal + largest |coordinate|) per
    // side.  This covers the FP32 rounding of the watertight test (a few
    // ulps of |vertex - origin|) for origins within a few scene sizes, and
    // the rounding of the fused slab test.
    float pad = 0.0f;
    if (nt > 0) {
        double diag = 0.0, amax = 0.0;
        for (int k = 0; k < 3; ++k) {
            diag += ((double)scene.hi[k] - scene.lo[k]) * ((double)scene.hi[k] - scene.lo[k]);
            amax = std::max(amax, (double)std::max(std::fabs(scene.lo[k]), std::fabs(scene.hi[k])));
        }
        pad = (float)((std::sqrt(diag) + amax) * std::ldexp(1.0, -19));
    }

    B.nodes.resize(nt > 0 ? 2 * (size_t)nt : 1);
    uint32_t root = B.alloc();
    if (nt > 0) {
        unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        int par_depth = 0;
        while ((1u << par_depth) < hw && par_depth < 6) ++par_depth;
        B.build(root, 0, nt, par_depth);
    } else {
        B.nodes[root].count = 0;
    }

    // Flatten depth-first.  A build leaf at the root becomes child 0 of a
    // synthetic root whose child 1 is the degenerate sentinel leaf.
    out = HostBvh();
    out.num_tris = nt;
    out.tris.reserve(nt + 1);
    out.nodes.reserve(std::max<size_t>(1, B.next.load()));
    uint32_t depth_max = 0;
    double sah = 0.0;
    const double root_area = nt ? B.nodes[root].box.area() : 0.0;

    auto emit_leaf = [&](const BNode &bn) -> int32_t {
        uint32_t first = (uint32_t)out.tris.size();
        for (uint32_t k = 0; k < bn.count; ++k) {
            const Ref &r = B.refs[bn.first + k];
            Tri3 t{};
            for (int v = 0; v < 3; ++v) {
                const float *p = xyz + 3 * (size_t)idx[3 * r.id + v];
                float *dst = v == 0 ? t.v0 : (v == 1 ? t.v1 : t.v2);
                dst[0] = p[0]; dst[1] = p[1]; dst[2] = p[2];
            }
            t.id = (int32_t)r.id;
            out.tris.push_back(t);
        }
        ++out.num_leaves;
        if (root_area > 0) sah += bn.box.area() / root_area * bn.count;
        return encode_leaf(first, bn.count);
    };

    // Depth-first emission: a node's near child follows it in memory.
    std::function<int32_t(uint32_t, uint32_t)> emit = [&](uint32_t bi, uint32_t depth) -> int32_t {
        const BNode &bn = B.nodes[bi];
        if (bn.left < 0) return emit_leaf(bn);
        uint32_t my = (uint32_t)out.nodes.size();
        out.nodes.emplace_back();
        depth_max = std::max(depth_max, depth);
        if (root_area > 0) sah += 1.2 * bn.box.area() / root_area;
        const BNode &l = B.nodes[bn.left], &r = B.nodes[bn.right];
        float lo0[3], hi0[3], lo1[3], hi1[3];
        pad_box(l.box, pad, lo0, hi0);
        pad_box(r.box, pad, lo1, hi1);
        int32_t c0 = emit((uint32_t)bn.left, depth + 1);
        int32_t c1 = emit((uint32_t)bn.right, depth + 1);
        Node2 &n = out.nodes[my];
        n.f[0] = lo0[0]; n.f[1] = hi0[0]; n.f[2] = lo0[1]; n.f[3] = hi0[1];
        n.f[4] = lo1[0]; n.f[5] = hi1[0]; n.f[6] = lo1[1]; n.f[7] = hi1[1];
        n.f[8] = lo0[2]; n.f[9] = hi0[2]; n.f[10] = lo1[2]; n.f[11] = hi1[2];
        n.child[0] = c0;
        n.child[1] = c1;
        n.pad[0] = n.pad[1] = 0;
        return (int32_t)my;
    };

    const BNode &rn = B.nodes[root];
    if (nt > 0 && rn.left >= 0) {
        emit(root, 1);
    } else {
        // Synthetic root: child 0 = the single real leaf (if any), child 1 =
        // the sentinel.  Boxes of empty slots enclose only the sentinel.
        out.nodes.emplace_back();
        depth_max = 1;
        int32_t c0;
        float lo0[3], hi0[3];
        if (nt > 0) {
            c0 = emit_leaf(rn);
            pad_box(rn.box, pad, lo0, hi0);
        } else {
            c0 = -1;  // patched to the sentinel below
            for (int k = 0; k < 3; ++k) { lo0[k] = 0.0f; hi0[k] = 0.0f; }
        }
        Node2 &n = out.nodes[0];
        n.f[0] = lo0[0]; n.f[1] = hi0[0]; n.f[2] = lo0[1]; n.f[3] = hi0[1];
        n.f[8] = lo0[2]; n.f[9] = hi0[2];
        n.f[4] = n.f[5] = n.f[6] = n.f[7] = 0.0f;
        n.f[10] = n.f[11] = 0.0f;
        n.child[0] = c0;
        n.child[1] = 0;  // patched below
        n.pad[0] = n.pad[1] = 0;
    }
    // Degenerate sentinel triangle (all vertices at the origin): every
    // watertight test rejects it (det == 0), so pointing at it is harmless.
    Tri3 sentinel{};
    sentinel.id = -1;
    uint32_t sidx = (uint32_t)out.tris.size();
    out.tris.push_back(sentinel);
    if (!(nt > 0 && rn.left >= 0)) {
        out.nodes[0].child[1] = encode_leaf(sidx, 1);
        if (nt == 0) out.nodes[0].child[0] = encode_leaf(sidx, 1);
    }
    out.max_depth = depth_max;
    out.sah_cost = sah;
    out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return true;
}


namespace {

struct Kid {
    float lo[3], hi[3];
    int32_t ref;
};

double kid_area(const Kid &k)
{
    double d[3];
    for (int a = 0; a < 3; ++a) d[a] = std::max(0.0, (double)k.hi[a] - (double)k.lo[a]);
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
}

void node2_kids(const Node2 &n, Kid &k0, Kid &k1)
{
    k0.lo[0] = n.f[0]; k0.hi[0] = n.f[1]; k0.lo[1] = n.f[2]; k0.hi[1] = n.f[3]; k0.lo[2] = n.f[8]; k0.hi[2] = n.f[9];
    k1.lo[0] = n.f[4]; k1.hi[0] = n.f[5]; k1.lo[1] = n.f[6]; k1.hi[1] = n.f[7]; k1.lo[2] = n.f[10]; k1.hi[2] = n.f[11];
    k0.ref = n.child[0];
    k1.ref = n.child[1];
}

// Smallest power-of-two scale s = 2^e (e >= -126) with 255 * s >= extent.
int scale_exponent(double extent)
{
    if (!(extent > 0.0)) return -126;
    int e = (int)std::ceil(std::log2(extent / 255.0));
    if (e < -126) e = -126;
    while (255.0 * std::ldexp(1.0, e) < extent) ++e;
    while (e > -126 && 255.0 * std::ldexp(1.0, e - 1) >= extent) --e;
    return e;
}

struct Collapser {
    const HostBvh &in;
    std::vector<Node4> &out;
    uint32_t depth_max = 0;

    // Emits the wide node for two-wide node n2; returns its index and the
    // pending-entry bound of its subtree.
    int32_t emit(uint32_t n2, uint32_t depth, uint32_t &need)
    {
        Kid kids[4];
        Node2 root = in.nodes[n2];
        node2_kids(root, kids[0], kids[1]);
        int nk = 2;
        while (nk < 4) {
            int best = -1;
            double best_area = -1.0;
            for (int i = 0; i < nk; ++i)
                if (kids[i].ref >= 0 && kid_area(kids[i]) > best_area) {
                    best_area = kid_area(kids[i]);
                    best = i;
                }
            if (best < 0) break;
            Kid a, b;
            node2_kids(in.nodes[kids[best].ref], a, b);
            kids[best] = a;
            kids[nk++] = b;
        }
        const uint32_t my = (uint32_t)out.size();
        out.emplace_back();
        depth_max = std::max(depth_max, depth);
        int32_t refs[4] = {0, 0, 0, 0};
        uint32_t sub = 0;
        for (int i = 0; i < nk; ++i) {
            if (kids[i].ref >= 0) {
                uint32_t nd = 0;
                refs[i] = emit((uint32_t)kids[i].ref, depth + 1, nd);
                sub = std::max(sub, nd);
            } else {
                refs[i] = kids[i].ref;
            }
        }
        need = (uint32_t)(nk - 1) + sub;
        // unused slots: an empty box (lo byte 255 > hi byte 0 on every axis,
        // which no ray enters) around the degenerate sentinel leaf, so even a
        // box test that rounds an empty box open only tests a triangle that
        // every test rejects
        const uint32_t sentinel = (uint32_t)in.tris.size() - 1u;
        for (int i = nk; i < 4; ++i) refs[i] = encode_leaf(sentinel, 1);
        Node4 &n = out[my];
        std::memset(&n, 0, sizeof n);
        uint32_t meta = 0;
        for (int a = 0; a < 3; ++a) {
            float p = kids[0].lo[a];
            for (int i = 1; i < nk; ++i) p = std::min(p, kids[i].lo[a]);
            double ext = 0.0;
            for (int i = 0; i < nk; ++i) ext = std::max(ext, (double)kids[i].hi[a] - (double)p);
            const int e = scale_exponent(ext);
            const double sc = std::ldexp(1.0, e);
            n.origin[a] = p;
            meta |= (uint32_t)(e + 127) << (8 * a);
            uint32_t qlo = 0, qhi = 0;
            for (int i = 0; i < nk; ++i) {
                double l = std::floor(((double)kids[i].lo[a] - (double)p) / sc);
                double h = std::ceil(((double)kids[i].hi[a] - (double)p) / sc);
                l = std::min(std::max(l, 0.0), 255.0);
                h = std::min(std::max(h, 0.0), 255.0);
                // outward: the dequantized bounds enclose the box exactly
                while (l > 0.0 && (double)p + l * sc > (double)kids[i].lo[a]) l -= 1.0;
                while (h < 255.0 && (double)p + h * sc < (double)kids[i].hi[a]) h += 1.0;
                qlo |= (uint32_t)l << (8 * i);
                qhi |= (uint32_t)h << (8 * i);
            }
            for (int i = nk; i < 4; ++i) qlo |= 255u << (8 * i);
            n.q[2 * a] = qlo;
            n.q[2 * a + 1] = qhi;
        }
        meta |= (uint32_t)((1u << nk) - 1u) << 24;
        n.meta = meta;
        for (int i = 0; i < 4; ++i) n.child[i] = refs[i];
        return (int32_t)my;
    }
};

}  // namespace

void build_bvh4(HostBvh &h)
{
    h.nodes4.clear();
    h.nodes4.reserve(h.nodes.size() / 2 + 1);
    Collapser c{h, h.nodes4};
    uint32_t need = 0;
    c.emit(0, 1, need);
    h.max_stack4 = need;
    h.depth4 = c.depth_max;
}
}  // namespace rt
