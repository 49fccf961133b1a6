// bvh.cpp -- host binned-SAH BVH over facet triangles, flattened to the
// child-pair node layout the traversal kernels read (engine.hpp).
//
// Boxes are padded by 1e-5 * max(1, max |coordinate|): the closest hit is
// defined over all triangles (smallest t, ties to the smaller facet), so box
// culling must never drop a triangle whose computed t is within the bound.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <numeric>
#include <thread>

#include "engine.hpp"

namespace fh {
namespace {

struct Box {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const float* p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    float area() const {
        if (lo[0] > hi[0]) return 0.0f;
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Ref {             // result of building a subtree
    bool leaf = true;
    int32_t child = -1;  // leaf: first slot; internal: node index
    int32_t count = 0;   // leaf primitive count
    Box box;
};

// Large ranges near the root build their left subtree on another thread
// into a node list of their own, appended afterwards with its internal
// child indices shifted: the partition of idx is the sequential one (each
// thread only reorders its own range), so the tree and the leaf order are
// the same as a one-thread build, only numbered differently.
constexpr uint32_t kParallelPrims = 1u << 15;
constexpr int kParallelDepth = 5;  // up to 32 concurrent subtrees

struct Shared {
    std::vector<Box> pbox;
    std::vector<float> cen;  // 3 per prim
    std::vector<uint32_t> idx;
    float pad = 0.0f;
};

struct Builder {
    Shared* sh;
    std::vector<Box>& pbox;
    std::vector<float>& cen;
    std::vector<uint32_t>& idx;
    std::vector<BvhNode> nodes;
    float pad = 0.0f;
    int maxDepth = 0;

    explicit Builder(Shared* s) : sh(s), pbox(s->pbox), cen(s->cen), idx(s->idx), pad(s->pad) {}

    Box range_box(uint32_t b, uint32_t e) const {
        Box x;
        for (uint32_t i = b; i < e; ++i) x.grow(pbox[idx[i]]);
        return x;
    }

    Ref build(uint32_t b, uint32_t e, int depth) {
        maxDepth = std::max(maxDepth, depth);
        const uint32_t n = e - b;
        Box box = range_box(b, e);
        if (n <= 4 || depth >= kBvhMaxDepth - 2) return Ref{true, int32_t(b), int32_t(n), box};
        Box cb;
        for (uint32_t i = b; i < e; ++i) cb.grow(&cen[3 * size_t(idx[i])]);
        constexpr int kBins = 16;
        float bestCost = FLT_MAX;
        int bestAxis = -1, bestSplit = -1;
        for (int axis = 0; axis < 3; ++axis) {
            const float ext = cb.hi[axis] - cb.lo[axis];
            if (!(ext > 0.0f)) continue;
            Box bb[kBins];
            uint32_t bc[kBins] = {};
            const float scale = kBins / ext;
            for (uint32_t i = b; i < e; ++i) {
                const uint32_t p = idx[i];
                int k = int((cen[3 * size_t(p) + axis] - cb.lo[axis]) * scale);
                k = std::min(std::max(k, 0), kBins - 1);
                bb[k].grow(pbox[p]);
                ++bc[k];
            }
            float rightArea[kBins];
            uint32_t rightCount[kBins];
            Box acc;
            uint32_t cnt = 0;
            for (int k = kBins - 1; k > 0; --k) {
                acc.grow(bb[k]);
                cnt += bc[k];
                rightArea[k] = acc.area();
                rightCount[k] = cnt;
            }
            Box lacc;
            uint32_t lcnt = 0;
            for (int k = 0; k < kBins - 1; ++k) {
                lacc.grow(bb[k]);
                lcnt += bc[k];
                if (lcnt == 0 || rightCount[k + 1] == 0) continue;
                const float cost = lacc.area() * lcnt + rightArea[k + 1] * rightCount[k + 1];
                if (cost < bestCost) {
                    bestCost = cost;
                    bestAxis = axis;
                    bestSplit = k;
                }
            }
        }
        uint32_t mid;
        if (bestAxis < 0) {
            mid = b + n / 2;  // all centroids coincide: split by index
        } else {
            const float ext = cb.hi[bestAxis] - cb.lo[bestAxis];
            const float scale = kBins / ext;
            auto it = std::partition(idx.begin() + b, idx.begin() + e, [&](uint32_t p) {
                int k = int((cen[3 * size_t(p) + bestAxis] - cb.lo[bestAxis]) * scale);
                k = std::min(std::max(k, 0), kBins - 1);
                return k <= bestSplit;
            });
            mid = uint32_t(it - idx.begin());
            if (mid == b || mid == e) mid = b + n / 2;
        }
        const int32_t me = int32_t(nodes.size());
        nodes.emplace_back();
        Ref l, r;
        if (n >= kParallelPrims && depth < kParallelDepth) {
            Builder left(sh);
            std::thread worker([&] { l = left.build(b, mid, depth + 1); });
            r = build(mid, e, depth + 1);
            worker.join();
            const int32_t off = int32_t(nodes.size());
            for (BvhNode nd : left.nodes) {
                if (nd.meta.z == 0 && nd.meta.x >= 0) nd.meta.x += off;
                if (nd.meta.w == 0 && nd.meta.y >= 0) nd.meta.y += off;
                nodes.push_back(nd);
            }
            if (!l.leaf) l.child += off;
            maxDepth = std::max(maxDepth, left.maxDepth);
        } else {
            l = build(b, mid, depth + 1);
            r = build(mid, e, depth + 1);
        }
        fill(nodes[size_t(me)], l, r);
        return Ref{false, me, 0, box};
    }

    void fill(BvhNode& nd, const Ref& l, const Ref& r) const {
        const Ref* c[2] = {&l, &r};
        float lo[2][3], hi[2][3];
        int32_t ch[2], ct[2];
        for (int i = 0; i < 2; ++i) {
            for (int k = 0; k < 3; ++k) {
                lo[i][k] = c[i]->box.lo[k] - pad;
                hi[i][k] = c[i]->box.hi[k] + pad;
            }
            ch[i] = c[i]->child;
            ct[i] = c[i]->leaf ? c[i]->count : 0;
        }
        nd.x = make_float4(lo[0][0], hi[0][0], lo[1][0], hi[1][0]);
        nd.y = make_float4(lo[0][1], hi[0][1], lo[1][1], hi[1][1]);
        nd.z = make_float4(lo[0][2], hi[0][2], lo[1][2], hi[1][2]);
        nd.meta = make_int4(ch[0], ch[1], ct[0], ct[1]);
    }
};

}  // namespace

void build_bvh(const float* tri, uint32_t F, BvhHost& out) {
    Shared sh;
    sh.pbox.resize(F);
    sh.cen.resize(3 * size_t(F));
    sh.idx.resize(F);
    float m = 1.0f;
    for (uint32_t f = 0; f < F; ++f) {
        const float* t = tri + 9 * size_t(f);
        for (int c = 0; c < 3; ++c) sh.pbox[f].grow(t + 3 * c);
        for (int k = 0; k < 3; ++k) {
            sh.cen[3 * size_t(f) + k] = (t[k] + t[3 + k] + t[6 + k]) * (1.0f / 3.0f);
            for (int c = 0; c < 3; ++c) m = std::max(m, std::fabs(t[3 * c + k]));
        }
    }
    sh.pad = 1.0e-5f * m;
    std::iota(sh.idx.begin(), sh.idx.end(), 0u);
    Builder bd(&sh);
    bd.nodes.reserve(size_t(F) / 2 + 8);
    if (F <= 4) {
        Ref leaf{true, 0, int32_t(F), bd.range_box(0, F)};
        Ref empty{true, -1, 0, Box{}};
        bd.nodes.emplace_back();
        bd.fill(bd.nodes[0], leaf, empty);
        // empty slot: invert its box so it never hits
        bd.nodes[0].x.z = FLT_MAX;
        bd.nodes[0].x.w = -FLT_MAX;
        bd.nodes[0].meta.w = 0;
    } else {
        bd.build(0, F, 0);
    }
    out.nodes = std::move(bd.nodes);
    out.depth = bd.maxDepth;
    out.primFacet = bd.idx;
    out.tris.resize(3 * size_t(F));
    for (uint32_t s = 0; s < F; ++s) {
        const uint32_t f = bd.idx[s];
        const float* t = tri + 9 * size_t(f);
        float fid;
        std::memcpy(&fid, &f, sizeof(float));
        out.tris[3 * size_t(s) + 0] = make_float4(t[0], t[1], t[2], fid);
        out.tris[3 * size_t(s) + 1] = make_float4(t[3], t[4], t[5], 0.0f);
        out.tris[3 * size_t(s) + 2] = make_float4(t[6], t[7], t[8], 0.0f);
    }
}

}  // namespace fh
