// bvh.cpp -- binned-SAH binary build, collapsed into 128-byte 4-wide nodes.
//
// The binary tree uses 32 centroid bins per axis (Wald 2007); below depth
// kSahDepth it switches to object-median splits so the depth, and with it
// the traversal stack, stays bounded.  The collapse pulls grandchildren up
// by largest surface area until a node has four children (Wald et al.,
// "Getting rid of packets", 2008), so a node fetch is one 128-byte line that
// tests four boxes.  Child boxes are widened by two ulps outward so the
// kernel's conservative slab test never drops a triangle the watertight
// test would hit.
#include "bvh.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace omv {

namespace {

struct BinNode {
    float lo[3], hi[3];
    int left = -1, right = -1;
    int start = 0, count = 0;  // count > 0: leaf
};

constexpr int kBins = 32;
constexpr int kSahDepth = 40;

inline void growBox(float lo[3], float hi[3], const float l[3], const float h[3]) {
    for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], l[a]); hi[a] = std::max(hi[a], h[a]); }
}
inline float halfArea(const float lo[3], const float hi[3]) {
    float d0 = std::max(0.0f, hi[0] - lo[0]), d1 = std::max(0.0f, hi[1] - lo[1]), d2 = std::max(0.0f, hi[2] - lo[2]);
    return d0 * d1 + d1 * d2 + d2 * d0;
}
inline float padDown(float v) { return std::nextafter(std::nextafter(v, -INFINITY), -INFINITY); }
inline float padUp(float v) { return std::nextafter(std::nextafter(v, INFINITY), INFINITY); }

struct Builder {
    const std::vector<BuildBox>& boxes;
    int leafMax;
    std::vector<float> cen;
    std::vector<int> order;
    std::vector<BinNode> nodes;

    Builder(const std::vector<BuildBox>& b, int lm) : boxes(b), leafMax(lm) {
        const int n = (int)b.size();
        cen.resize((size_t)n * 3);
        order.resize(n);
        for (int i = 0; i < n; ++i) {
            order[i] = i;
            for (int a = 0; a < 3; ++a) cen[(size_t)i * 3 + a] = 0.5f * b[i].lo[a] + 0.5f * b[i].hi[a];
        }
    }

    void build() {
        struct Job { int node, start, count, depth; };
        std::vector<Job> jobs;
        nodes.emplace_back();
        jobs.push_back({0, 0, (int)order.size(), 0});
        while (!jobs.empty()) {
            Job j = jobs.back();
            jobs.pop_back();
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int i = j.start; i < j.start + j.count; ++i) {
                const int p = order[i];
                growBox(lo, hi, boxes[p].lo, boxes[p].hi);
                const float* c = &cen[(size_t)p * 3];
                growBox(clo, chi, c, c);
            }
            BinNode& nd = nodes[j.node];
            std::memcpy(nd.lo, lo, sizeof lo);
            std::memcpy(nd.hi, hi, sizeof hi);
            nd.start = j.start;
            nd.count = j.count;
            if (j.count <= leafMax) continue;

            int mid = -1;
            if (j.depth < kSahDepth) {
                int bestAxis = -1, bestBin = -1;
                float bestCost = INFINITY;
                for (int a = 0; a < 3; ++a) {
                    if (!(chi[a] > clo[a])) continue;
                    float blo[kBins][3], bhi[kBins][3];
                    int bcnt[kBins];
                    for (int b = 0; b < kBins; ++b) {
                        bcnt[b] = 0;
                        for (int k = 0; k < 3; ++k) { blo[b][k] = INFINITY; bhi[b][k] = -INFINITY; }
                    }
                    const float scale = kBins / (chi[a] - clo[a]);
                    for (int i = j.start; i < j.start + j.count; ++i) {
                        const int p = order[i];
                        int b = (int)((cen[(size_t)p * 3 + a] - clo[a]) * scale);
                        b = std::min(std::max(b, 0), kBins - 1);
                        growBox(blo[b], bhi[b], boxes[p].lo, boxes[p].hi);
                        ++bcnt[b];
                    }
                    float rArea[kBins];
                    int rCnt[kBins];
                    float alo[3] = {INFINITY, INFINITY, INFINITY}, ahi[3] = {-INFINITY, -INFINITY, -INFINITY};
                    int acc = 0;
                    for (int b = kBins - 1; b > 0; --b) {
                        growBox(alo, ahi, blo[b], bhi[b]);
                        acc += bcnt[b];
                        rArea[b] = halfArea(alo, ahi);
                        rCnt[b] = acc;
                    }
                    for (int k = 0; k < 3; ++k) { alo[k] = INFINITY; ahi[k] = -INFINITY; }
                    acc = 0;
                    for (int b = 0; b < kBins - 1; ++b) {
                        growBox(alo, ahi, blo[b], bhi[b]);
                        acc += bcnt[b];
                        if (acc == 0 || rCnt[b + 1] == 0) continue;
                        const float cost = halfArea(alo, ahi) * (float)acc + rArea[b + 1] * (float)rCnt[b + 1];
                        if (cost < bestCost) { bestCost = cost; bestAxis = a; bestBin = b; }
                    }
                }
                if (bestAxis >= 0) {
                    const float scale = kBins / (chi[bestAxis] - clo[bestAxis]);
                    const float c0 = clo[bestAxis];
                    auto it = std::partition(order.begin() + j.start, order.begin() + j.start + j.count, [&](int p) {
                        int b = (int)((cen[(size_t)p * 3 + bestAxis] - c0) * scale);
                        b = std::min(std::max(b, 0), kBins - 1);
                        return b <= bestBin;
                    });
                    mid = (int)(it - order.begin());
                    if (mid == j.start || mid == j.start + j.count) mid = -1;
                }
            }
            if (mid < 0) {
                // object median along the widest centroid axis
                int axis = 0;
                float ext = chi[0] - clo[0];
                for (int a = 1; a < 3; ++a) if (chi[a] - clo[a] > ext) { ext = chi[a] - clo[a]; axis = a; }
                mid = j.start + j.count / 2;
                std::nth_element(order.begin() + j.start, order.begin() + mid, order.begin() + j.start + j.count, [&](int x, int y) {
                    float cx = cen[(size_t)x * 3 + axis], cy = cen[(size_t)y * 3 + axis];
                    return cx < cy || (cx == cy && x < y);
                });
            }
            const int l = (int)nodes.size();
            nodes.emplace_back();
            nodes.emplace_back();
            nodes[j.node].left = l;
            nodes[j.node].right = l + 1;
            nodes[j.node].count = 0;
            jobs.push_back({l + 1, mid, j.start + j.count - mid, j.depth + 1});
            jobs.push_back({l, j.start, mid - j.start, j.depth + 1});
        }
    }
};

struct Collapser {
    const std::vector<BinNode>& bin;
    int leafBase;
    std::vector<OmvNode4>& pool;
    bool tlas;
    int maxDepth = 0;

    int emit(int b, int depth) {
        maxDepth = std::max(maxDepth, depth);
        int kids[4];
        int nk = 0;
        if (bin[b].count > 0) {
            kids[nk++] = b;   // a leaf root: one-child node
        } else {
            kids[nk++] = bin[b].left;
            kids[nk++] = bin[b].right;
            while (nk < 4) {
                int best = -1;
                float bestA = -1.0f;
                for (int k = 0; k < nk; ++k) {
                    const BinNode& c = bin[kids[k]];
                    if (c.count > 0) continue;
                    float a = halfArea(c.lo, c.hi);
                    if (a > bestA) { bestA = a; best = k; }
                }
                if (best < 0) break;
                const int inner = kids[best];
                kids[best] = bin[inner].left;
                kids[nk++] = bin[inner].right;
            }
        }
        const int idx = (int)pool.size();
        pool.emplace_back();
        int32_t child[4];
        uint32_t entry[4];
        for (int k = 0; k < 4; ++k) {
            if (k >= nk) { child[k] = OMV_EMPTY_CHILD; entry[k] = OMV_ENT_NONE; continue; }
            const BinNode& c = bin[kids[k]];
            if (c.count > 0) {
                child[k] = -(leafBase + c.start + 1);
                entry[k] = tlas ? OMV_ENT(OMV_ENT_INST, leafBase + c.start)
                                : OMV_ENT(OMV_ENT_BLAS_LEAF, ((uint32_t)(leafBase + c.start) << 3) | ((uint32_t)(c.count - 1) & 7u));
            } else {
                child[k] = emit(kids[k], depth + 1);
                entry[k] = OMV_ENT(tlas ? OMV_ENT_TLAS_NODE : OMV_ENT_BLAS_NODE, child[k]);
            }
        }
        OmvNode4& nd = pool[idx];
        for (int k = 0; k < 4; ++k) {
            nd.child[k] = child[k];
            nd.entry[k] = entry[k];
            if (k >= nk) {   // empty slot: a box at +inf on every plane, which no ray hits
                nd.lox[k] = nd.loy[k] = nd.loz[k] = INFINITY;
                nd.hix[k] = nd.hiy[k] = nd.hiz[k] = INFINITY;
                continue;
            }
            const BinNode& c = bin[kids[k]];
            nd.lox[k] = padDown(c.lo[0]); nd.loy[k] = padDown(c.lo[1]); nd.loz[k] = padDown(c.lo[2]);
            nd.hix[k] = padUp(c.hi[0]); nd.hiy[k] = padUp(c.hi[1]); nd.hiz[k] = padUp(c.hi[2]);
        }
        return idx;
    }
};

}  // namespace

Bvh4Result buildBvh4(const std::vector<BuildBox>& boxes, int leafMax, int leafBase, std::vector<OmvNode4>& pool, bool tlas) {
    Bvh4Result r;
    if (boxes.empty()) return r;
    Builder b(boxes, leafMax);
    b.build();
    Collapser c{b.nodes, leafBase, pool, tlas};
    r.root = c.emit(0, 1);
    r.depth = c.maxDepth;
    r.order = std::move(b.order);
    return r;
}

}  // namespace omv
