// bvh.cpp - host BVH builder (replaces the reference's Embree build) and the
// packer that turns the reference's skip-link layout into the GPU layout.
//
// Reference: BvhBuilder::buildEmbree / repackEmbree (src/render/bvh.cpp:107-235),
// AABB padding (include/render/bvh.h:138-151).  The reference asks Embree for a
// binary SAH tree with one primitive per leaf and unit traversal and
// intersection costs; Embree is not available, so this is our own binned SAH
// builder emitting the same node layout: DFS preorder, the hit link of node i
// is i+1, nodeOffset is the miss (skip) link, a left child's skip link is its
// right sibling and a right child's skip link is its parent's.
#include "bvh.hpp"

#include <algorithm>
#include <memory>
#include <array>
#include <atomic>
#include <map>
#include <cmath>
#include <thread>

namespace stk {

StkBvhOptions defaultBvhOptions()
{
    StkBvhOptions o;
    o.maxDepth = 32;
    o.numBins = 32;
    o.traversalCost = 1.0f;
    o.intersectionCost = 1.0f;
    o.spatialSplitAlpha = 0.0f;
    return o;
}

bool flattenInstances(const StkVertex* vb, uint32_t nv, const uint32_t* ib, uint32_t ni, const StkInstance* inst,
                      uint32_t ninst, std::vector<BvhInputTri>& out, std::string& err)
{
    out.clear();
    size_t total = 0;
    for (uint32_t i = 0; i < ninst; ++i) {
        const StkInstance& ic = inst[i];
        if (ic.indexOffset < 0 || ic.indexCount < 0 || ic.indexCount % 3 != 0 ||
            (uint64_t)ic.indexOffset + (uint64_t)ic.indexCount > ni) {
            err = "instance " + std::to_string(i) + " has an index range outside the index buffer";
            return false;
        }
        total += (size_t)ic.indexCount / 3;
    }
    out.reserve(total);
    for (uint32_t i = 0; i < ninst; ++i) {
        const StkInstance& ic = inst[i];
        const uint32_t ntri = (uint32_t)ic.indexCount / 3;
        for (uint32_t p = 0; p < ntri; ++p) {
            BvhInputTri t;
            for (int k = 0; k < 3; ++k) {
                const uint32_t vi = ib[(uint32_t)ic.indexOffset + p * 3 + k];
                if (vi >= nv) {
                    err = "index buffer entry outside the vertex buffer";
                    return false;
                }
                t.v[k] = xform4(ic.objectToWorld, ld3(vb[vi].position), 1.0f);
            }
            t.instId = i;
            t.primId = p;
            out.push_back(t);
        }
    }
    return true;
}

namespace {

struct PrimRef {
    float lo[3], hi[3], c[3];
    uint32_t id;
};

struct Box {
    float lo[3] = {1e30f, 1e30f, 1e30f};
    float hi[3] = {-1e30f, -1e30f, -1e30f};
    void grow(const float* l, const float* h)
    {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], l[a]);
            hi[a] = std::max(hi[a], h[a]);
        }
    }
    // AABB::halfArea (include/render/bvh.h:64-67)
    float halfArea() const
    {
        if (hi[0] < lo[0]) return 0.0f;
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return dx * (dy + dz) + dy * dz;
    }
};

inline uint32_t ceilLog2(uint32_t n)
{
    uint32_t l = 0;
    while ((1u << l) < n) ++l;
    return l;
}

class Builder {
public:
    Builder(const std::vector<BvhInputTri>& tris, const StkBvhOptions& o, std::vector<StkBvhNode>& nodes)
        : tris_(tris), opts_(o), nodes_(nodes)
    {
    }

    void run()
    {
        const uint32_t n = (uint32_t)tris_.size();
        refs_.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            const BvhInputTri& t = tris_[i];
            PrimRef& r = refs_[i];
            // boundingBox (include/render/bvh.h:138-151): max padded by 1e-6
            const float* v0 = &t.v[0].x;
            const float* v1 = &t.v[1].x;
            const float* v2 = &t.v[2].x;
            for (int a = 0; a < 3; ++a) {
                r.lo[a] = std::min(v0[a], std::min(v1[a], v2[a]));
                r.hi[a] = std::max(v0[a], std::max(v1[a], v2[a])) + 1e-6f;
                r.c[a] = (r.lo[a] + r.hi[a]) * 0.5f;
            }
            r.id = i;
        }
        nodes_.assign(2 * (size_t)n - 1, StkBvhNode{});
        maxDepthSeen_ = 0;
        emit(0, n, 0, STK_INVALID_INDEX, 0);
    }

    uint32_t depth() const { return maxDepthSeen_.load(); }

private:
    void noteDepth(uint32_t d)
    {
        uint32_t cur = maxDepthSeen_.load();
        while (d > cur && !maxDepthSeen_.compare_exchange_weak(cur, d)) {
        }
    }

    void emit(uint32_t begin, uint32_t end, uint32_t idx, uint32_t skip, uint32_t depth)
    {
        const uint32_t n = end - begin;
        StkBvhNode& node = nodes_[idx];
        node.nodeOffset = (int32_t)skip;
        if (n == 1) {
            const BvhInputTri& t = tris_[refs_[begin].id];
            node.minBounds[0] = bits_to_float(t.primId);
            node.minBounds[1] = 0.0f;
            node.minBounds[2] = 0.0f;
            node.maxBounds[0] = node.maxBounds[1] = node.maxBounds[2] = 0.0f;
            node.instId = (int32_t)t.instId;
            noteDepth(depth);
            return;
        }
        Box b;
        for (uint32_t i = begin; i < end; ++i) b.grow(refs_[i].lo, refs_[i].hi);
        for (int a = 0; a < 3; ++a) {
            node.minBounds[a] = b.lo[a];
            node.maxBounds[a] = b.hi[a];
        }
        node.instId = (int32_t)STK_INVALID_INDEX;
        const uint32_t mid = split(begin, end, depth, b);
        const uint32_t nl = mid - begin;
        const uint32_t left = idx + 1, right = idx + 2 * nl; // left subtree holds 2*nl-1 nodes
        if (n > 65536) {
            std::thread th([&] { emit(begin, mid, left, right, depth + 1); });
            emit(mid, end, right, skip, depth + 1);
            th.join();
        } else {
            emit(begin, mid, left, right, depth + 1);
            emit(mid, end, right, skip, depth + 1);
        }
    }

    uint32_t medianSplit(uint32_t begin, uint32_t end, int axis)
    {
        const uint32_t mid = begin + (end - begin) / 2;
        std::nth_element(refs_.begin() + begin, refs_.begin() + mid, refs_.begin() + end,
                         [axis](const PrimRef& a, const PrimRef& b) {
                             return a.c[axis] < b.c[axis] || (a.c[axis] == b.c[axis] && a.id < b.id);
                         });
        return mid;
    }

    uint32_t split(uint32_t begin, uint32_t end, uint32_t depth, const Box& nodeBox)
    {
        const uint32_t n = end - begin;
        Box cb;
        for (uint32_t i = begin; i < end; ++i) cb.grow(refs_[i].c, refs_[i].c);
        float ext[3];
        int axis = 0;
        for (int a = 0; a < 3; ++a) {
            ext[a] = cb.hi[a] - cb.lo[a];
            if (ext[a] > ext[axis]) axis = a;
        }
        if (!(ext[axis] > 0.0f)) return begin + n / 2; // all centroids coincide
        if (n <= 2) return medianSplit(begin, end, axis);

        const uint32_t maxD = opts_.maxDepth;
        const uint32_t nbins = std::min<uint32_t>(opts_.numBins ? opts_.numBins : 32, 64);
        struct Bin {
            Box b;
            uint32_t n = 0;
        };
        float bestCost = INFINITY;
        int bestAxis = -1;
        uint32_t bestBin = 0;
        const float parentArea = nodeBox.halfArea();
        for (int a = 0; a < 3; ++a) {
            if (!(ext[a] > 0.0f)) continue;
            Bin bins[64];
            const float scale = (float)nbins / ext[a] * (1.0f - 1e-6f);
            for (uint32_t i = begin; i < end; ++i) {
                const PrimRef& r = refs_[i];
                uint32_t bi = (uint32_t)((r.c[a] - cb.lo[a]) * scale);
                if (bi >= nbins) bi = nbins - 1;
                bins[bi].b.grow(r.lo, r.hi);
                bins[bi].n++;
            }
            float rightArea[64];
            uint32_t rightN[64];
            Box acc;
            uint32_t cnt = 0;
            for (uint32_t i = nbins - 1; i > 0; --i) {
                acc.grow(bins[i].b.lo, bins[i].b.hi);
                cnt += bins[i].n;
                rightArea[i] = acc.halfArea();
                rightN[i] = cnt;
            }
            Box lacc;
            uint32_t lcnt = 0;
            for (uint32_t i = 0; i + 1 < nbins; ++i) {
                lacc.grow(bins[i].b.lo, bins[i].b.hi);
                lcnt += bins[i].n;
                const uint32_t rcnt = rightN[i + 1];
                if (lcnt == 0 || rcnt == 0) continue;
                const uint32_t bigger = std::max(lcnt, rcnt);
                if (depth + 1 + ceilLog2(bigger) > maxD) continue; // keep the depth bound
                const float cost = opts_.traversalCost + opts_.intersectionCost *
                                                             (lacc.halfArea() * (float)lcnt +
                                                              rightArea[i + 1] * (float)rcnt) /
                                                             (parentArea > 0.0f ? parentArea : 1.0f);
                if (cost < bestCost) {
                    bestCost = cost;
                    bestAxis = a;
                    bestBin = i;
                }
            }
        }
        if (bestAxis < 0) return medianSplit(begin, end, axis);
        const int a = bestAxis;
        const float scale = (float)nbins / ext[a] * (1.0f - 1e-6f);
        auto it = std::partition(refs_.begin() + begin, refs_.begin() + end, [&](const PrimRef& r) {
            uint32_t bi = (uint32_t)((r.c[a] - cb.lo[a]) * scale);
            if (bi >= nbins) bi = nbins - 1;
            return bi <= bestBin;
        });
        const uint32_t mid = (uint32_t)(it - refs_.begin());
        if (mid == begin || mid == end) return medianSplit(begin, end, axis);
        return mid;
    }

    const std::vector<BvhInputTri>& tris_;
    StkBvhOptions opts_;
    std::vector<StkBvhNode>& nodes_;
    std::vector<PrimRef> refs_;
    std::atomic<uint32_t> maxDepthSeen_{0};
};

// ---- spatial splits ----------------------------------------------------------
//
// Binned SAH with spatial splits (Stich, Friedrich and Dietrich, "Spatial
// Splits in Bounding Volume Hierarchies", HPG 2009): a node's references are
// split by the best object partition of their centroids, or - when the two
// children of that partition overlap by more than alpha of the root's area -
// by the best of nbins planes across the node box, the triangles straddling
// the plane clipped into both children (each side keeps the bounds of its part
// of the triangle within the reference's box).  Leaves still hold one
// reference of one triangle, so the tree has 2R - 1 nodes for R references
// (a triangle may sit in several leaves); references grow by at most the
// split budget.  The tree is built with per-node reference lists and then
// laid out in the same preorder with skip links as the in-place builder.

struct SNode {
    float lo[3], hi[3];
    std::unique_ptr<SNode> left, right;
    uint32_t tri = STK_INVALID_INDEX; // leaf: the input triangle
    uint32_t leaves = 1;
};

// The bounds of triangle v (double) clipped to the box [lo, hi], rounded
// outwards to float; false when the clipped part is empty.
bool clippedBounds(const double v[3][3], const double lo[3], const double hi[3], float olo[3], float ohi[3])
{
    double poly[2][12][3];
    int n = 3, cur = 0;
    for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) poly[0][k][a] = v[k][a];
    for (int a = 0; a < 3 && n > 0; ++a)
        for (int side = 0; side < 2 && n > 0; ++side) {
            const double plane = side ? hi[a] : lo[a];
            int m = 0;
            for (int k = 0; k < n; ++k) {
                const double* p = poly[cur][k];
                const double* q = poly[cur][(k + 1) % n];
                const bool pin = side ? p[a] <= plane : p[a] >= plane;
                const bool qin = side ? q[a] <= plane : q[a] >= plane;
                if (pin) {
                    for (int c = 0; c < 3; ++c) poly[1 - cur][m][c] = p[c];
                    ++m;
                }
                if (pin != qin) {
                    const double t = (plane - p[a]) / (q[a] - p[a]);
                    for (int c = 0; c < 3; ++c) poly[1 - cur][m][c] = p[c] + (q[c] - p[c]) * t;
                    poly[1 - cur][m][a] = plane;
                    ++m;
                }
            }
            n = m;
            cur = 1 - cur;
        }
    if (n == 0) return false;
    double blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) {
            blo[a] = std::min(blo[a], poly[cur][k][a]);
            bhi[a] = std::max(bhi[a], poly[cur][k][a]);
        }
    for (int a = 0; a < 3; ++a) {
        // a margin of 1e-6 relative outwards, then rounded outwards to float:
        // the parts' boxes overlap across the plane
        const double m = 1e-6 * (1.0 + std::max(std::fabs(blo[a]), std::fabs(bhi[a])));
        olo[a] = std::nextafter((float)(blo[a] - m), -INFINITY);
        ohi[a] = std::nextafter((float)(bhi[a] + m), INFINITY);
    }
    return true;
}

class SpatialBuilder {
public:
    SpatialBuilder(const std::vector<BvhInputTri>& tris, const StkBvhOptions& o, std::vector<StkBvhNode>& nodes)
        : tris_(tris), opts_(o), nodes_(nodes)
    {
    }

    void run()
    {
        const uint32_t n = (uint32_t)tris_.size();
        std::vector<PrimRef> refs(n);
        Box all;
        for (uint32_t i = 0; i < n; ++i) {
            const BvhInputTri& t = tris_[i];
            PrimRef& r = refs[i];
            const float* v0 = &t.v[0].x;
            const float* v1 = &t.v[1].x;
            const float* v2 = &t.v[2].x;
            for (int a = 0; a < 3; ++a) {
                r.lo[a] = std::min(v0[a], std::min(v1[a], v2[a]));
                r.hi[a] = std::max(v0[a], std::max(v1[a], v2[a])) + 1e-6f;
                r.c[a] = (r.lo[a] + r.hi[a]) * 0.5f;
            }
            r.id = i;
            all.grow(r.lo, r.hi);
        }
        rootArea_ = all.halfArea();
        std::unique_ptr<SNode> root = build(std::move(refs), 0, (uint64_t)((double)n * 0.3));
        nodes_.assign(2 * (size_t)root->leaves - 1, StkBvhNode{});
        emit(*root, 0, STK_INVALID_INDEX, 0);
    }

    uint32_t depth() const { return maxDepth_; }
    uint32_t references() const { return (uint32_t)((nodes_.size() + 1) / 2); }

private:
    struct Choice {
        float cost = INFINITY;
        int axis = -1;
        uint32_t bin = 0;
        bool spatial = false;
    };

    // budget: the references this subtree may add by splitting; a split hands
    // what is left to its children in proportion to their references, so the
    // tree does not depend on the threads' timing
    std::unique_ptr<SNode> build(std::vector<PrimRef> refs, uint32_t depth, uint64_t budget)
    {
        std::unique_ptr<SNode> node(new SNode());
        const uint32_t n = (uint32_t)refs.size();
        Box b;
        for (const PrimRef& r : refs) b.grow(r.lo, r.hi);
        for (int a = 0; a < 3; ++a) {
            node->lo[a] = b.lo[a];
            node->hi[a] = b.hi[a];
        }
        if (n == 1) {
            node->tri = refs[0].id;
            return node;
        }
        std::vector<PrimRef> L, R;
        const uint64_t used = partition(refs, b, depth, budget, L, R);
        refs.clear();
        refs.shrink_to_fit();
        const uint64_t rest = budget - used;
        const uint64_t bl = (uint64_t)((double)rest * (double)L.size() / (double)(L.size() + R.size()));
        const uint64_t br = rest - bl;
        if (n > 65536) {
            std::unique_ptr<SNode> lc;
            std::thread th([&] { lc = build(std::move(L), depth + 1, bl); });
            node->right = build(std::move(R), depth + 1, br);
            th.join();
            node->left = std::move(lc);
        } else {
            node->left = build(std::move(L), depth + 1, bl);
            node->right = build(std::move(R), depth + 1, br);
        }
        node->leaves = node->left->leaves + node->right->leaves;
        return node;
    }

    // Object split by binned SAH over the centroids (as the in-place builder),
    // then the spatial split where the children would overlap.  Returns the
    // references the split added.
    uint64_t partition(std::vector<PrimRef>& refs, const Box& nodeBox, uint32_t depth, uint64_t budget,
                       std::vector<PrimRef>& L, std::vector<PrimRef>& R)
    {
        const uint32_t n = (uint32_t)refs.size();
        const uint32_t nbins = std::min<uint32_t>(opts_.numBins ? opts_.numBins : 32, 64);
        const uint32_t maxD = opts_.maxDepth;
        const float parentArea = nodeBox.halfArea() > 0.0f ? nodeBox.halfArea() : 1.0f;
        Box cb;
        for (const PrimRef& r : refs) cb.grow(r.c, r.c);
        Choice obj;
        Box objL, objR;
        for (int a = 0; a < 3; ++a) {
            const float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0.0f)) continue;
            Box bins[64];
            uint32_t cnt[64] = {};
            const float scale = (float)nbins / ext * (1.0f - 1e-6f);
            for (const PrimRef& r : refs) {
                uint32_t bi = (uint32_t)((r.c[a] - cb.lo[a]) * scale);
                if (bi >= nbins) bi = nbins - 1;
                bins[bi].grow(r.lo, r.hi);
                cnt[bi]++;
            }
            Box racc[64];
            uint32_t rn[64];
            Box acc;
            uint32_t c = 0;
            for (uint32_t i = nbins - 1; i > 0; --i) {
                acc.grow(bins[i].lo, bins[i].hi);
                c += cnt[i];
                racc[i] = acc;
                rn[i] = c;
            }
            Box lacc;
            uint32_t lc = 0;
            for (uint32_t i = 0; i + 1 < nbins; ++i) {
                lacc.grow(bins[i].lo, bins[i].hi);
                lc += cnt[i];
                if (lc == 0 || rn[i + 1] == 0) continue;
                if (depth + 1 + ceilLog2(std::max(lc, rn[i + 1])) > maxD) continue;
                const float cost = opts_.traversalCost +
                                   opts_.intersectionCost *
                                       (lacc.halfArea() * (float)lc + racc[i + 1].halfArea() * (float)rn[i + 1]) /
                                       parentArea;
                if (cost < obj.cost) {
                    obj.cost = cost;
                    obj.axis = a;
                    obj.bin = i;
                    objL = lacc;
                    objR = racc[i + 1];
                }
            }
        }
        // spatial split where the object split's children overlap
        Choice sp;
        if (obj.axis >= 0 && budget > 0) {
            Box ov;
            for (int a = 0; a < 3; ++a) {
                ov.lo[a] = std::max(objL.lo[a], objR.lo[a]);
                ov.hi[a] = std::min(objL.hi[a], objR.hi[a]);
            }
            const bool overlap = ov.hi[0] > ov.lo[0] && ov.hi[1] > ov.lo[1] && ov.hi[2] > ov.lo[2];
            if (overlap && ov.halfArea() > opts_.spatialSplitAlpha * rootArea_) sp = spatialChoice(refs, nodeBox, depth);
        }
        uint64_t dup = 0;
        if (sp.axis >= 0 && sp.cost < obj.cost && spatialSplit(refs, nodeBox, sp, budget, L, R, dup)) return dup;
        L.clear();
        R.clear();
        if (obj.axis >= 0) {
            const int a = obj.axis;
            const float ext = cb.hi[a] - cb.lo[a];
            const float scale = (float)nbins / ext * (1.0f - 1e-6f);
            for (const PrimRef& r : refs) {
                uint32_t bi = (uint32_t)((r.c[a] - cb.lo[a]) * scale);
                if (bi >= nbins) bi = nbins - 1;
                (bi <= obj.bin ? L : R).push_back(r);
            }
            if (!L.empty() && !R.empty()) return 0;
            L.clear();
            R.clear();
        }
        // median by centroid on the widest axis (index order on ties)
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (cb.hi[a] - cb.lo[a] > cb.hi[axis] - cb.lo[axis]) axis = a;
        std::sort(refs.begin(), refs.end(), [axis](const PrimRef& x, const PrimRef& y) {
            return x.c[axis] < y.c[axis] || (x.c[axis] == y.c[axis] && x.id < y.id);
        });
        L.assign(refs.begin(), refs.begin() + n / 2);
        R.assign(refs.begin() + n / 2, refs.end());
        return 0;
    }

    double planeOf(const Box& nodeBox, int a, uint32_t nbins, uint32_t bin) const
    {
        const double w = ((double)nodeBox.hi[a] - nodeBox.lo[a]) / nbins;
        return (double)nodeBox.lo[a] + w * (bin + 1);
    }

    void triangle(uint32_t id, double v[3][3]) const
    {
        const BvhInputTri& t = tris_[id];
        for (int k = 0; k < 3; ++k) {
            const float* p = &t.v[k].x;
            for (int a = 0; a < 3; ++a) v[k][a] = p[a];
        }
    }

    // Binned spatial split: each reference chopped into the bins it spans.
    Choice spatialChoice(const std::vector<PrimRef>& refs, const Box& nodeBox, uint32_t depth) const
    {
        Choice best;
        const uint32_t n = (uint32_t)refs.size();
        const uint32_t nbins = std::min<uint32_t>(opts_.numBins ? opts_.numBins : 32, 64);
        const float parentArea = nodeBox.halfArea() > 0.0f ? nodeBox.halfArea() : 1.0f;
        for (int a = 0; a < 3; ++a) {
            const double lo = nodeBox.lo[a], w = ((double)nodeBox.hi[a] - lo) / nbins;
            if (!(w > 0.0)) continue;
            Box bins[64];
            uint32_t enter[64] = {}, exit[64] = {};
            for (const PrimRef& r : refs) {
                int b0 = (int)std::floor((r.lo[a] - lo) / w), b1 = (int)std::floor((r.hi[a] - lo) / w);
                b0 = std::clamp(b0, 0, (int)nbins - 1);
                b1 = std::clamp(b1, b0, (int)nbins - 1);
                if (b0 == b1) {
                    bins[b0].grow(r.lo, r.hi);
                } else {
                    double v[3][3];
                    triangle(r.id, v);
                    for (int b = b0; b <= b1; ++b) {
                        double blo[3], bhi[3];
                        for (int c = 0; c < 3; ++c) {
                            blo[c] = r.lo[c];
                            bhi[c] = r.hi[c];
                        }
                        blo[a] = std::max(blo[a], lo + w * b);
                        bhi[a] = std::min(bhi[a], lo + w * (b + 1));
                        float clo[3], chi[3];
                        if (clippedBounds(v, blo, bhi, clo, chi)) bins[b].grow(clo, chi);
                    }
                }
                enter[b0]++;
                exit[b1]++;
            }
            Box racc[64];
            uint32_t rn[64];
            Box acc;
            uint32_t c = 0;
            for (uint32_t i = nbins - 1; i > 0; --i) {
                acc.grow(bins[i].lo, bins[i].hi);
                c += exit[i];
                racc[i] = acc;
                rn[i] = c;
            }
            Box lacc;
            uint32_t lc = 0;
            for (uint32_t i = 0; i + 1 < nbins; ++i) {
                lacc.grow(bins[i].lo, bins[i].hi);
                lc += enter[i];
                const uint32_t rc = rn[i + 1];
                if (lc == 0 || rc == 0 || lc >= n || rc >= n) continue;
                if (depth + 1 + ceilLog2(std::max(lc, rc)) > opts_.maxDepth) continue;
                const float cost =
                    opts_.traversalCost +
                    opts_.intersectionCost * (lacc.halfArea() * (float)lc + racc[i + 1].halfArea() * (float)rc) /
                        parentArea;
                if (cost < best.cost) {
                    best.cost = cost;
                    best.axis = a;
                    best.bin = i;
                    best.spatial = true;
                }
            }
        }
        return best;
    }

    // Splits the references at the chosen plane; false (nothing changed) when
    // a side would be empty or keep every reference.
    bool spatialSplit(const std::vector<PrimRef>& refs, const Box& nodeBox, const Choice& ch, uint64_t budget,
                      std::vector<PrimRef>& L, std::vector<PrimRef>& R, uint64_t& dupOut)
    {
        const uint32_t nbins = std::min<uint32_t>(opts_.numBins ? opts_.numBins : 32, 64);
        const int a = ch.axis;
        const double plane = planeOf(nodeBox, a, nbins, ch.bin);
        L.clear();
        R.clear();
        uint64_t dup = 0;
        for (const PrimRef& r : refs) {
            if ((double)r.hi[a] <= plane) {
                L.push_back(r);
                continue;
            }
            if ((double)r.lo[a] >= plane) {
                R.push_back(r);
                continue;
            }
            double v[3][3], blo[3], bhi[3];
            triangle(r.id, v);
            for (int c = 0; c < 3; ++c) {
                blo[c] = r.lo[c];
                bhi[c] = r.hi[c];
            }
            PrimRef pl = r, pr = r;
            bool hasL, hasR;
            {
                double h2[3] = {bhi[0], bhi[1], bhi[2]};
                h2[a] = plane;
                hasL = clippedBounds(v, blo, h2, pl.lo, pl.hi);
            }
            {
                double l2[3] = {blo[0], blo[1], blo[2]};
                l2[a] = plane;
                hasR = clippedBounds(v, l2, bhi, pr.lo, pr.hi);
            }
            // the parts stay inside the reference's box
            for (int c = 0; c < 3; ++c) {
                pl.lo[c] = std::max(pl.lo[c], r.lo[c]);
                pl.hi[c] = std::min(pl.hi[c], r.hi[c]);
                pr.lo[c] = std::max(pr.lo[c], r.lo[c]);
                pr.hi[c] = std::min(pr.hi[c], r.hi[c]);
            }
            if (!hasL && !hasR) {
                (r.c[a] < plane ? L : R).push_back(r);
                continue;
            }
            for (PrimRef* q : {&pl, &pr})
                for (int c = 0; c < 3; ++c) q->c[c] = (q->lo[c] + q->hi[c]) * 0.5f;
            if (hasL) L.push_back(pl);
            if (hasR) R.push_back(pr);
            if (hasL && hasR) ++dup;
        }
        const size_t n = refs.size();
        if (L.empty() || R.empty() || L.size() >= n || R.size() >= n || dup > budget) {
            L.clear();
            R.clear();
            return false;
        }
        dupOut = dup;
        return true;
    }

    void emit(const SNode& s, uint32_t idx, uint32_t skip, uint32_t depth)
    {
        StkBvhNode& node = nodes_[idx];
        node.nodeOffset = (int32_t)skip;
        maxDepth_ = std::max(maxDepth_, depth);
        if (!s.left) {
            const BvhInputTri& t = tris_[s.tri];
            node.minBounds[0] = bits_to_float(t.primId);
            node.minBounds[1] = node.minBounds[2] = 0.0f;
            node.maxBounds[0] = node.maxBounds[1] = node.maxBounds[2] = 0.0f;
            node.instId = (int32_t)t.instId;
            return;
        }
        for (int a = 0; a < 3; ++a) {
            node.minBounds[a] = s.lo[a];
            node.maxBounds[a] = s.hi[a];
        }
        node.instId = (int32_t)STK_INVALID_INDEX;
        const uint32_t left = idx + 1, right = idx + 2 * s.left->leaves;
        emit(*s.left, left, right, depth + 1);
        emit(*s.right, right, skip, depth + 1);
    }

    const std::vector<BvhInputTri>& tris_;
    StkBvhOptions opts_;
    std::vector<StkBvhNode>& nodes_;
    float rootArea_ = 0.0f;
    uint32_t maxDepth_ = 0;
};

} // namespace

bool buildBvh(const std::vector<BvhInputTri>& tris, const StkBvhOptions& optsIn, Bvh& out, std::string& err)
{
    out = Bvh{};
    StkBvhOptions opts = optsIn;
    if (opts.maxDepth == 0) opts.maxDepth = 32;
    if (opts.maxDepth > 32) {
        err = "maxDepth above 32 is not supported by the GPU traversal stack";
        return false;
    }
    if (tris.empty()) {
        out.header.root = INT32_MIN;
        return true;
    }
    if (ceilLog2((uint32_t)tris.size()) > opts.maxDepth) {
        err = "too many triangles for the requested maxDepth";
        return false;
    }
    if (opts.spatialSplitAlpha > 0.0f) {
        SpatialBuilder b(tris, opts, out.nodes);
        b.run();
        out.depth = b.depth();
        if (out.depth > opts.maxDepth) {
            err = "spatial-split build exceeded maxDepth";
            return false;
        }
        return true;
    }
    Builder b(tris, opts, out.nodes);
    b.run();
    out.depth = b.depth();
    return true;
}

bool packBvh(const StkVertex* vb, uint32_t nv, const uint32_t* ib, uint32_t ni, const StkInstance* inst,
             uint32_t ninst, Bvh& bvh, std::string& err)
{
    const uint32_t n = (uint32_t)bvh.nodes.size();
    bvh.packed.clear();
    bvh.tris.clear();
    bvh.leafToTri.assign(n, STK_INVALID_INDEX);
    bvh.header = StkBvhHeader{};
    if (n == 0) {
        bvh.header.root = INT32_MIN;
        bvh.depth = 0;
        return true;
    }
    // Validate the skip-link structure by walking it as a binary tree.
    std::vector<uint32_t> packedIndex(n, STK_INVALID_INDEX);
    std::vector<uint8_t> seen(n, 0);
    struct Item {
        uint32_t idx, skip, depth;
    };
    std::vector<Item> stack;
    stack.push_back({0, STK_INVALID_INDEX, 0});
    uint32_t nleaf = 0, ninner = 0, maxDepth = 0;
    std::vector<uint32_t> order; // preorder as walked
    order.reserve(n);
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        if (it.idx >= n || seen[it.idx]) {
            err = "BVH node index out of range or visited twice";
            return false;
        }
        seen[it.idx] = 1;
        order.push_back(it.idx);
        const StkBvhNode& nd = bvh.nodes[it.idx];
        if ((uint32_t)nd.nodeOffset != it.skip) {
            err = "BVH skip link of node " + std::to_string(it.idx) + " is inconsistent";
            return false;
        }
        maxDepth = std::max(maxDepth, it.depth);
        if ((uint32_t)nd.instId != STK_INVALID_INDEX) {
            const uint32_t prim = float_to_bits(nd.minBounds[0]);
            if ((uint32_t)nd.instId >= ninst || prim >= (uint32_t)inst[nd.instId].indexCount / 3) {
                err = "BVH leaf references a missing instance or primitive";
                return false;
            }
            ++nleaf;
            continue;
        }
        ++ninner;
        for (int a = 0; a < 3; ++a)
            if (!std::isfinite(nd.minBounds[a]) || !std::isfinite(nd.maxBounds[a]) || nd.minBounds[a] > nd.maxBounds[a]) {
                err = "BVH inner node " + std::to_string(it.idx) + " has a non-finite or inverted box";
                return false;
            }
        const uint32_t left = it.idx + 1;
        if (left >= n) {
            err = "BVH inner node without children";
            return false;
        }
        const uint32_t right = (uint32_t)bvh.nodes[left].nodeOffset;
        if (right >= n || right <= left) {
            err = "BVH right child link is invalid";
            return false;
        }
        stack.push_back({right, it.skip, it.depth + 1});
        stack.push_back({left, right, it.depth + 1});
    }
    if (nleaf + ninner != n || ninner + 1 != nleaf) {
        err = "BVH is not a full binary tree over its nodes";
        return false;
    }
    for (uint32_t i = 0; i < n; ++i)
        if (order[i] != i) {
            err = "BVH nodes are not stored in preorder";
            return false;
        }
    bvh.depth = maxDepth;
    // Number inner nodes and leaves in preorder.
    uint32_t pi = 0, ti = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if ((uint32_t)bvh.nodes[i].instId == STK_INVALID_INDEX)
            packedIndex[i] = pi++;
        else
            bvh.leafToTri[i] = ti++;
    }
    bvh.tris.resize(ti);
    for (uint32_t i = 0; i < n; ++i) {
        const StkBvhNode& nd = bvh.nodes[i];
        if ((uint32_t)nd.instId == STK_INVALID_INDEX) continue;
        const StkInstance& ic = inst[nd.instId];
        const uint32_t prim = float_to_bits(nd.minBounds[0]);
        uint32_t vi[3];
        for (int k = 0; k < 3; ++k) {
            const uint64_t at = (uint64_t)(uint32_t)ic.indexOffset + (uint64_t)prim * 3 + (uint64_t)k;
            if (at >= ni || ib[at] >= nv) {
                err = "BVH leaf triangle indexes outside the geometry buffers";
                return false;
            }
            vi[k] = ib[at];
        }
        // getTriangle (shaders/raytracing.h:134-153)
        const F3 v0 = xform4(ic.objectToWorld, ld3(vb[vi[0]].position), 1.0f);
        const F3 e0 = xform4(ic.objectToWorld, ld3(vb[vi[1]].position), 1.0f) - v0;
        const F3 e1 = xform4(ic.objectToWorld, ld3(vb[vi[2]].position), 1.0f) - v0;
        StkTriangle& t = bvh.tris[bvh.leafToTri[i]];
        t.v0[0] = v0.x;
        t.v0[1] = v0.y;
        t.v0[2] = v0.z;
        t.e0[0] = e0.x;
        t.e0[1] = e0.y;
        t.e0[2] = e0.z;
        t.e1[0] = e1.x;
        t.e1[1] = e1.y;
        t.e1[2] = e1.z;
        t.leaf = i;
        t.instId = (uint32_t)nd.instId;
        t.primId = prim;
    }
    bvh.packed.resize(pi);
    for (uint32_t i = 0; i < n; ++i) {
        const StkBvhNode& nd = bvh.nodes[i];
        if ((uint32_t)nd.instId != STK_INVALID_INDEX) continue;
        StkPackedNode& p = bvh.packed[packedIndex[i]];
        memset(&p, 0, sizeof(p));
        const uint32_t c[2] = {i + 1, (uint32_t)bvh.nodes[i + 1].nodeOffset};
        for (int k = 0; k < 2; ++k) {
            const StkBvhNode& cn = bvh.nodes[c[k]];
            if ((uint32_t)cn.instId != STK_INVALID_INDEX) {
                p.child[k] = ~(int32_t)bvh.leafToTri[c[k]];
                continue;
            }
            p.child[k] = (int32_t)packedIndex[c[k]];
            p.x[2 * k + 0] = cn.minBounds[0];
            p.x[2 * k + 1] = cn.maxBounds[0];
            p.y[2 * k + 0] = cn.minBounds[1];
            p.y[2 * k + 1] = cn.maxBounds[1];
            p.z[2 * k + 0] = cn.minBounds[2];
            p.z[2 * k + 1] = cn.maxBounds[2];
        }
    }
    StkBvhHeader& h = bvh.header;
    h.numTriangles = ti;
    if ((uint32_t)bvh.nodes[0].instId != STK_INVALID_INDEX) {
        h.root = ~(int32_t)0;
    } else {
        h.root = 0;
        for (int a = 0; a < 3; ++a) {
            h.lo[a] = bvh.nodes[0].minBounds[a];
            h.hi[a] = bvh.nodes[0].maxBounds[a];
        }
    }
    return true;
}

// ---- compressed 8-wide layout ---------------------------------------------------

namespace {

struct WideChild {
    uint32_t node; // skip-link node index
    bool leaf;
    float lo[3], hi[3];
};

float decodePlane(float origin, uint32_t q, float scale)
{
    const float prod = (float)q * scale; // exact: q < 2^8, scale a power of two
    return origin + prod;
}

float powerOfTwo(int e) { return bits_to_float((uint32_t)e << 23); } // 2^(e - 127), 1 <= e <= 254

// Smallest exponent byte e with 2^(e-127) * 253 >= extent, at least 2^-100
// and no finer than a quarter ulp of the origin's magnitude.
uint8_t chooseExponent(float origin, float extent)
{
    const double need = std::max((double)extent / 253.0, std::ldexp(1.0, -100));
    const double ulp = std::fabs((double)origin) * std::ldexp(1.0, -25);
    const double target = std::max(need, ulp);
    if (!(target < 1e300)) return 254; // non-finite boxes (rejected before the collapse)
    int e = (int)std::ceil(std::log2(target)) + 127;
    while (e > 1 && std::ldexp(1.0, e - 1 - 127) * 253.0 >= (double)extent && std::ldexp(1.0, e - 1 - 127) >= target)
        --e;
    while (std::ldexp(1.0, e - 127) < target) ++e;
    return (uint8_t)std::min(254, std::max(1, e));
}

// Which binary nodes become wide nodes: the collapse that minimises the summed
// surface area of the wide nodes (the walk's expected node tests; with one
// triangle per leaf slot the triangle tests do not depend on the collapse), by
// dynamic programming over the binary tree (Ylitie, Karras and Laine 2017,
// "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs").
// F(c, a): least cost of the subtree of c in at most a slots of its parent -
// 0 for a leaf, SA(c) + G(c) for c kept as one wide node, or c opened and its
// two children given a1 + a2 <= a slots; G(c): least cost inside a wide node
// at c (its children in at most 8 slots).
class WideCollapsePlan {
public:
    explicit WideCollapsePlan(const std::vector<StkBvhNode>& nodes) : nodes_(nodes)
    {
        const size_t n = nodes.size();
        F_.assign(n * 8, 0.0);
        openSplit_.assign(n * 8, 0);
        innerSplit_.assign(n, 0);
        // children have larger preorder indices than their parent
        for (size_t i = n; i-- > 0;) {
            if (isLeaf((uint32_t)i)) continue;
            const uint32_t L = (uint32_t)i + 1, R = (uint32_t)nodes[i + 1].nodeOffset;
            double g = kInf;
            for (int a1 = 1; a1 <= 7; ++a1) {
                const double c = F(L, a1) + F(R, 8 - a1);
                if (c < g) {
                    g = c;
                    innerSplit_[i] = (uint8_t)a1;
                }
            }
            const double keep = area((uint32_t)i) + g;
            for (int a = 1; a <= 8; ++a) {
                double best = keep;
                uint8_t split = 0; // 0: kept as one wide node
                for (int a1 = 1; a1 < a; ++a1) {
                    const double c = F(L, a1) + F(R, a - a1);
                    if (c < best) {
                        best = c;
                        split = (uint8_t)a1;
                    }
                }
                F_[i * 8 + (a - 1)] = best;
                openSplit_[i * 8 + (a - 1)] = split;
            }
        }
    }
    // The slots of a wide node at inner node i: binary nodes, in preorder.
    void slotsOf(uint32_t i, std::vector<uint32_t>& out) const
    {
        out.clear();
        const uint32_t L = i + 1, R = (uint32_t)nodes_[i + 1].nodeOffset;
        expand(L, innerSplit_[i], out);
        expand(R, 8 - innerSplit_[i], out);
    }

private:
    static constexpr double kInf = 1e300;
    bool isLeaf(uint32_t i) const { return (uint32_t)nodes_[i].instId != STK_INVALID_INDEX; }
    double F(uint32_t i, int a) const { return isLeaf(i) ? 0.0 : F_[(size_t)i * 8 + (a - 1)]; }
    double area(uint32_t i) const
    {
        const StkBvhNode& nd = nodes_[i];
        const double dx = (double)nd.maxBounds[0] - nd.minBounds[0], dy = (double)nd.maxBounds[1] - nd.minBounds[1],
                     dz = (double)nd.maxBounds[2] - nd.minBounds[2];
        const double s = dx * dy + dy * dz + dz * dx;
        return std::isfinite(s) ? s : kInf / 1e10; // non-finite boxes are rejected later; keep the sums finite
    }
    void expand(uint32_t c, int a, std::vector<uint32_t>& out) const
    {
        if (isLeaf(c) || a <= 1 || openSplit_[(size_t)c * 8 + (a - 1)] == 0) {
            out.push_back(c);
            return;
        }
        const int a1 = openSplit_[(size_t)c * 8 + (a - 1)];
        expand(c + 1, a1, out);
        expand((uint32_t)nodes_[c + 1].nodeOffset, a - a1, out);
    }
    const std::vector<StkBvhNode>& nodes_;
    std::vector<double> F_;
    std::vector<uint8_t> openSplit_, innerSplit_;
};

// Collapses a binary skip-link tree (preorder, hit link i + 1, miss link
// nodeOffset) into compressed 8-wide nodes.  leafBox(i, lo, hi) gives the box
// of leaf node i and emitLeaf(i) appends its leaf record; leaves are emitted
// in wide-node slot order, node by node in breadth-first order, and triBase
// counts them from 0.  Each wide node's children follow WideCollapsePlan.
template <class LeafBox, class EmitLeaf>
bool collapseWide(const std::vector<StkBvhNode>& nodes, LeafBox leafBox, EmitLeaf emitLeaf,
                  std::vector<StkWideNode>& wide, uint32_t& wideDepth, std::string& err)
{
    wide.clear();
    wideDepth = 0;
    if (nodes.empty()) return true;
    const WideCollapsePlan plan(nodes);
    std::vector<uint32_t> slotNodes;
    auto childOf = [&](uint32_t idx) {
        WideChild c;
        c.node = idx;
        const StkBvhNode& nd = nodes[idx];
        c.leaf = (uint32_t)nd.instId != STK_INVALID_INDEX;
        if (!c.leaf) {
            for (int a = 0; a < 3; ++a) {
                c.lo[a] = nd.minBounds[a];
                c.hi[a] = nd.maxBounds[a];
            }
            return c;
        }
        leafBox(idx, c.lo, c.hi);
        return c;
    };
    struct Job {
        uint32_t node; // skip-link inner node (or leaf for a single-leaf tree)
        uint32_t wide; // wide node to fill
        uint32_t level;
    };
    std::vector<Job> queue;
    uint32_t emitted = 0;
    wide.emplace_back();
    queue.push_back({0, 0, 1});
    for (size_t qi = 0; qi < queue.size(); ++qi) {
        const Job job = queue[qi];
        wideDepth = std::max(wideDepth, job.level);
        std::vector<WideChild> kids;
        if ((uint32_t)nodes[job.node].instId != STK_INVALID_INDEX) {
            kids.push_back(childOf(job.node)); // the whole tree is one leaf
        } else {
            plan.slotsOf(job.node, slotNodes);
            for (uint32_t c : slotNodes) kids.push_back(childOf(c));
        }
        // node box and slot assignment by the child's centre relative to the node's
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = kids[0].lo[a];
            hi[a] = kids[0].hi[a];
            for (const WideChild& c : kids) {
                lo[a] = std::min(lo[a], c.lo[a]);
                hi[a] = std::max(hi[a], c.hi[a]);
            }
        }
        int slotOf[8];
        bool slotUsed[8] = {false, false, false, false, false, false, false, false};
        std::vector<bool> placed(kids.size(), false);
        for (size_t round = 0; round < kids.size(); ++round) {
            double bestScore = -1e300;
            int bk = -1, bs = -1;
            for (size_t k = 0; k < kids.size(); ++k) {
                if (placed[k]) continue;
                double d[3];
                for (int a = 0; a < 3; ++a)
                    d[a] = 0.5 * ((double)kids[k].lo[a] + kids[k].hi[a]) - 0.5 * ((double)lo[a] + hi[a]);
                for (int sl = 0; sl < 8; ++sl) {
                    if (slotUsed[sl]) continue;
                    double score = 0.0;
                    for (int a = 0; a < 3; ++a) score += ((sl >> a) & 1 ? d[a] : -d[a]);
                    if (score > bestScore || bk < 0) { // the first free pair stands in for non-finite scores
                        bestScore = score;
                        bk = (int)k;
                        bs = sl;
                    }
                }
            }
            placed[bk] = true;
            slotUsed[bs] = true;
            slotOf[bk] = bs;
        }
        int kidAt[8];
        for (int sl = 0; sl < 8; ++sl) kidAt[sl] = -1;
        for (size_t k = 0; k < kids.size(); ++k) kidAt[slotOf[k]] = (int)k;
        StkWideNode w;
        memset(&w, 0, sizeof(w));
        float scale[3];
        for (int a = 0; a < 3; ++a) {
            w.origin[a] = lo[a];
            w.exps[a] = chooseExponent(lo[a], hi[a] - lo[a]);
            scale[a] = powerOfTwo(w.exps[a]);
        }
        w.childBase = (uint32_t)wide.size();
        w.triBase = emitted;
        uint32_t inner = 0;
        for (int sl = 0; sl < 8; ++sl) {
            if (kidAt[sl] < 0) continue;
            const WideChild& c = kids[kidAt[sl]];
            w.vmask |= 1u << sl;
            for (int a = 0; a < 3; ++a) {
                // outward rounding checked against the decode the traversal performs
                double ql = std::floor(((double)c.lo[a] - (double)w.origin[a]) / (double)scale[a]);
                double qh = std::ceil(((double)c.hi[a] - (double)w.origin[a]) / (double)scale[a]);
                uint32_t qlo = (uint32_t)std::min(255.0, std::max(0.0, ql));
                uint32_t qhi = (uint32_t)std::min(255.0, std::max(0.0, qh));
                while (qlo > 0 && decodePlane(w.origin[a], qlo, scale[a]) > c.lo[a]) --qlo;
                while (qhi < 255 && decodePlane(w.origin[a], qhi, scale[a]) < c.hi[a]) ++qhi;
                if (decodePlane(w.origin[a], qlo, scale[a]) > c.lo[a] ||
                    decodePlane(w.origin[a], qhi, scale[a]) < c.hi[a]) {
                    err = "wide BVH: a child box does not fit its node's quantisation grid";
                    return false;
                }
                w.qlo[a][sl] = (uint8_t)qlo;
                w.qhi[a][sl] = (uint8_t)qhi;
            }
            if (c.leaf) {
                emitLeaf(c.node);
                ++emitted;
            } else {
                w.imask |= (uint8_t)(1u << sl);
                ++inner;
            }
        }
        const uint32_t firstChild = (uint32_t)wide.size();
        wide.resize(wide.size() + inner);
        uint32_t r = 0;
        for (int sl = 0; sl < 8; ++sl)
            if (kidAt[sl] >= 0 && !kids[kidAt[sl]].leaf) queue.push_back({kids[kidAt[sl]].node, firstChild + r++, job.level + 1});
        wide[job.wide] = w;
    }
    return true;
}

// A leaf's box: its triangle's box (the vertices as given), widened by 1e-5
// of the larger of its extent and magnitude.
void paddedTriangleBox(const float* v0, const float* v1, const float* v2, float* lo, float* hi)
{
    for (int a = 0; a < 3; ++a) {
        const float l = std::min(v0[a], std::min(v1[a], v2[a])), h = std::max(v0[a], std::max(v1[a], v2[a]));
        const float mag = std::max(std::fabs(l), std::fabs(h));
        const float pad = 1e-5f * std::max(h - l, mag) + 1e-30f;
        lo[a] = l - pad;
        hi[a] = h + pad;
    }
}

} // namespace

bool buildWide(const StkInstance* inst, uint32_t ninst, Bvh& bvh, std::string& err)
{
    bvh.wide.clear();
    bvh.wideTris.clear();
    bvh.emitters.clear();
    bvh.wideDepth = 0;
    for (const StkTriangle& t : bvh.tris)
        if (t.instId < ninst && inst[t.instId].lightId != -1) bvh.emitters.push_back(t);
    // A tree with spatial splits holds a triangle in several leaves, each for
    // the part of it inside its parent's box: there a leaf's box is the
    // triangle's box cut to the parent's box widened by the same margin (it
    // still holds the leaf's part; hits outside it belong to the triangle's
    // other leaves).
    const uint32_t nn = (uint32_t)bvh.nodes.size();
    std::vector<uint32_t> parent;
    {
        std::vector<uint64_t> keys;
        keys.reserve(bvh.tris.size());
        for (const StkTriangle& t : bvh.tris) keys.push_back(((uint64_t)t.instId << 32) | t.primId);
        std::sort(keys.begin(), keys.end());
        if (std::adjacent_find(keys.begin(), keys.end()) != keys.end()) {
            parent.assign(nn, STK_INVALID_INDEX);
            for (uint32_t i = 0; i < nn; ++i)
                if ((uint32_t)bvh.nodes[i].instId == STK_INVALID_INDEX && i + 1 < nn) {
                    parent[i + 1] = i;
                    const uint32_t r = (uint32_t)bvh.nodes[i + 1].nodeOffset;
                    if (r < nn) parent[r] = i;
                }
        }
    }
    // a leaf's box: its triangle's box, the vertices as the hit test forms them
    auto leafBox = [&](uint32_t idx, float* lo, float* hi) {
        const StkTriangle& t = bvh.tris[bvh.leafToTri[idx]];
        float v1[3], v2[3];
        for (int a = 0; a < 3; ++a) {
            v1[a] = t.v0[a] + t.e0[a];
            v2[a] = t.v0[a] + t.e1[a];
        }
        paddedTriangleBox(t.v0, v1, v2, lo, hi);
        if (!parent.empty() && parent[idx] != STK_INVALID_INDEX) {
            const StkBvhNode& pn = bvh.nodes[parent[idx]];
            float plo[3], phi[3];
            float c0[3], c1[3], c2[3];
            for (int a = 0; a < 3; ++a) {
                c0[a] = pn.minBounds[a];
                c1[a] = pn.maxBounds[a];
                c2[a] = pn.maxBounds[a];
            }
            paddedTriangleBox(c0, c1, c2, plo, phi); // the parent's box with the leaf margin
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::max(lo[a], plo[a]);
                hi[a] = std::min(hi[a], phi[a]);
                if (lo[a] > hi[a]) lo[a] = hi[a] = 0.5f * (lo[a] + hi[a]);
            }
        }
    };
    auto emitLeaf = [&](uint32_t idx) { bvh.wideTris.push_back(bvh.tris[bvh.leafToTri[idx]]); };
    if (!collapseWide(bvh.nodes, leafBox, emitLeaf, bvh.wide, bvh.wideDepth, err)) return false;
    if (bvh.wideTris.size() != bvh.tris.size()) {
        err = "wide BVH: leaf count differs from the binary tree";
        return false;
    }
    return true;
}

// ---- two-level (TLAS / BLAS) ----------------------------------------------------

namespace {

// The world box of a box [lo, hi] under the affine map m (column-major 4x4),
// computed in double around the centre and rounded outwards to float, then
// widened by 1e-5 of the larger of its extent and magnitude per axis: it holds
// every point getTriangle's float arithmetic forms from points of [lo, hi].
void transformBox(const float* m, const float* lo, const float* hi, float* olo, float* ohi)
{
    double c[3], e[3];
    for (int a = 0; a < 3; ++a) {
        c[a] = 0.5 * ((double)lo[a] + hi[a]);
        e[a] = 0.5 * ((double)hi[a] - lo[a]);
    }
    for (int r = 0; r < 3; ++r) {
        double cr = m[12 + r], er = 0.0;
        for (int k = 0; k < 3; ++k) {
            cr += (double)m[4 * k + r] * c[k];
            er += std::fabs((double)m[4 * k + r]) * e[k];
        }
        const double l = cr - er, h = cr + er;
        const double pad = 1e-5 * std::max(h - l, std::max(std::fabs(l), std::fabs(h))) + 1e-30;
        olo[r] = std::nextafter((float)(l - pad), -INFINITY);
        ohi[r] = std::nextafter((float)(h + pad), INFINITY);
    }
}

} // namespace

bool buildTwoLevel(const StkVertex* vb, uint32_t nv, const uint32_t* ib, uint32_t ni, const StkInstance* inst,
                   uint32_t ninst, const StkBvhOptions& optsIn, TwoLevel& out, std::string& err)
{
    // spatial splits serve flat trees: the instance tree's records are boxes
    StkBvhOptions opts = optsIn;
    opts.spatialSplitAlpha = 0.0f;
    out = TwoLevel{};
    out.instMesh.assign(ninst, STK_INVALID_INDEX);
    out.instKey.assign(ninst, STK_INVALID_INDEX);
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> meshOf;
    for (uint32_t i = 0; i < ninst; ++i) {
        const StkInstance& ic = inst[i];
        if (ic.indexOffset < 0 || ic.indexCount < 0 || ic.indexCount % 3 != 0 ||
            (uint64_t)ic.indexOffset + (uint64_t)ic.indexCount > ni) {
            err = "instance " + std::to_string(i) + " has an index range outside the index buffer";
            return false;
        }
        if (ic.indexCount == 0) continue;
        const auto key = std::make_pair((uint32_t)ic.indexOffset, (uint32_t)ic.indexCount);
        auto it = meshOf.find(key);
        if (it == meshOf.end()) {
            it = meshOf.emplace(key, (uint32_t)out.meshOffset.size()).first;
            out.meshOffset.push_back(key.first);
            out.meshCount.push_back(key.second);
        }
        out.instMesh[i] = it->second;
    }
    const uint32_t nmesh = (uint32_t)out.meshOffset.size();
    if (nmesh == 0) return true;
    for (uint32_t m = 0; m < nmesh; ++m)
        for (uint32_t k = 0; k < out.meshCount[m]; ++k)
            if (ib[out.meshOffset[m] + k] >= nv) {
                err = "index buffer entry outside the vertex buffer";
                return false;
            }
    auto pos = [&](uint32_t m, uint32_t prim, int k) { return vb[ib[out.meshOffset[m] + 3 * prim + k]].position; };

    // one binary tree and one wide tree per mesh, in object space
    out.blasNodes.resize(nmesh);
    out.meshRoot.resize(nmesh);
    out.meshTriBase.assign(nmesh, 0u);
    out.meshTriCount.assign(nmesh, 0u);
    std::vector<std::array<float, 6>> meshBox(nmesh);
    for (uint32_t m = 0; m < nmesh; ++m) {
        const uint32_t ntri = out.meshCount[m] / 3;
        std::vector<BvhInputTri> tris(ntri);
        std::array<float, 6> box = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (uint32_t p = 0; p < ntri; ++p) {
            for (int k = 0; k < 3; ++k) {
                tris[p].v[k] = ld3(pos(m, p, k));
                for (int a = 0; a < 3; ++a) {
                    box[a] = std::min(box[a], pos(m, p, k)[a]);
                    box[3 + a] = std::max(box[3 + a], pos(m, p, k)[a]);
                }
            }
            tris[p].instId = 0;
            tris[p].primId = p;
        }
        meshBox[m] = box;
        Bvh b;
        if (!buildBvh(tris, opts, b, err)) return false;
        std::vector<StkWideNode> wide;
        std::vector<StkBlasTriangle> leaves;
        uint32_t depth = 0;
        auto leafBox = [&](uint32_t idx, float* lo, float* hi) {
            const uint32_t p = float_to_bits(b.nodes[idx].minBounds[0]);
            paddedTriangleBox(pos(m, p, 0), pos(m, p, 1), pos(m, p, 2), lo, hi);
        };
        auto emitLeaf = [&](uint32_t idx) {
            const uint32_t p = float_to_bits(b.nodes[idx].minBounds[0]);
            StkBlasTriangle t;
            memset(&t, 0, sizeof(t));
            for (int a = 0; a < 3; ++a) {
                t.p0[a] = pos(m, p, 0)[a];
                t.p1[a] = pos(m, p, 1)[a];
                t.p2[a] = pos(m, p, 2)[a];
            }
            t.primId = p;
            t.leaf = idx;
            leaves.push_back(t);
        };
        if (!collapseWide(b.nodes, leafBox, emitLeaf, wide, depth, err)) return false;
        const uint32_t nodeBase = (uint32_t)out.blasWide.size(), triBase = (uint32_t)out.blasTris.size();
        for (StkWideNode& w : wide) {
            w.childBase += nodeBase;
            w.triBase += triBase;
        }
        out.meshRoot[m] = nodeBase;
        out.meshTriBase[m] = triBase;
        out.meshTriCount[m] = (uint32_t)leaves.size();
        out.blasWide.insert(out.blasWide.end(), wide.begin(), wide.end());
        out.blasTris.insert(out.blasTris.end(), leaves.begin(), leaves.end());
        out.blasDepth = std::max(out.blasDepth, depth);
        out.blasNodes[m] = std::move(b.nodes);
    }

    // each instance's inverse transform as the walk uses it (rows of the
    // linear part and translation, rounded to float) and a bound on how far
    // it moves the image of a mesh point from the point itself
    out.instW.assign((size_t)ninst * 16, 0.0f);
    for (uint32_t i = 0; i < ninst; ++i) {
        const uint32_t m = out.instMesh[i];
        if (m == STK_INVALID_INDEX) continue;
        const float* M = inst[i].objectToWorld;
        double L[3][3], Li[3][3], t[3];
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) L[r][c] = M[4 * c + r];
            t[r] = M[12 + r];
        }
        const double det = L[0][0] * (L[1][1] * L[2][2] - L[1][2] * L[2][1]) -
                           L[0][1] * (L[1][0] * L[2][2] - L[1][2] * L[2][0]) +
                           L[0][2] * (L[1][0] * L[2][1] - L[1][1] * L[2][0]);
        if (!(std::fabs(det) > 1e-30) || !std::isfinite(det)) {
            err = "instance " + std::to_string(i) + " has a singular objectToWorld";
            return false;
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                const int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
                Li[r][c] = (L[r1][c1] * L[r2][c2] - L[r1][c2] * L[r2][c1]) / det;
            }
        float Wf[3][4];
        for (int r = 0; r < 3; ++r) {
            double wt = 0.0;
            for (int c = 0; c < 3; ++c) {
                Wf[r][c] = (float)Li[r][c];
                wt -= Li[r][c] * t[c];
            }
            Wf[r][3] = (float)wt;
        }
        // D(p) = (Wf Lf - I) p + (Wf t + wf) for p in the mesh box
        double worst = 0.0;
        for (int r = 0; r < 3; ++r) {
            double d = 0.0, b = (double)Wf[r][3];
            for (int c = 0; c < 3; ++c) {
                double a = -(r == c ? 1.0 : 0.0);
                for (int k = 0; k < 3; ++k) a += (double)Wf[r][k] * L[k][c];
                d += std::fabs(a) * std::max(std::fabs((double)meshBox[m][c]), std::fabs((double)meshBox[m][3 + c]));
                b += (double)Wf[r][c] * t[c];
            }
            worst = std::max(worst, d + std::fabs(b));
        }
        float* w = &out.instW[(size_t)i * 16];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) w[4 * r + c] = Wf[r][c];
        w[12] = std::nextafter((float)(2.0 * worst), INFINITY);
    }

    // the instances' world boxes, one binary tree and one wide tree over them
    std::vector<std::array<float, 6>> instBox(ninst);
    std::vector<BvhInputTri> prims;
    for (int a = 0; a < 3; ++a) {
        out.lo[a] = INFINITY;
        out.hi[a] = -INFINITY;
    }
    for (uint32_t i = 0; i < ninst; ++i) {
        const uint32_t m = out.instMesh[i];
        if (m == STK_INVALID_INDEX) continue;
        float* b = instBox[i].data();
        transformBox(inst[i].objectToWorld, meshBox[m].data(), meshBox[m].data() + 3, b, b + 3);
        BvhInputTri t;
        t.v[0] = mk3(b[0], b[1], b[2]);
        t.v[1] = mk3(b[3], b[4], b[5]);
        t.v[2] = t.v[0];
        t.instId = i;
        t.primId = 0;
        prims.push_back(t);
        for (int a = 0; a < 3; ++a) {
            out.lo[a] = std::min(out.lo[a], b[a]);
            out.hi[a] = std::max(out.hi[a], b[3 + a]);
        }
    }
    Bvh top;
    if (!buildBvh(prims, opts, top, err)) return false;
    auto leafBox = [&](uint32_t idx, float* lo, float* hi) {
        const float* b = instBox[(uint32_t)top.nodes[idx].instId].data();
        for (int a = 0; a < 3; ++a) {
            lo[a] = b[a];
            hi[a] = b[3 + a];
        }
    };
    auto emitLeaf = [&](uint32_t idx) {
        const uint32_t i = (uint32_t)top.nodes[idx].instId;
        StkTlasLeaf l;
        memset(&l, 0, sizeof(l));
        for (int a = 0; a < 3; ++a) {
            l.lo[a] = instBox[i][a];
            l.hi[a] = instBox[i][3 + a];
        }
        l.instId = i;
        l.blasRoot = out.meshRoot[out.instMesh[i]];
        l.key = idx;
        out.tlasLeaves.push_back(l);
    };
    if (!collapseWide(top.nodes, leafBox, emitLeaf, out.tlasWide, out.tlasDepth, err)) return false;
    for (uint32_t idx = 0; idx < (uint32_t)top.nodes.size(); ++idx)
        if ((uint32_t)top.nodes[idx].instId != STK_INVALID_INDEX) out.instKey[(uint32_t)top.nodes[idx].instId] = idx;
    out.tlasNodes = std::move(top.nodes);

    // emitter triangles in world space as getTriangle forms them (shaders/raytracing.h:134-153)
    for (uint32_t i = 0; i < ninst; ++i) {
        const uint32_t m = out.instMesh[i];
        if (m == STK_INVALID_INDEX || inst[i].lightId == -1) continue;
        const std::vector<StkBvhNode>& nodes = out.blasNodes[m];
        for (uint32_t idx = 0; idx < (uint32_t)nodes.size(); ++idx) {
            if ((uint32_t)nodes[idx].instId == STK_INVALID_INDEX) continue;
            const uint32_t p = float_to_bits(nodes[idx].minBounds[0]);
            const F3 v0 = xform4(inst[i].objectToWorld, ld3(pos(m, p, 0)), 1.0f);
            const F3 e0 = xform4(inst[i].objectToWorld, ld3(pos(m, p, 1)), 1.0f) - v0;
            const F3 e1 = xform4(inst[i].objectToWorld, ld3(pos(m, p, 2)), 1.0f) - v0;
            StkTriangle t;
            t.v0[0] = v0.x;
            t.v0[1] = v0.y;
            t.v0[2] = v0.z;
            t.e0[0] = e0.x;
            t.e0[1] = e0.y;
            t.e0[2] = e0.z;
            t.e1[0] = e1.x;
            t.e1[1] = e1.y;
            t.e1[2] = e1.z;
            t.leaf = idx;
            t.instId = i;
            t.primId = p;
            out.emitters.push_back(t);
        }
    }
    return true;
}

} // namespace stk
