// bvh.hpp - host BVH builder and packer.
#pragma once

#include <string>
#include <vector>

#include "../../include/stk/stk.h"
#include "geom.hpp"

namespace stk {

// One world-space triangle of the BVH input; the three positions of
// BVHInputPosition (include/render/bvh.h:31-36) with their (instId, primId).
struct BvhInputTri {
    F3 v[3];
    uint32_t instId;
    uint32_t primId;
};

struct Bvh {
    std::vector<StkBvhNode> nodes;     // reference skip-link layout, preorder
    std::vector<StkPackedNode> packed; // BVH2 nodes of the GPU traversal
    std::vector<StkTriangle> tris;     // pre-transformed triangles, leaf order
    std::vector<uint32_t> leafToTri;   // preorder node index -> triangle index
    StkBvhHeader header{};
    uint32_t depth = 0;
};

// createBvhBuffer (src/render/ptrender.cpp:460-512): world-space positions of
// every instance's triangles, primId = triangle index within its instance.
bool flattenInstances(const StkVertex* vb, uint32_t nv, const uint32_t* ib, uint32_t ni, const StkInstance* inst,
                      uint32_t ninst, std::vector<BvhInputTri>& out, std::string& err);

// Binned-SAH build, one triangle per leaf, preorder + skip links.
bool buildBvh(const std::vector<BvhInputTri>& tris, const StkBvhOptions& opts, Bvh& out, std::string& err);

// Validates a skip-link array and derives packed nodes and triangles from it.
bool packBvh(const StkVertex* vb, uint32_t nv, const uint32_t* ib, uint32_t ni, const StkInstance* inst,
             uint32_t ninst, Bvh& bvh, std::string& err);

StkBvhOptions defaultBvhOptions();

} // namespace stk
