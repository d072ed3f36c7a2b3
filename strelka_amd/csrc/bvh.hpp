// bvh.hpp - host BVH builder and packer.
#pragma once

#include <cstdint>
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
    // compressed 8-wide layout (buildWide)
    std::vector<StkWideNode> wide;     // node 0 is the root
    std::vector<StkTriangle> wideTris; // leaf triangles in wide-node slot order
    std::vector<StkTriangle> emitters; // triangles of emitter instances (lightId != -1)
    uint32_t wideDepth = 0;            // inner-node levels of the wide tree
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

// Collapses the packed binary BVH into the compressed 8-wide layout and
// collects the emitter triangles.  Needs packBvh's output.
bool buildWide(const StkInstance* inst, uint32_t ninst, Bvh& bvh, std::string& err);

// Two-level acceleration structure: one tree per distinct mesh (index range)
// in object space, one over the instances' world boxes.
struct TwoLevel {
    std::vector<uint32_t> meshOffset, meshCount;      // index range of each mesh
    std::vector<uint32_t> instMesh;                   // mesh of each instance (STK_INVALID_INDEX: no triangles)
    std::vector<std::vector<StkBvhNode>> blasNodes;   // binary skip-link tree per mesh, object space
    std::vector<StkWideNode> blasWide;                // every mesh's wide tree, indices rebased
    std::vector<StkBlasTriangle> blasTris;            // leaves of blasWide, slot order
    std::vector<uint32_t> meshRoot;                   // wide root of each mesh in blasWide
    std::vector<uint32_t> meshTriBase, meshTriCount;  // each mesh's leaves in blasTris
    uint32_t blasDepth = 0;                           // deepest mesh tree (wide levels)
    std::vector<StkBvhNode> tlasNodes;                // binary skip-link tree over instance boxes
    std::vector<StkWideNode> tlasWide;
    std::vector<StkTlasLeaf> tlasLeaves;              // leaves of tlasWide, slot order
    uint32_t tlasDepth = 0;
    std::vector<uint32_t> instKey;                    // TLAS preorder leaf index of each instance
    std::vector<float> instW;                         // 16 per instance: inverse rows (3 x 4), {pad, 0, 0, 0}
    std::vector<StkTriangle> emitters;                // world triangles of emitter instances (leaf = mesh-tree leaf)
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};       // scene box
};

// Builds the two-level structure with the binned-SAH builder per mesh and over
// the instances.  Instances with an empty index range are left out.
bool buildTwoLevel(const StkVertex* vb, uint32_t nv, const uint32_t* ib, uint32_t ni, const StkInstance* inst,
                   uint32_t ninst, const StkBvhOptions& opts, TwoLevel& out, std::string& err);

StkBvhOptions defaultBvhOptions();

} // namespace stk
