// bvh_api.cpp - host-only BVH entry points of the C ABI (no device needed).
#include <new>

#include "../../include/stk/stk.h"
#include "bvh.hpp"

struct stk_bvh {
    stk::Bvh bvh;
};

struct stk_tlas {
    stk::TwoLevel tl;
};

extern "C" {

STK_API stk_status stk_bvh_build_host(const StkVertex* vb, uint32_t nv, const uint32_t* ib, uint32_t ni,
                                      const StkInstance* inst, uint32_t ninst, const StkBvhOptions* opts,
                                      stk_bvh** out)
{
    if (!out || (nv && !vb) || (ni && !ib) || (ninst && !inst)) return STK_E_INVALID;
    *out = nullptr;
    stk_bvh* b = new (std::nothrow) stk_bvh();
    if (!b) return STK_E_OOM;
    try {
        std::vector<stk::BvhInputTri> tris;
        std::string err;
        const StkBvhOptions o = opts ? *opts : stk::defaultBvhOptions();
        if (!stk::flattenInstances(vb, nv, ib, ni, inst, ninst, tris, err) || !stk::buildBvh(tris, o, b->bvh, err) ||
            !stk::packBvh(vb, nv, ib, ni, inst, ninst, b->bvh, err) || !stk::buildWide(inst, ninst, b->bvh, err)) {
            delete b;
            return STK_E_INVALID;
        }
    } catch (const std::bad_alloc&) {
        delete b;
        return STK_E_OOM;
    }
    *out = b;
    return STK_OK;
}

STK_API stk_status stk_bvh_from_nodes(const StkVertex* vb, uint32_t nv, const uint32_t* ib, uint32_t ni,
                                      const StkInstance* inst, uint32_t ninst, const StkBvhNode* nodes,
                                      uint32_t numNodes, stk_bvh** out)
{
    if (!out || (numNodes && !nodes)) return STK_E_INVALID;
    *out = nullptr;
    stk_bvh* b = new (std::nothrow) stk_bvh();
    if (!b) return STK_E_OOM;
    try {
        std::string err;
        b->bvh.nodes.assign(nodes, nodes + numNodes);
        if (!stk::packBvh(vb, nv, ib, ni, inst, ninst, b->bvh, err) || !stk::buildWide(inst, ninst, b->bvh, err)) {
            delete b;
            return STK_E_INVALID;
        }
    } catch (const std::bad_alloc&) {
        delete b;
        return STK_E_OOM;
    }
    *out = b;
    return STK_OK;
}

STK_API void stk_bvh_free(stk_bvh* b) { delete b; }
STK_API uint32_t stk_bvh_num_nodes(const stk_bvh* b) { return b ? (uint32_t)b->bvh.nodes.size() : 0; }
STK_API const StkBvhNode* stk_bvh_nodes(const stk_bvh* b) { return b ? b->bvh.nodes.data() : nullptr; }
STK_API uint32_t stk_bvh_num_packed(const stk_bvh* b) { return b ? (uint32_t)b->bvh.packed.size() : 0; }
STK_API const StkPackedNode* stk_bvh_packed(const stk_bvh* b) { return b ? b->bvh.packed.data() : nullptr; }
STK_API uint32_t stk_bvh_num_triangles(const stk_bvh* b) { return b ? (uint32_t)b->bvh.tris.size() : 0; }
STK_API const StkTriangle* stk_bvh_triangles(const stk_bvh* b) { return b ? b->bvh.tris.data() : nullptr; }
STK_API StkBvhHeader stk_bvh_header(const stk_bvh* b)
{
    StkBvhHeader h{};
    h.root = INT32_MIN;
    return b ? b->bvh.header : h;
}
STK_API uint32_t stk_bvh_depth(const stk_bvh* b) { return b ? b->bvh.depth : 0; }
STK_API uint32_t stk_bvh_num_wide(const stk_bvh* b) { return b ? (uint32_t)b->bvh.wide.size() : 0; }
STK_API const StkWideNode* stk_bvh_wide_nodes(const stk_bvh* b) { return b ? b->bvh.wide.data() : nullptr; }
STK_API const StkTriangle* stk_bvh_wide_triangles(const stk_bvh* b) { return b ? b->bvh.wideTris.data() : nullptr; }
STK_API uint32_t stk_bvh_wide_depth(const stk_bvh* b) { return b ? b->bvh.wideDepth : 0; }
STK_API uint32_t stk_bvh_num_emitters(const stk_bvh* b) { return b ? (uint32_t)b->bvh.emitters.size() : 0; }

STK_API stk_status stk_tlas_build_host(const StkVertex* vb, uint32_t nv, const uint32_t* ib, uint32_t ni,
                                       const StkInstance* inst, uint32_t ninst, const StkBvhOptions* opts,
                                       stk_tlas** out)
{
    if (!out || (nv && !vb) || (ni && !ib) || (ninst && !inst)) return STK_E_INVALID;
    *out = nullptr;
    stk_tlas* t = new (std::nothrow) stk_tlas();
    if (!t) return STK_E_OOM;
    try {
        std::string err;
        const StkBvhOptions o = opts ? *opts : stk::defaultBvhOptions();
        if (!stk::buildTwoLevel(vb, nv, ib, ni, inst, ninst, o, t->tl, err)) {
            delete t;
            return STK_E_INVALID;
        }
    } catch (const std::bad_alloc&) {
        delete t;
        return STK_E_OOM;
    }
    *out = t;
    return STK_OK;
}

STK_API void stk_tlas_free(stk_tlas* t) { delete t; }
STK_API uint32_t stk_tlas_num_meshes(const stk_tlas* t) { return t ? (uint32_t)t->tl.meshOffset.size() : 0; }
STK_API uint32_t stk_tlas_instance_mesh(const stk_tlas* t, uint32_t i)
{
    return (t && i < t->tl.instMesh.size()) ? t->tl.instMesh[i] : STK_INVALID_INDEX;
}
STK_API const StkBvhNode* stk_tlas_mesh_nodes(const stk_tlas* t, uint32_t mesh, uint32_t* count)
{
    if (!t || mesh >= t->tl.blasNodes.size()) {
        if (count) *count = 0;
        return nullptr;
    }
    if (count) *count = (uint32_t)t->tl.blasNodes[mesh].size();
    return t->tl.blasNodes[mesh].data();
}
STK_API const StkBvhNode* stk_tlas_top_nodes(const stk_tlas* t, uint32_t* count)
{
    if (count) *count = t ? (uint32_t)t->tl.tlasNodes.size() : 0;
    return t ? t->tl.tlasNodes.data() : nullptr;
}
STK_API const StkWideNode* stk_tlas_top_wide(const stk_tlas* t, uint32_t* count)
{
    if (count) *count = t ? (uint32_t)t->tl.tlasWide.size() : 0;
    return t ? t->tl.tlasWide.data() : nullptr;
}
STK_API const StkTlasLeaf* stk_tlas_top_leaves(const stk_tlas* t, uint32_t* count)
{
    if (count) *count = t ? (uint32_t)t->tl.tlasLeaves.size() : 0;
    return t ? t->tl.tlasLeaves.data() : nullptr;
}
STK_API const StkWideNode* stk_tlas_mesh_wide(const stk_tlas* t, uint32_t* count)
{
    if (count) *count = t ? (uint32_t)t->tl.blasWide.size() : 0;
    return t ? t->tl.blasWide.data() : nullptr;
}
STK_API const StkBlasTriangle* stk_tlas_mesh_triangles(const stk_tlas* t, uint32_t* count)
{
    if (count) *count = t ? (uint32_t)t->tl.blasTris.size() : 0;
    return t ? t->tl.blasTris.data() : nullptr;
}
STK_API const uint32_t* stk_tlas_instance_keys(const stk_tlas* t) { return t ? t->tl.instKey.data() : nullptr; }
STK_API const float* stk_tlas_instance_inverses(const stk_tlas* t) { return t ? t->tl.instW.data() : nullptr; }
STK_API const StkTriangle* stk_tlas_emitters(const stk_tlas* t, uint32_t* count)
{
    if (count) *count = t ? (uint32_t)t->tl.emitters.size() : 0;
    return t ? t->tl.emitters.data() : nullptr;
}
STK_API void stk_tlas_depths(const stk_tlas* t, uint32_t* top, uint32_t* mesh)
{
    if (top) *top = t ? t->tl.tlasDepth : 0;
    if (mesh) *mesh = t ? t->tl.blasDepth : 0;
}

} // extern "C"
