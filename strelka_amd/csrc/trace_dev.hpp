// trace_dev.hpp - device-side traversal for gfx950 (wave64).
//
// Two traversals share the reference's box and triangle tests bit for bit:
//  * closest_packed: near-first BVH2 traversal over 64-B packed nodes with
//    closest-t pruning and a per-lane LDS stack.  Result rule: smallest t, ties
//    to the smaller preorder leaf index - identical to the strict '<' first-in-
//    DFS acceptance of closestHit (shaders/raytracing.h:198-247) on the same
//    BVH.  Leaf boxes are never tested (the reference tests a leaf's triangle
//    whenever its parent box is entered, :211-231).
//  * closest_ref / any_ref: the reference's stackless skip-link walk.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/stk/stk.h"
#include "geom.hpp"

namespace stk {
namespace dev {

constexpr int kStackDepth = 32;
// Boxes whose entry t exceeds best*(1 + 2^-18) are pruned.
constexpr float kPruneSlack = 1.0f + 1.0f / 262144.0f;

struct BvhView {
    const float4* packed;         // 4 x float4 per StkPackedNode
    const float4* tris;           // 3 x float4 per StkTriangle
    const StkBvhNode* refNodes;   // skip-link nodes; leaves carry the triangle index in maxBounds[0]
    const StkInstance* instances; // for the any-hit light test
    StkBvhHeader hdr;
};

struct RayF {
    float ox, oy, oz, tmax;
    float dx, dy, dz;
    float ix, iy, iz; // 1/d, IEEE division (inf allowed, shaders/raytracing.h:200)
};

struct HitF {
    float u, v, t;
    uint32_t inst, prim;
    uint32_t leaf;
};

__device__ __forceinline__ RayF make_ray(float4 o, float4 d)
{
    RayF r;
    r.ox = o.x;
    r.oy = o.y;
    r.oz = o.z;
    r.tmax = o.w;
    r.dx = d.x;
    r.dy = d.y;
    r.dz = d.z;
    r.ix = 1.0f / d.x;
    r.iy = 1.0f / d.y;
    r.iz = 1.0f / d.z;
    return r;
}

// intersectRayBox (shaders/ray.h:35-48); min/max are IEEE minNum/maxNum.
__device__ __forceinline__ bool slab(const RayF& r, float lox, float hix, float loy, float hiy, float loz, float hiz,
                                     float& t)
{
    const float fx = (hix - r.ox) * r.ix;
    const float fy = (hiy - r.oy) * r.iy;
    const float fz = (hiz - r.oz) * r.iz;
    const float nx = (lox - r.ox) * r.ix;
    const float ny = (loy - r.oy) * r.iy;
    const float nz = (loz - r.oz) * r.iz;
    const float t1 = fminf(fmaxf(fx, nx), fminf(fmaxf(fy, ny), fmaxf(fz, nz)));
    const float t0 = fmaxf(fmaxf(fminf(fx, nx), fmaxf(fminf(fy, ny), fminf(fz, nz))), 0.0f);
    t = t0;
    return t1 >= t0;
}

// RayTriangleIntersect (shaders/raytracing.h:36-88).
__device__ __forceinline__ bool tri_test(const RayF& r, float4 a, float4 b, float4 c, float& t, float& u, float& v)
{
    t = 1e9f;
    const F3 dir = mk3(r.dx, r.dy, r.dz);
    const F3 e0 = mk3(b.x, b.y, b.z), e1 = mk3(c.x, c.y, c.z);
    const F3 pvec = cross(dir, e1);
    const float det = dot(e0, pvec);
    if (fabsf(det) < 1e-6f) return false;
    const float invDet = 1.0f / det;
    const F3 tvec = mk3(r.ox, r.oy, r.oz) - mk3(a.x, a.y, a.z);
    u = dot(tvec, pvec) * invDet;
    if (u < 0.0f || u > 1.0f) return false;
    const F3 qvec = cross(tvec, e0);
    v = dot(dir, qvec) * invDet;
    if (v < 0.0f || (u + v) > 1.0f) return false;
    t = dot(e1, qvec) * invDet;
    if (t < 1e-6f) return false;
    return true;
}

__device__ __forceinline__ void leaf_closest(const BvhView& bvh, const RayF& r, uint32_t tri, bool& found, HitF& h)
{
    const float4* p = bvh.tris + 3 * (size_t)tri;
    const float4 a = p[0], b = p[1], c = p[2];
    float t, u, v;
    if (tri_test(r, a, b, c, t, u, v) && t < r.tmax) {
        const uint32_t leaf = float_to_bits(a.w);
        const bool acc = found ? (t < h.t || (t == h.t && leaf < h.leaf)) : (t < 1e9f);
        if (acc) {
            found = true;
            h.t = t;
            h.u = u;
            h.v = v;
            h.leaf = leaf;
            h.inst = float_to_bits(b.w);
            h.prim = float_to_bits(c.w);
        }
    }
}

__device__ __forceinline__ void miss(HitF& h)
{
    h.u = 0.0f;
    h.v = 0.0f;
    h.t = 1e9f;
    h.inst = STK_INVALID_INDEX;
    h.prim = STK_INVALID_INDEX;
    h.leaf = STK_INVALID_INDEX;
}

// Near-first packed traversal, resumable: a walk is its ray, the node `cur`
// (-1 when done), the stack depth `sp`, and the best hit so far.  `stack`
// points at this lane's first slot; the slots are `stride` words apart (LDS
// layout [depth][lane], conflict free).
struct PackedWalk {
    int cur;
    int sp;
    bool found;
};

// Starts a walk: handles the empty tree, a single-leaf tree and the root box.
__device__ __forceinline__ void closest_packed_begin(const BvhView& bvh, const StkBvhHeader& hd, const RayF& r,
                                                     PackedWalk& w, HitF& h)
{
    miss(h);
    w.found = false;
    w.cur = -1;
    w.sp = 0;
    if (hd.root == INT32_MIN) return;
    if (hd.root < 0) {
        leaf_closest(bvh, r, (uint32_t)~hd.root, w.found, h);
        return;
    }
    float t0;
    if (slab(r, hd.lo[0], hd.hi[0], hd.lo[1], hd.hi[1], hd.lo[2], hd.hi[2], t0) && t0 <= r.tmax) w.cur = hd.root;
}

// Visits at most `budget` nodes (budget < 0: until the walk ends).
__device__ __forceinline__ void closest_packed_steps(const BvhView& bvh, const RayF& r, uint32_t* stack, int stride,
                                                     PackedWalk& w, HitF& h, int budget)
{
    int cur = w.cur, sp = w.sp;
    bool found = w.found;
    while (cur >= 0 && budget-- != 0) {
        const float4* p = bvh.packed + 4 * (size_t)cur;
        const float4 bx = p[0], by = p[1], bz = p[2];
        const int4 ch = *reinterpret_cast<const int4*>(p + 3);
        const int c0 = ch.x, c1 = ch.y;
        if (c0 < 0) leaf_closest(bvh, r, (uint32_t)~c0, found, h);
        if (c1 < 0) leaf_closest(bvh, r, (uint32_t)~c1, found, h);
        const float lim = (found ? h.t : 1e9f) * kPruneSlack;
        bool ok0 = false, ok1 = false;
        float ta = 0.0f, tb = 0.0f;
        if (c0 >= 0) ok0 = slab(r, bx.x, bx.y, by.x, by.y, bz.x, bz.y, ta) && ta <= r.tmax && !(ta > lim);
        if (c1 >= 0) ok1 = slab(r, bx.z, bx.w, by.z, by.w, bz.z, bz.w, tb) && tb <= r.tmax && !(tb > lim);
        if (ok0 && ok1) {
            const bool oneNear = tb < ta;
            stack[sp * stride] = (uint32_t)(oneNear ? c0 : c1);
            ++sp;
            cur = oneNear ? c1 : c0;
        } else if (ok0) {
            cur = c0;
        } else if (ok1) {
            cur = c1;
        } else {
            cur = sp > 0 ? (int)stack[--sp * stride] : -1;
        }
    }
    w.cur = cur;
    w.sp = sp;
    w.found = found;
}

__device__ __forceinline__ bool closest_packed(const BvhView& bvh, const StkBvhHeader& hd, const RayF& r,
                                               uint32_t* stack, int stride, HitF& h)
{
    PackedWalk w;
    closest_packed_begin(bvh, hd, r, w, h);
    closest_packed_steps(bvh, r, stack, stride, w, h, -1);
    return w.found;
}

// closestHit (shaders/raytracing.h:198-247), reference order, unpruned,
// starting at skip-link node `root` (0 for a single scene).
__device__ __forceinline__ bool closest_ref(const BvhView& bvh, const StkBvhHeader& hd, uint32_t root, const RayF& r,
                                            HitF& h)
{
    miss(h);
    if (hd.root == INT32_MIN) return false;
    float minHit = 1e9f;
    bool found = false;
    uint32_t idx = root;
    while (idx != STK_INVALID_INDEX) {
        const float4* np = reinterpret_cast<const float4*>(bvh.refNodes + idx);
        const float4 lo = np[0], hi = np[1];
        const uint32_t inst = float_to_bits(lo.w);
        const uint32_t skip = float_to_bits(hi.w);
        if (inst != STK_INVALID_INDEX) {
            const uint32_t tri = float_to_bits(hi.x);
            const float4* p = bvh.tris + 3 * (size_t)tri;
            const float4 a = p[0], b = p[1], c = p[2];
            float t, u, v;
            if (tri_test(r, a, b, c, t, u, v) && t < r.tmax && t < minHit) {
                minHit = t;
                h.t = t;
                h.u = u;
                h.v = v;
                h.inst = inst;
                h.prim = float_to_bits(lo.x);
                h.leaf = idx;
                found = true;
            }
            idx = skip;
            continue;
        }
        float bt;
        if (slab(r, lo.x, hi.x, lo.y, hi.y, lo.z, hi.z, bt)) {
            idx = (bt > r.tmax) ? skip : idx + 1;
            continue;
        }
        idx = skip;
    }
    return found;
}

// anyHit (shaders/raytracing.h:249-286): first leaf in DFS order hit with
// t < tmax.  h.inst/h.prim receive that leaf's ids.
__device__ __forceinline__ bool any_ref(const BvhView& bvh, const StkBvhHeader& hd, uint32_t root, const RayF& r,
                                        HitF& h)
{
    miss(h);
    if (hd.root == INT32_MIN) return false;
    uint32_t idx = root;
    while (idx != STK_INVALID_INDEX) {
        const float4* np = reinterpret_cast<const float4*>(bvh.refNodes + idx);
        const float4 lo = np[0], hi = np[1];
        const uint32_t inst = float_to_bits(lo.w);
        const uint32_t skip = float_to_bits(hi.w);
        if (inst != STK_INVALID_INDEX) {
            const uint32_t tri = float_to_bits(hi.x);
            const float4* p = bvh.tris + 3 * (size_t)tri;
            const float4 a = p[0], b = p[1], c = p[2];
            float t, u, v;
            const bool isect = tri_test(r, a, b, c, t, u, v);
            if (isect && t < r.tmax) {
                h.t = t;
                h.inst = inst;
                h.prim = float_to_bits(lo.x);
                h.leaf = idx;
                return true;
            }
            idx = skip;
            continue;
        }
        float bt;
        if (slab(r, lo.x, hi.x, lo.y, hi.y, lo.z, hi.z, bt)) {
            idx = (bt > r.tmax) ? skip : idx + 1;
            continue;
        }
        idx = skip;
    }
    return false;
}

// anyHit with the reference's result - the first leaf in DFS preorder whose
// triangle is hit with t < tmax - over the packed nodes: children are visited
// left first and a leaf right child waits on the stack while its inner left
// sibling's subtree is walked, so leaves are tested in preorder and the walk
// stops at the first hit, two boxes per 64-B fetch.
__device__ __forceinline__ bool any_packed(const BvhView& bvh, const StkBvhHeader& hd, const RayF& r,
                                           uint32_t* stack, int stride, HitF& h)
{
    miss(h);
    if (hd.root == INT32_MIN) return false;
    int next = hd.root;
    if (hd.root >= 0) {
        float t0;
        if (!(slab(r, hd.lo[0], hd.hi[0], hd.lo[1], hd.hi[1], hd.lo[2], hd.hi[2], t0) && t0 <= r.tmax)) return false;
    }
    int sp = 0;
    for (;;) {
        // resolve leaves in DFS order until an inner node is next
        while (next < 0) {
            const uint32_t tri = (uint32_t)~next;
            const float4* p = bvh.tris + 3 * (size_t)tri;
            const float4 a = p[0], b = p[1], c = p[2];
            float t, u, v;
            if (tri_test(r, a, b, c, t, u, v) && t < r.tmax) {
                h.t = t;
                h.leaf = float_to_bits(a.w);
                h.inst = float_to_bits(b.w);
                h.prim = float_to_bits(c.w);
                return true;
            }
            if (sp == 0) return false;
            next = (int)stack[--sp * stride];
        }
        const float4* p = bvh.packed + 4 * (size_t)next;
        const float4 bx = p[0], by = p[1], bz = p[2];
        const int4 ch = *reinterpret_cast<const int4*>(p + 3);
        const int c0 = ch.x, c1 = ch.y;
        float ta, tb;
        const bool ok0 = c0 < 0 || (slab(r, bx.x, bx.y, by.x, by.y, bz.x, bz.y, ta) && ta <= r.tmax);
        const bool ok1 = c1 < 0 || (slab(r, bx.z, bx.w, by.z, by.w, bz.z, bz.w, tb) && tb <= r.tmax);
        if (ok0) {
            if (ok1) stack[sp++ * stride] = (uint32_t)c1;
            next = c0;
        } else if (ok1) {
            next = c1;
        } else {
            if (sp == 0) return false;
            next = (int)stack[--sp * stride];
        }
    }
}

// XCD-aware block order: hardware deals workgroups round-robin over the 8
// XCDs; remap so XCD k owns a contiguous range of logical blocks (and so a
// contiguous range of rays, which share BVH nodes in that XCD's L2).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t g)
{
    const uint32_t q = g / 8, rem = g % 8, xcd = b % 8, local = b / 8;
    return (xcd < rem) ? xcd * (q + 1) + local : rem * (q + 1) + (xcd - rem) * q + local;
}

} // namespace dev
using dev::BvhView;
} // namespace stk
