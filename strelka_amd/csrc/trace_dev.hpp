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
//  * closest_wide / visible_wide: the compressed 8-wide tree (StkWideNode),
//    80-B nodes with 8-bit child boxes, node-group stack entries.  Same
//    result rule as closest_packed; visible_wide answers the shadow-ray
//    question of the path tracer with the reference's any-hit semantics
//    (first leaf in preorder decides) in any visiting order.
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
    const float4* wide;           // 5 x float4 per StkWideNode
    const float4* wideTris;       // 3 x float4 per StkTriangle, wide slot order
    const float4* emitters;       // 3 x float4 per emitter triangle
    uint32_t wideRoot;            // STK_INVALID_INDEX: no usable wide tree
    uint32_t emitCount;
    uint32_t packedOk;            // 0: deeper than the packed walks' stack (kStackDepth)
    uint32_t wideDepth;           // levels of the wide tree (a walk holds at most one stack entry per level)
    // two-level structure (tlas == nullptr: none); emitters then holds the
    // emitter instances' world triangles (leaf = mesh-tree leaf)
    const float4* tlas;           // 5 x float4 per StkWideNode, root at 0
    const float4* tlasLeaves;     // 3 x float4 per StkTlasLeaf
    const float4* blas;           // 5 x float4 per StkWideNode, all meshes
    const float4* blasTris;       // 3 x float4 per StkBlasTriangle
    const uint32_t* instKey;      // TLAS preorder leaf index of each instance
    const float4* instW;          // 4 x float4 per instance: the rows of the inverse of
                                  // objectToWorld (linear part | translation), then {pad, 0, 0, 0}
    // optional: every instance's world triangles as world_triangle forms them
    // (3 x float4 each), mesh-tree leaf t of instance i at t + instTriOff[i]
    // (modulo 2^32; t indexes blasTris); nullptr: formed per test
    const float4* worldTris;
    const uint32_t* instTriOff;
};

constexpr int kWideStack = 16; // node-group entries; the builder's wide depth must not exceed it
// Stack entries of the wide walks over trees of at most kWideStackShallow + 1
// levels (a walk holds at most one entry per level below the root): 5.5 KiB
// of LDS per wave instead of 8, so a 72-register walk runs 7 waves per SIMD.
constexpr int kWideStackShallow = 11;

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

// Whether a triangle of a non-emitter instance is hit with t < tmax: the
// skip-link walk of anyHit that does not stop at emitter triangles (the
// shadow test of the corrected estimator, where only non-emitters occlude).
__device__ __forceinline__ bool occluded_ref(const BvhView& bvh, const StkBvhHeader& hd, uint32_t root,
                                             const RayF& r)
{
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
            if (tri_test(r, a, b, c, t, u, v) && t < r.tmax && bvh.instances[inst].lightId == -1) return true;
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

// ---- compressed 8-wide traversal ----------------------------------------------

// One visited wide node: which slots' boxes the ray enters, their entry
// distances, and the node's child ranges.  Slots are evaluated in a fixed
// order (so byte lanes and register indices are static and the code is
// branch-free); octant order only decides the order in which a pushed node
// group is popped.
struct WideVisit {
    uint32_t childBase, triBase, imask, leafMask;
    uint32_t innerHits, leafHits; // bit s: the box of slot s is entered
    float t[8];                   // entry distance of slot s
};

// Byte K of w as a float (one v_cvt_f32_ubyteK).
template <int K>
__device__ __forceinline__ float ubyte(uint32_t w)
{
    return (float)((w >> (8 * K)) & 0xFFu);
}

// The slab test of slot S: per axis the near and far planes' distances as
// one fused multiply-add each, t = q * A + B, with the node's per-axis terms
// A and B (near and far) of wide_axis.  nearW[2a] / nearW[2a + 1] hold axis
// a's near-plane bytes of slots 0-3 / 4-7, farW its far-plane bytes.
template <int S>
__device__ __forceinline__ void wide_slot(const uint32_t (&nearW)[6], const uint32_t (&farW)[6], const float (&A)[3],
                                          const float (&BN)[3], const float (&BF)[3], float L, uint32_t& hits,
                                          float& t)
{
    const int w = S < 4 ? 0 : 1;
    const float nx = fmaf(ubyte<S & 3>(nearW[0 + w]), A[0], BN[0]);
    const float ny = fmaf(ubyte<S & 3>(nearW[2 + w]), A[1], BN[1]);
    const float nz = fmaf(ubyte<S & 3>(nearW[4 + w]), A[2], BN[2]);
    const float fx = fmaf(ubyte<S & 3>(farW[0 + w]), A[0], BF[0]);
    const float fy = fmaf(ubyte<S & 3>(farW[2 + w]), A[1], BF[1]);
    const float fz = fmaf(ubyte<S & 3>(farW[4 + w]), A[2], BF[2]);
    const float t0 = fmaxf(fmaxf(nx, ny), fmaxf(nz, 0.0f));
    const float t1 = fminf(fminf(fx, fy), fminf(fz, L));
    hits |= (t0 <= t1) ? (1u << S) : 0u;
    t = t0;
}

// Per-axis terms of a visited node for the ray: A = scale x inverse
// direction, and the near and far planes' offsets B = (plane origin - ray
// origin) x inverse direction moved out by a bound on the rounding of the
// fused form and of the decode the builder checks its boxes with (u = 2^-24):
//   |t - exact| <= u (|origin| + 255 scale) |inv|   the decode's rounding
//                + u 255 |A|                        A's rounding, times q
//                + 2u |B| + u |t|                   B's roundings, the fma's
//              <= u (765 scale + 4 |origin| + 3 |ray origin|) |inv|,
// taken twice; so a slot's test never rejects a box the exact slab test of
// its decoded box accepts.  The inverse direction is clamped to +-2^80 (a
// zero component's infinite inverse would make q A + B undefined): for a
// ray parallel to an axis's planes every distance on that axis is then
// beyond 2^30 (further than any hit) unless the origin lies within the
// margin of a plane, so the axis still accepts exactly the slabs that hold
// the origin, and nothing overflows.  The near plane is the low one unless
// the inverse direction is negative.  PAD: the plane origins moved out by
// pad (the widened boxes of a mesh tree).
constexpr float kInvClamp = 1208925819614629174706176.0f; // 2^80
template <bool PAD>
__device__ __forceinline__ void wide_axis(float origin, float pad, float scale, float ro, float inv, float& A,
                                          float& BN, float& BF)
{
    A = scale * inv;
    const float mag = PAD ? fabsf(origin) + pad : fabsf(origin);
    const float e = fmaf(765.0f, scale, fmaf(4.0f, mag, 3.0f * fabsf(ro))) * (fabsf(inv) * (1.0f / 8388608.0f));
    if (PAD) {
        const bool neg = signbit(inv);
        BN = ((neg ? origin + pad : origin - pad) - ro) * inv - e;
        BF = ((neg ? origin - pad : origin + pad) - ro) * inv + e;
    } else {
        const float b = (origin - ro) * inv;
        BN = b - e;
        BF = b + e;
    }
}

// PAD: every child box widened by pad[a] per axis (the plane origins moved
// out by pad before the decode).  pad = 0 decodes exactly like the unwidened
// form.
template <bool PAD>
__device__ __forceinline__ void wide_visit_t(const float4* node, const RayF& r, float lim, const float (&pad)[3],
                                             WideVisit& v)
{
    const float4 w0 = node[0], w1 = node[1];
    const uint4 q0 = *reinterpret_cast<const uint4*>(node + 2); // qlo x[0..7], qlo y[0..7]
    const uint4 q1 = *reinterpret_cast<const uint4*>(node + 3); // qlo z[0..7], qhi x[0..7]
    const uint4 q2 = *reinterpret_cast<const uint4*>(node + 4); // qhi y[0..7], qhi z[0..7]
    const uint32_t meta = float_to_bits(w0.w);
    // scale 2^(e - 127) built from the exponent byte: exact powers of two
    const float sx = bits_to_float((meta & 0xFFu) << 23);
    const float sy = bits_to_float(((meta >> 8) & 0xFFu) << 23);
    const float sz = bits_to_float(((meta >> 16) & 0xFFu) << 23);
    v.imask = meta >> 24;
    v.childBase = float_to_bits(w1.x);
    v.triBase = float_to_bits(w1.y);
    const uint32_t vmask = float_to_bits(w1.z) & 0xFFu;
    v.leafMask = vmask & ~v.imask;
    float A[3], BN[3], BF[3];
    const float ix = fminf(fmaxf(r.ix, -kInvClamp), kInvClamp);
    const float iy = fminf(fmaxf(r.iy, -kInvClamp), kInvClamp);
    const float iz = fminf(fmaxf(r.iz, -kInvClamp), kInvClamp);
    wide_axis<PAD>(w0.x, pad[0], sx, r.ox, ix, A[0], BN[0], BF[0]);
    wide_axis<PAD>(w0.y, pad[1], sy, r.oy, iy, A[1], BN[1], BF[1]);
    wide_axis<PAD>(w0.z, pad[2], sz, r.oz, iz, A[2], BN[2], BF[2]);
    const bool nx = signbit(ix), ny = signbit(iy), nz = signbit(iz);
    const uint32_t nearW[6] = {nx ? q1.z : q0.x, nx ? q1.w : q0.y, ny ? q2.x : q0.z,
                               ny ? q2.y : q0.w, nz ? q2.z : q1.x, nz ? q2.w : q1.y};
    const uint32_t farW[6] = {nx ? q0.x : q1.z, nx ? q0.y : q1.w, ny ? q0.z : q2.x,
                              ny ? q0.w : q2.y, nz ? q1.x : q2.z, nz ? q1.y : q2.w};
    const float L = fminf(r.tmax, lim);
    uint32_t hits = 0;
    wide_slot<0>(nearW, farW, A, BN, BF, L, hits, v.t[0]);
    wide_slot<1>(nearW, farW, A, BN, BF, L, hits, v.t[1]);
    wide_slot<2>(nearW, farW, A, BN, BF, L, hits, v.t[2]);
    wide_slot<3>(nearW, farW, A, BN, BF, L, hits, v.t[3]);
    wide_slot<4>(nearW, farW, A, BN, BF, L, hits, v.t[4]);
    wide_slot<5>(nearW, farW, A, BN, BF, L, hits, v.t[5]);
    wide_slot<6>(nearW, farW, A, BN, BF, L, hits, v.t[6]);
    wide_slot<7>(nearW, farW, A, BN, BF, L, hits, v.t[7]);
    hits &= vmask;
    v.innerHits = hits & v.imask;
    v.leafHits = hits & v.leafMask;
}

__device__ __forceinline__ void wide_visit(const float4* node, const RayF& r, float lim, WideVisit& v)
{
    const float none[3] = {0.0f, 0.0f, 0.0f};
    wide_visit_t<false>(node, r, lim, none, v);
}

__device__ __forceinline__ uint32_t ray_octant(const RayF& r)
{
    return (signbit(r.dx) ? 1u : 0u) | (signbit(r.dy) ? 2u : 0u) | (signbit(r.dz) ? 4u : 0u);
}

__device__ __forceinline__ uint32_t rank_below(uint32_t mask, uint32_t s) { return __popc(mask & ((1u << s) - 1u)); }

// Bit s of m moved to bit s ^ oct (an involution on 8-bit masks).
__device__ __forceinline__ uint32_t xor_permute(uint32_t m, uint32_t oct)
{
    if (oct & 1u) m = ((m & 0x55u) << 1) | ((m >> 1) & 0x55u);
    if (oct & 2u) m = ((m & 0x33u) << 2) | ((m >> 2) & 0x33u);
    if (oct & 4u) m = ((m & 0x0Fu) << 4) | ((m >> 4) & 0x0Fu);
    return m;
}

// Next node of the walk: the nearest entered inner child of the visited node,
// the rest of its entered inner children pushed as one node-group entry
// (childBase, imask | remaining << 8, remaining in visiting order k = s ^ oct);
// otherwise the next child of the top group.  Returns false when the walk is
// over (the stack is back at `base`: a nested walk stacks on its caller's).
__device__ __forceinline__ bool wide_next(const WideVisit& v, uint32_t oct, uint32_t* sA, uint32_t* sB, int stride,
                                          int& sp, uint32_t& node, int base = 0)
{
    if (v.innerHits) {
        uint32_t m = v.innerHits;
        int sn = 0;
        float tn = INFINITY;
#pragma unroll
        for (int s = 0; s < 8; ++s) // static indices keep t[] in registers
            if (((m >> s) & 1u) && v.t[s] < tn) {
                tn = v.t[s];
                sn = s;
            }
        m &= ~(1u << sn);
        if (m) {
            sA[sp * stride] = v.childBase;
            sB[sp * stride] = v.imask | (xor_permute(m, oct) << 8);
            ++sp;
        }
        node = v.childBase + rank_below(v.imask, (uint32_t)sn);
        return true;
    }
    if (sp == base) return false;
    const uint32_t first = sA[(sp - 1) * stride];
    const uint32_t word = sB[(sp - 1) * stride];
    const uint32_t imask = word & 0xFFu;
    uint32_t m = word >> 8;
    const int k = __ffs(m) - 1;
    m &= m - 1u;
    if (m)
        sB[(sp - 1) * stride] = imask | (m << 8);
    else
        --sp;
    node = first + rank_below(imask, (uint32_t)k ^ oct);
    return true;
}

// Closest hit over the wide tree rooted at `root`.  sA/sB: this lane's first
// slots of two LDS stacks of kWideStack words, `stride` words apart.
__device__ __forceinline__ bool closest_wide(const BvhView& bvh, uint32_t root, const RayF& r, uint32_t* sA,
                                             uint32_t* sB, int stride, HitF& h)
{
    miss(h);
    bool found = false;
    if (root == STK_INVALID_INDEX) return false;
    const uint32_t oct = ray_octant(r);
    int sp = 0;
    uint32_t node = root;
    for (;;) {
        WideVisit v;
        wide_visit(bvh.wide + 5 * (size_t)node, r, (found ? h.t : 1e9f) * kPruneSlack, v);
        for (uint32_t m = v.leafHits; m; m &= m - 1u) {
            const uint32_t s = (uint32_t)(__ffs(m) - 1);
            const uint32_t tri = v.triBase + rank_below(v.leafMask, s);
            const float4* p = bvh.wideTris + 3 * (size_t)tri;
            const float4 a = p[0], b = p[1], c = p[2];
            float t, u, w;
            if (tri_test(r, a, b, c, t, u, w) && t < r.tmax) {
                const uint32_t leaf = float_to_bits(a.w);
                const bool acc = found ? (t < h.t || (t == h.t && leaf < h.leaf)) : (t < 1e9f);
                if (acc) {
                    found = true;
                    h.t = t;
                    h.u = u;
                    h.v = w;
                    h.leaf = leaf;
                    h.inst = float_to_bits(b.w);
                    h.prim = float_to_bits(c.w);
                }
            }
        }
        if (!wide_next(v, oct, sA, sB, stride, sp, node)) return found;
    }
}

// Shadow-ray visibility with the reference's any-hit rule: anyHit returns the
// first leaf in preorder whose triangle is hit with t < tmax, and the light
// counts as visible when there is none or when that leaf belongs to an emitter
// (shaders/pathtracerMdl.hlsl:31-92).  Equivalently: occluded iff some
// non-emitter leaf is hit before, in preorder, every hit emitter leaf.  The
// emitter triangles (few) are tested first; the walk then stops at the first
// qualifying occluder, in any order.
__device__ __forceinline__ bool visible_wide(const BvhView& bvh, uint32_t root, const float4* emit, uint32_t emitCount,
                                             const RayF& r, uint32_t* sA, uint32_t* sB, int stride)
{
    uint32_t lightLeaf = STK_INVALID_INDEX;
    for (uint32_t e = 0; e < emitCount; ++e) {
        const float4* p = emit + 3 * (size_t)e;
        const float4 a = p[0], b = p[1], c = p[2];
        float t, u, w;
        if (tri_test(r, a, b, c, t, u, w) && t < r.tmax) lightLeaf = min(lightLeaf, float_to_bits(a.w));
    }
    if (root == STK_INVALID_INDEX) return true;
    const uint32_t oct = ray_octant(r);
    int sp = 0;
    uint32_t node = root;
    for (;;) {
        WideVisit v;
        wide_visit(bvh.wide + 5 * (size_t)node, r, 1e9f * kPruneSlack, v);
        for (uint32_t m = v.leafHits; m; m &= m - 1u) {
            const uint32_t s = (uint32_t)(__ffs(m) - 1);
            const uint32_t tri = v.triBase + rank_below(v.leafMask, s);
            const float4* p = bvh.wideTris + 3 * (size_t)tri;
            const float4 a = p[0], b = p[1], c = p[2];
            if (float_to_bits(a.w) >= lightLeaf) continue; // preorder after a hit emitter: cannot decide
            float t, u, w;
            if (tri_test(r, a, b, c, t, u, w) && t < r.tmax &&
                bvh.instances[float_to_bits(b.w)].lightId == -1)
                return false;
        }
        if (!wide_next(v, oct, sA, sB, stride, sp, node)) return true;
    }
}

// ---- two-level walks ---------------------------------------------------------------

// Exit distance of the ray from the box [lo, hi] (intersectRayBox's t1).
__device__ __forceinline__ float slab_exit(const RayF& r, float4 lo, float4 hi)
{
    const float fx = (hi.x - r.ox) * r.ix, nx = (lo.x - r.ox) * r.ix;
    const float fy = (hi.y - r.oy) * r.iy, ny = (lo.y - r.oy) * r.iy;
    const float fz = (hi.z - r.oz) * r.iz, nz = (lo.z - r.oz) * r.iz;
    return fminf(fmaxf(fx, nx), fminf(fmaxf(fy, ny), fmaxf(fz, nz)));
}

// The ray in an instance's object space: origin and direction through W, the
// inverse of objectToWorld the builder computes (rows: linear part |
// translation), the same parameter t, and the widening of the mesh tree's
// boxes that keeps its walk conservative.  The object image of a world point
// o + t d (t up to T, the exit from the instance's world box) is off by the
// rounding of the transformed ray, and the world triangle getTriangle forms
// (with its hit test's rounding) lies off the object triangle by float
// rounding of M p and of the hit point; both stay below
// 2^-16 (|W| (|o| + T |d|) + |w|) per axis.  W M differs from the identity
// by the inverse's own rounding: the builder bounds what that moves the
// mesh's points by in the record's pad term, added on top.
struct InstRay {
    RayF r;
    float pad[3];
};

__device__ __forceinline__ void object_ray(const float4* W, const RayF& r, float T, InstRay& o)
{
    const float4 w0 = W[0], w1 = W[1], w2 = W[2], w3 = W[3];
    o.r.ox = (w0.x * r.ox + w0.y * r.oy) + (w0.z * r.oz + w0.w);
    o.r.oy = (w1.x * r.ox + w1.y * r.oy) + (w1.z * r.oz + w1.w);
    o.r.oz = (w2.x * r.ox + w2.y * r.oy) + (w2.z * r.oz + w2.w);
    o.r.tmax = r.tmax;
    o.r.dx = (w0.x * r.dx + w0.y * r.dy) + w0.z * r.dz;
    o.r.dy = (w1.x * r.dx + w1.y * r.dy) + w1.z * r.dz;
    o.r.dz = (w2.x * r.dx + w2.y * r.dy) + w2.z * r.dz;
    // The hardware reciprocal (v_rcp_f32, within 1 ulp; 0 -> inf like the
    // division): it scales an axis's plane distances by at most 1 + 2^-23,
    // moving the crossing points it tests by at most 2^-23 T |d|, 2^-7 of
    // the widening below, and the closest-hit pruning allows 2^-18 on top.
    o.r.ix = __builtin_amdgcn_rcpf(o.r.dx);
    o.r.iy = __builtin_amdgcn_rcpf(o.r.dy);
    o.r.iz = __builtin_amdgcn_rcpf(o.r.dz);
    const float ax = fabsf(r.ox) + T * fabsf(r.dx), ay = fabsf(r.oy) + T * fabsf(r.dy),
                az = fabsf(r.oz) + T * fabsf(r.dz);
    const float k = 1.0f / 65536.0f;
    o.pad[0] = ((fabsf(w0.x) * ax + fabsf(w0.y) * ay) + (fabsf(w0.z) * az + fabsf(w0.w))) * k + w3.x;
    o.pad[1] = ((fabsf(w1.x) * ax + fabsf(w1.y) * ay) + (fabsf(w1.z) * az + fabsf(w1.w))) * k + w3.x;
    o.pad[2] = ((fabsf(w2.x) * ax + fabsf(w2.y) * ay) + (fabsf(w2.z) * az + fabsf(w2.w))) * k + w3.x;
}

// The widening's parameter bound for an instance entered through its TLAS
// leaf: the exit from the leaf's world box, at most tmax (an unbounded or
// undefined exit falls back to tmax).
__device__ __forceinline__ float exit_bound(const RayF& r, float4 lo, float4 hi)
{
    return fminf(slab_exit(r, lo, hi) * 1.001f, r.tmax);
}

// getTriangle (shaders/raytracing.h:134-153) on a mesh-tree leaf record: the
// world triangle {v0, e0, e1} from the object positions through objectToWorld.
__device__ __forceinline__ void world_triangle(const float* M, const float4* p, float4& a, float4& b, float4& c)
{
    const float4 p0 = p[0], p1 = p[1], p2 = p[2];
    const F3 v0 = xform4(M, mk3(p0.x, p0.y, p0.z), 1.0f);
    const F3 e0 = xform4(M, mk3(p1.x, p1.y, p1.z), 1.0f) - v0;
    const F3 e1 = xform4(M, mk3(p2.x, p2.y, p2.z), 1.0f) - v0;
    a = make_float4(v0.x, v0.y, v0.z, p0.w); // .w: primId
    b = make_float4(e0.x, e0.y, e0.z, p1.w); // .w: mesh-tree leaf
    c = make_float4(e1.x, e1.y, e1.z, 0.0f);
}

// Order of the leaves of the expanded tree: the instance's TLAS leaf, then the
// triangle's leaf in its mesh tree.
__device__ __forceinline__ bool key_less(uint32_t ikA, uint32_t leafA, uint32_t ikB, uint32_t leafB)
{
    return ikA < ikB || (ikA == ikB && leafA < leafB);
}

// The two-level walks run as one loop of wide-node visits: the current ray is
// the world ray in the instance tree and the instance's object ray (with its
// box widening) in a mesh tree, so every iteration of a wave does the same
// kind of work.  Stack entries are node groups tagged by kind: 0 instance-tree
// nodes, 1 mesh-tree nodes of the current instance, 2 entered instances of an
// instance-tree node (slot order).  A node's entered instances are taken
// before its inner children; popping an instance sets up its object ray,
// popping an instance-tree group from inside a mesh tree returns to the world
// ray.  (Entering the nearest instance at once, as inner children are, costs
// more in registers and divergence than its ordering saves.)  Modes: closest hit (t, then the expanded preorder: instance leaf, then
// mesh-tree leaf), visibility (occluded by a non-emitter triangle ordered
// before the first hit emitter triangle, lightIk / lightLeaf) and any hit
// (the hit triangle first in expanded preorder).
enum { kTlClosest = 0, kTlVisible = 1, kTlAny = 2 };

template <int MODE>
__device__ __forceinline__ bool two_level_walk(const BvhView& bvh, const RayF& r, uint32_t* sA, uint32_t* sB,
                                               int stride, HitF& h, uint32_t& hIk, uint32_t lightIk,
                                               uint32_t lightLeaf)
{
    miss(h);
    hIk = STK_INVALID_INDEX;
    bool found = false;
    const uint32_t octW = ray_octant(r);
    RayF c = r;
    float pad[3] = {0.0f, 0.0f, 0.0f};
    uint32_t oct = octW, inst = 0, ik = 0, triOff = 0;
    bool inMesh = false;
    int sp = 0;
    uint32_t node = 0;
    // enters the instance of instance-tree leaf record li (false: skipped)
    auto enter = [&](uint32_t li) -> bool {
        const float4* L = bvh.tlasLeaves + 3 * (size_t)li;
        const float4 L0 = L[0], L1 = L[1], L2 = L[2];
        const uint32_t ni = float_to_bits(L0.w), nk = float_to_bits(L2.x);
        // visibility: an emitter's hit triangles, and every triangle of an
        // instance whose leaf follows the light's, come after the light in
        // preorder; any hit: nothing after the best
        if (MODE == kTlVisible && (nk > lightIk || bvh.instances[ni].lightId != -1)) return false;
        if (MODE == kTlAny && found && nk > hIk) return false;
        inst = ni;
        ik = nk;
        if (bvh.worldTris) triOff = bvh.instTriOff[ni];
        InstRay o;
        object_ray(bvh.instW + 4 * (size_t)ni, r, exit_bound(r, L0, L1), o);
        c = o.r;
        pad[0] = o.pad[0];
        pad[1] = o.pad[1];
        pad[2] = o.pad[2];
        oct = ray_octant(c);
        inMesh = true;
        node = float_to_bits(L1.w);
        return true;
    };
    for (;;) {
        WideVisit v;
        const float lim = ((MODE == kTlClosest && found) ? h.t : 1e9f) * kPruneSlack;
        wide_visit_t<true>((inMesh ? bvh.blas : bvh.tlas) + 5 * (size_t)node, c, lim, pad, v);
        uint32_t inner = v.innerHits;
        if (inMesh) {
            const float* M = bvh.instances[inst].objectToWorld;
            for (uint32_t lm = v.leafHits; lm; lm &= lm - 1u) {
                const uint32_t ls = (uint32_t)(__ffs(lm) - 1);
                float4 a, b, cc;
                const uint32_t gt = v.triBase + rank_below(v.leafMask, ls);
                if (bvh.worldTris) {
                    const float4* q = bvh.worldTris + 3 * (size_t)(gt + triOff);
                    a = q[0];
                    b = q[1];
                    cc = q[2];
                } else {
                    world_triangle(M, bvh.blasTris + 3 * (size_t)gt, a, b, cc);
                }
                const uint32_t leaf = float_to_bits(b.w);
                if (MODE == kTlVisible && ik == lightIk && leaf >= lightLeaf) continue;
                if (MODE == kTlAny && found && !key_less(ik, leaf, hIk, h.leaf)) continue;
                float t, u, w;
                if (tri_test(r, a, b, cc, t, u, w) && t < r.tmax) {
                    if (MODE == kTlVisible) return false;
                    const bool acc = MODE == kTlAny
                                         ? true
                                         : (found ? (t < h.t || (t == h.t && key_less(ik, leaf, hIk, h.leaf)))
                                                  : (t < 1e9f));
                    if (acc) {
                        found = true;
                        h.t = t;
                        h.u = u;
                        h.v = w;
                        h.leaf = leaf;
                        h.inst = inst;
                        h.prim = float_to_bits(a.w);
                        hIk = ik;
                    }
                }
            }
        } else if (v.leafHits) {
            if (inner) {
                sA[sp * stride] = v.childBase;
                sB[sp * stride] = v.imask | (xor_permute(inner, oct) << 8);
                ++sp;
                inner = 0;
            }
            sA[sp * stride] = v.triBase;
            sB[sp * stride] = v.leafMask | (v.leafHits << 8) | (2u << 16);
            ++sp;
        }
        if (inner) {
            int sn = 0;
            float tn = INFINITY;
#pragma unroll
            for (int s = 0; s < 8; ++s) // static indices keep t[] in registers
                if (((inner >> s) & 1u) && v.t[s] < tn) {
                    tn = v.t[s];
                    sn = s;
                }
            const uint32_t rest = inner & ~(1u << sn);
            if (rest) {
                sA[sp * stride] = v.childBase;
                sB[sp * stride] = v.imask | (xor_permute(rest, oct) << 8) | ((inMesh ? 1u : 0u) << 16);
                ++sp;
            }
            node = v.childBase + rank_below(v.imask, (uint32_t)sn);
            continue;
        }
        for (;;) {
            if (sp == 0) return MODE == kTlVisible ? true : found;
            const uint32_t first = sA[(sp - 1) * stride], word = sB[(sp - 1) * stride];
            const uint32_t mask = word & 0xFFu, kind = word >> 16;
            uint32_t m = (word >> 8) & 0xFFu;
            const uint32_t k = (uint32_t)(__ffs(m) - 1);
            m &= m - 1u;
            if (m)
                sB[(sp - 1) * stride] = (word & 0xFFFF00FFu) | (m << 8);
            else
                --sp;
            if (kind == 2) {
                if (enter(first + rank_below(mask, k))) break;
                continue;
            }
            if (kind == 0 && inMesh) {
                c = r;
                pad[0] = pad[1] = pad[2] = 0.0f;
                oct = octW;
                inMesh = false;
            }
            node = first + rank_below(mask, k ^ oct);
            break;
        }
    }
}

// Closest hit over the two-level structure; h.leaf is the mesh-tree leaf and
// hIk the instance's leaf in the instance tree.
__device__ __forceinline__ bool closest_two_level(const BvhView& bvh, const RayF& r, uint32_t* sA, uint32_t* sB,
                                                  int stride, HitF& h, uint32_t& hIk)
{
    return two_level_walk<kTlClosest>(bvh, r, sA, sB, stride, h, hIk, 0u, 0u);
}

// Shadow-ray visibility with the reference's any-hit rule on the expanded
// tree (as visible_wide): the emitter triangles first (the first hit one in
// expanded preorder), then the walk in any order.
__device__ __forceinline__ bool visible_two_level(const BvhView& bvh, const float4* emit, uint32_t emitCount,
                                                  const RayF& r, uint32_t* sA, uint32_t* sB, int stride)
{
    uint32_t lightIk = STK_INVALID_INDEX, lightLeaf = STK_INVALID_INDEX;
    for (uint32_t e = 0; e < emitCount; ++e) {
        const float4* p = emit + 3 * (size_t)e;
        const float4 a = p[0], b = p[1], c = p[2];
        float t, u, w;
        if (tri_test(r, a, b, c, t, u, w) && t < r.tmax) {
            const uint32_t ik = bvh.instKey[float_to_bits(b.w)], leaf = float_to_bits(a.w);
            if (key_less(ik, leaf, lightIk, lightLeaf)) {
                lightIk = ik;
                lightLeaf = leaf;
            }
        }
    }
    HitF h;
    uint32_t hIk;
    return two_level_walk<kTlVisible>(bvh, r, sA, sB, stride, h, hIk, lightIk, lightLeaf);
}

// anyHit (shaders/raytracing.h:249-286) over the expanded tree: the first leaf
// in expanded preorder whose triangle is hit with t < tmax, found by walking
// every entered box (no order of the wide walk follows the preorder).
__device__ __forceinline__ bool any_two_level(const BvhView& bvh, const RayF& r, uint32_t* sA, uint32_t* sB,
                                              int stride, HitF& h)
{
    uint32_t hIk;
    return two_level_walk<kTlAny>(bvh, r, sA, sB, stride, h, hIk, 0u, 0u);
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
