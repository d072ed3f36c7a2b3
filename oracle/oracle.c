/*
 * oracle.c - CPU oracle for the Strelka path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU
 * baseline of the benchmark as the checker; never linked into the product.
 *
 * Each function cites the reference file:line it restates.  Build:
 *   gcc -O2 -ffp-contract=off -fno-fast-math -fPIC -shared -pthread
 * (see oracle/build.py).
 */
#include "oracle.h"
#include "../strelka_amd/csrc/fmath.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* float3 helpers; evaluation order is fixed and contraction is off.          */
/* ------------------------------------------------------------------------ */
typedef struct { float x, y, z; } f3;

static inline f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 ld(const float* p) { return mk(p[0], p[1], p[2]); }
static inline void st3(float* p, f3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
static inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 mul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline f3 scl(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline f3 neg(f3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross(f3 a, f3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float len3(f3 a) { return sqrtf(dot(a, a)); }
/* normalize(v) = v * (1 / |v|) */
static inline f3 nrm(f3 a)
{
    float inv = 1.0f / sqrtf(dot(a, a));
    return scl(a, inv);
}
static inline float saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
static inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline float as_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t as_uint(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline int32_t as_int(float f) { int32_t u; memcpy(&u, &f, 4); return u; }
static inline float int_as_float(int32_t u) { float f; memcpy(&f, &u, 4); return f; }

#define ORC_PI 3.1415926535897f
#define ORC_INV_PI (1.0f / ORC_PI)

/* mul(M, float4(p, w)).xyz with glm's pairwise summation
 * (c0*x + c1*y) + (c2*z + c3*w); the reference builds its BVH input with
 * glm `m * float4(v, 1)` (src/render/ptrender.cpp:488-490) and the shader
 * applies the same objectToWorld (shaders/raytracing.h:143-145). */
static inline f3 xform4(const float m[16], f3 p, float w)
{
    f3 r;
    r.x = (m[0] * p.x + m[4] * p.y) + (m[8] * p.z + m[12] * w);
    r.y = (m[1] * p.x + m[5] * p.y) + (m[9] * p.z + m[13] * w);
    r.z = (m[2] * p.x + m[6] * p.y) + (m[10] * p.z + m[14] * w);
    return r;
}
/* mul((float3x3)M, v) = (c0*x + c1*y) + c2*z */
static inline f3 xform3(const float m[16], f3 v)
{
    f3 r;
    r.x = (m[0] * v.x + m[4] * v.y) + m[8] * v.z;
    r.y = (m[1] * v.x + m[5] * v.y) + m[9] * v.z;
    r.z = (m[2] * v.x + m[6] * v.y) + m[10] * v.z;
    return r;
}

void orc_transform_point(const float m[16], const float p[3], float out[3]) { st3(out, xform4(m, ld(p), 1.0f)); }

/* ------------------------------------------------------------------------ */
/* Primitives                                                                 */
/* ------------------------------------------------------------------------ */

/* intersectRayBox (shaders/ray.h:20-33 C++ branch, :35-48 HLSL branch). */
int orc_intersect_ray_box(const StkRay* r, const float invdir[3], const float pmin[3], const float pmax[3], float* t)
{
    const float fx = (pmax[0] - r->o[0]) * invdir[0];
    const float fy = (pmax[1] - r->o[1]) * invdir[1];
    const float fz = (pmax[2] - r->o[2]) * invdir[2];
    const float nx = (pmin[0] - r->o[0]) * invdir[0];
    const float ny = (pmin[1] - r->o[1]) * invdir[1];
    const float nz = (pmin[2] - r->o[2]) * invdir[2];
    const float tmaxx = fmaxf(fx, nx), tmaxy = fmaxf(fy, ny), tmaxz = fmaxf(fz, nz);
    const float tminx = fminf(fx, nx), tminy = fminf(fy, ny), tminz = fminf(fz, nz);
    const float t1 = fminf(tmaxx, fminf(tmaxy, tmaxz));
    const float t0 = fmaxf(fmaxf(tminx, fmaxf(tminy, tminz)), 0.0f);
    *t = t0;
    return t1 >= t0;
}

/* RayTriangleIntersect (shaders/raytracing.h:36-88): Moller-Trumbore, no
 * culling, |det| < 1e-6 rejected, t preset to 1e9, t < 1e-6 rejected. */
int orc_ray_triangle(const float o[3], const float d[3], const float v0[3], const float e0[3], const float e1[3],
                     float* t, float bary[2])
{
    *t = 1e9f;
    const f3 dir = ld(d), E0 = ld(e0), E1 = ld(e1);
    const f3 pvec = cross(dir, E1);
    const float det = dot(E0, pvec);
    if (fabsf(det) < 1e-6f) return 0;
    const float invDet = 1.0f / det;
    const f3 tvec = sub(ld(o), ld(v0));
    const float u = dot(tvec, pvec) * invDet;
    if (u < 0.0f || u > 1.0f) return 0;
    const f3 qvec = cross(tvec, E0);
    const float v = dot(dir, qvec) * invDet;
    if (v < 0.0f || (u + v) > 1.0f) return 0;
    *t = dot(E1, qvec) * invDet;
    if (*t < 1e-6f) return 0;
    bary[0] = u;
    bary[1] = v;
    return 1;
}

/* getTriangle (shaders/raytracing.h:134-153). */
void orc_get_triangle(const OrcScene* s, uint32_t inst, uint32_t prim, float v0[3], float e0[3], float e1[3])
{
    const StkInstance* ic = &s->inst[inst];
    const uint32_t i0 = s->ib[(uint32_t)ic->indexOffset + prim * 3 + 0];
    const uint32_t i1 = s->ib[(uint32_t)ic->indexOffset + prim * 3 + 1];
    const uint32_t i2 = s->ib[(uint32_t)ic->indexOffset + prim * 3 + 2];
    const f3 a = xform4(ic->objectToWorld, ld(s->vb[i0].position), 1.0f);
    const f3 b = sub(xform4(ic->objectToWorld, ld(s->vb[i1].position), 1.0f), a);
    const f3 c = sub(xform4(ic->objectToWorld, ld(s->vb[i2].position), 1.0f), a);
    st3(v0, a);
    st3(e0, b);
    st3(e1, c);
}

static inline int is_leaf(const StkBvhNode* n) { return (uint32_t)n->instId != STK_INVALID_INDEX; }

static inline void ray_invdir(const StkRay* r, float inv[3])
{
    inv[0] = 1.0f / r->d[0];
    inv[1] = 1.0f / r->d[1];
    inv[2] = 1.0f / r->d[2];
}

static inline void miss_hit(StkHit* h)
{
    h->bary[0] = 0.0f;
    h->bary[1] = 0.0f;
    h->t = 1e9f;
    h->instId = STK_INVALID_INDEX;
    h->primId = STK_INVALID_INDEX;
}

/* ------------------------------------------------------------------------ */
/* Traversal                                                                  */
/* ------------------------------------------------------------------------ */

/* closestHit (shaders/raytracing.h:198-247): stackless walk over the
 * preorder/skip-link layout; leaf accepted if hit && t < tmax && t < minHit
 * (strict), minHit starts at 1e9; an inner node is entered iff the slab test
 * hits and boxT <= tmax (:232-240). */
int orc_closest_hit(const OrcScene* s, const StkRay* r, StkHit* hit, OrcStats* st)
{
    miss_hit(hit);
    if (s->nnodes == 0) return 0;
    float inv[3];
    ray_invdir(r, inv);
    float minHit = 1e9f;
    int found = 0;
    uint32_t idx = 0;
    uint64_t ninner = 0, nleaf = 0;
    while (idx != STK_INVALID_INDEX) {
        const StkBvhNode* n = &s->nodes[idx];
        if (is_leaf(n)) {
            const uint32_t prim = as_uint(n->minBounds[0]);
            float v0[3], e0[3], e1[3], t, b[2];
            orc_get_triangle(s, (uint32_t)n->instId, prim, v0, e0, e1);
            ++nleaf;
            if (orc_ray_triangle(r->o, r->d, v0, e0, e1, &t, b) && t < r->o[3] && t < minHit) {
                minHit = t;
                hit->t = t;
                hit->bary[0] = b[0];
                hit->bary[1] = b[1];
                hit->instId = (uint32_t)n->instId;
                hit->primId = prim;
                found = 1;
            }
            idx = (uint32_t)n->nodeOffset;
            continue;
        }
        float boxT;
        ++ninner;
        if (orc_intersect_ray_box(r, inv, n->minBounds, n->maxBounds, &boxT)) {
            if (boxT > r->o[3]) {
                idx = (uint32_t)n->nodeOffset;
                continue;
            }
            ++idx;
            continue;
        }
        idx = (uint32_t)n->nodeOffset;
    }
    if (st) {
        st->inner += ninner;
        st->leaf += nleaf;
        st->rays += 1;
        st->hits += (uint64_t)found;
    }
    return found;
}

/* Acceptance rule shared by the order-independent traversals: smallest t,
 * ties broken by the smaller preorder leaf index.  This equals the strict '<'
 * first-in-DFS rule of closestHit on the same BVH. */
static inline int accept(int found, float t, uint32_t leaf, float best, uint32_t bestLeaf)
{
    if (!found) return t < 1e9f;
    return t < best || (t == best && leaf < bestLeaf);
}

/* Box pruning threshold: a box is skipped when its entry t exceeds the current
 * best t by more than a relative 2^-18, which leaves a margin for the rounding
 * of the triangle t against the box t. */
#define ORC_PRUNE_SLACK (1.0f + 1.0f / 262144.0f)

static inline int leaf_test(const OrcScene* s, const StkRay* r, uint32_t idx, int* found, float* best,
                            uint32_t* bestLeaf, StkHit* hit)
{
    const StkBvhNode* n = &s->nodes[idx];
    const uint32_t prim = as_uint(n->minBounds[0]);
    float v0[3], e0[3], e1[3], t, b[2];
    orc_get_triangle(s, (uint32_t)n->instId, prim, v0, e0, e1);
    if (orc_ray_triangle(r->o, r->d, v0, e0, e1, &t, b) && t < r->o[3] && accept(*found, t, idx, *best, *bestLeaf)) {
        *found = 1;
        *best = t;
        *bestLeaf = idx;
        hit->t = t;
        hit->bary[0] = b[0];
        hit->bary[1] = b[1];
        hit->instId = (uint32_t)n->instId;
        hit->primId = prim;
        return 1;
    }
    return 0;
}

/* Near-first traversal with closest-t pruning over the same skip-link array:
 * the children of inner node i are i+1 and nodes[i+1].nodeOffset
 * (src/render/bvh.cpp:182-235 writes the right sibling's preorder index as the
 * left child's skip link).  This is the algorithm of the GPU kernel; its
 * counters give the bytes that kernel must move. */
#define ORC_STACK 64
int orc_closest_hit_pruned(const OrcScene* s, const StkRay* r, StkHit* hit, OrcStats* st)
{
    miss_hit(hit);
    if (s->nnodes == 0) return 0;
    float inv[3];
    ray_invdir(r, inv);
    int found = 0;
    float best = 1e9f;
    uint32_t bestLeaf = STK_INVALID_INDEX;
    uint64_t nfetch = 0, nleaf = 0;
    const float tmax = r->o[3];

    if (is_leaf(&s->nodes[0])) {
        ++nleaf;
        leaf_test(s, r, 0, &found, &best, &bestLeaf, hit);
    } else {
        float t0;
        uint32_t stack[ORC_STACK];
        int sp = 0;
        uint32_t cur = STK_INVALID_INDEX;
        if (orc_intersect_ray_box(r, inv, s->nodes[0].minBounds, s->nodes[0].maxBounds, &t0) && t0 <= tmax)
            cur = 0;
        while (cur != STK_INVALID_INDEX) {
            ++nfetch;
            const uint32_t c[2] = {cur + 1, (uint32_t)s->nodes[cur + 1].nodeOffset};
            int leafc[2];
            for (int i = 0; i < 2; ++i) {
                leafc[i] = is_leaf(&s->nodes[c[i]]);
                if (leafc[i]) {
                    ++nleaf;
                    leaf_test(s, r, c[i], &found, &best, &bestLeaf, hit);
                }
            }
            int ok[2] = {0, 0};
            float tc[2] = {0.0f, 0.0f};
            for (int i = 0; i < 2; ++i) {
                if (leafc[i]) continue;
                const StkBvhNode* n = &s->nodes[c[i]];
                ok[i] = orc_intersect_ray_box(r, inv, n->minBounds, n->maxBounds, &tc[i]) && tc[i] <= tmax &&
                        !(tc[i] > best * ORC_PRUNE_SLACK);
            }
            if (ok[0] && ok[1]) {
                const int nearIsOne = tc[1] < tc[0];
                stack[sp++] = c[nearIsOne ? 0 : 1];
                cur = c[nearIsOne ? 1 : 0];
            } else if (ok[0]) {
                cur = c[0];
            } else if (ok[1]) {
                cur = c[1];
            } else {
                cur = sp > 0 ? stack[--sp] : STK_INVALID_INDEX;
            }
            if (sp >= ORC_STACK) abort(); /* trees are built with depth <= 32 */
        }
    }
    if (st) {
        st->inner += nfetch;
        st->leaf += nleaf;
        st->rays += 1;
        st->hits += (uint64_t)found;
    }
    return found;
}

int orc_closest_hit_brute(const OrcScene* s, const StkRay* r, StkHit* hit)
{
    miss_hit(hit);
    int found = 0;
    float best = 1e9f;
    uint32_t bestLeaf = STK_INVALID_INDEX;
    for (uint32_t i = 0; i < s->nnodes; ++i)
        if (is_leaf(&s->nodes[i])) leaf_test(s, r, i, &found, &best, &bestLeaf, hit);
    return found;
}

/* The any-hit rule by brute force: the first leaf of the preorder array whose
 * triangle is hit with t < tmax.  The walk below visits leaves in preorder and
 * only skips subtrees whose box it misses or enters beyond tmax; a triangle
 * hit at t < tmax lies in every ancestor box, so both give the same leaf
 * except where a slab test rejects a box that holds a valid hit. */
int orc_any_hit_brute(const OrcScene* s, const StkRay* r, StkHit* hit)
{
    miss_hit(hit);
    for (uint32_t i = 0; i < s->nnodes; ++i) {
        const StkBvhNode* n = &s->nodes[i];
        if (!is_leaf(n)) continue;
        const uint32_t prim = as_uint(n->minBounds[0]);
        float v0[3], e0[3], e1[3], t, b[2];
        orc_get_triangle(s, (uint32_t)n->instId, prim, v0, e0, e1);
        if (orc_ray_triangle(r->o, r->d, v0, e0, e1, &t, b) && t < r->o[3]) {
            hit->t = t;
            hit->instId = (uint32_t)n->instId;
            hit->primId = prim;
            return 1;
        }
    }
    return 0;
}

/* anyHit (shaders/raytracing.h:249-286): first leaf in DFS order whose
 * triangle is hit with t < tmax. */
int orc_any_hit(const OrcScene* s, const StkRay* r, StkHit* hit, OrcStats* st)
{
    miss_hit(hit);
    if (s->nnodes == 0) return 0;
    float inv[3];
    ray_invdir(r, inv);
    uint32_t idx = 0;
    uint64_t ninner = 0, nleaf = 0;
    int found = 0;
    while (idx != STK_INVALID_INDEX) {
        const StkBvhNode* n = &s->nodes[idx];
        if (is_leaf(n)) {
            const uint32_t prim = as_uint(n->minBounds[0]);
            float v0[3], e0[3], e1[3], t, b[2];
            orc_get_triangle(s, (uint32_t)n->instId, prim, v0, e0, e1);
            ++nleaf;
            const int isect = orc_ray_triangle(r->o, r->d, v0, e0, e1, &t, b);
            hit->t = t;
            if (isect && t < r->o[3]) {
                hit->instId = (uint32_t)n->instId;
                hit->primId = prim;
                found = 1;
                break;
            }
        } else {
            float boxT;
            ++ninner;
            if (orc_intersect_ray_box(r, inv, n->minBounds, n->maxBounds, &boxT)) {
                if (boxT > r->o[3]) {
                    idx = (uint32_t)n->nodeOffset;
                    continue;
                }
                ++idx;
                continue;
            }
        }
        idx = (uint32_t)n->nodeOffset;
    }
    if (!found) hit->t = 1e9f;
    if (st) {
        st->inner += ninner;
        st->leaf += nleaf;
        st->rays += 1;
        st->hits += (uint64_t)found;
    }
    return found;
}

/* The shadow test of the corrected estimator: whether a triangle of a
 * non-emitter instance is hit with t < tmax (the anyHit walk of
 * shaders/raytracing.h:249-286 continued past emitter triangles). */
int orc_occluded(const OrcScene* s, const StkRay* r)
{
    if (s->nnodes == 0) return 0;
    float inv[3];
    ray_invdir(r, inv);
    uint32_t idx = 0;
    while (idx != STK_INVALID_INDEX) {
        const StkBvhNode* n = &s->nodes[idx];
        if (is_leaf(n)) {
            const uint32_t prim = as_uint(n->minBounds[0]);
            float v0[3], e0[3], e1[3], t, b[2];
            orc_get_triangle(s, (uint32_t)n->instId, prim, v0, e0, e1);
            if (orc_ray_triangle(r->o, r->d, v0, e0, e1, &t, b) && t < r->o[3] && s->inst[n->instId].lightId == -1)
                return 1;
        } else {
            float boxT;
            if (orc_intersect_ray_box(r, inv, n->minBounds, n->maxBounds, &boxT)) {
                idx = boxT > r->o[3] ? (uint32_t)n->nodeOffset : idx + 1;
                continue;
            }
        }
        idx = (uint32_t)n->nodeOffset;
    }
    return 0;
}

/* ---- threaded batch ------------------------------------------------------- */
typedef struct {
    const OrcScene* s;
    const StkRay* rays;
    StkHit* hits;
    uint32_t n;
    int mode;
    volatile uint32_t* next;
    pthread_mutex_t* mu;
    OrcStats stats;
} BatchArg;

#define ORC_CHUNK 4096u

/* ------------------------------------------------------------------------ */
/* The compressed 8-wide walk of the GPU kernel (its design, not the          */
/* reference's); hits must equal orc_closest_hit's.                           */
/* ------------------------------------------------------------------------ */
#define ORC_WIDE_STACK 64

static inline float wide_scale(uint8_t e)
{
    const uint32_t bits = (uint32_t)e << 23;
    float f;
    memcpy(&f, &bits, sizeof(f));
    return f;
}

static inline uint32_t xor_permute8(uint32_t m, uint32_t oct)
{
    if (oct & 1u) m = ((m & 0x55u) << 1) | ((m >> 1) & 0x55u);
    if (oct & 2u) m = ((m & 0x33u) << 2) | ((m >> 2) & 0x33u);
    if (oct & 4u) m = ((m & 0x0Fu) << 4) | ((m >> 4) & 0x0Fu);
    return m;
}

static inline uint32_t popc8(uint32_t m) { return (uint32_t)__builtin_popcount(m & 0xFFu); }

/* The kernel's slot tests (trace_dev.hpp wide_axis / wide_slot): per axis
 * A = scale x inverse direction and the near / far planes' offsets
 * B = (plane origin - ray origin) x inverse direction moved out by
 * e = (765 scale + 4 |origin| + 3 |ray origin|) |inv| 2^-23, twice a bound on
 * the rounding of t = q A + B and of the builder's decode, so no box the
 * exact slab test of the decoded box accepts is rejected (the inverse
 * direction clamped to +-2^80, so a zero component gives distances beyond
 * any hit instead of undefined ones); per slot the near and far distances
 * are one fused multiply-add each, reduced with 0 and min(tmax, lim) as
 * bounds.  pad (may be NULL): the plane origins moved out
 * by pad[a] (a mesh tree's widened boxes).  tch[s] gets slot s's entry
 * distance; returns the entered mask. */
#define ORC_INV_CLAMP 1208925819614629174706176.0f /* 2^80 */
static uint32_t wide_slots(const StkWideNode* w, const StkRay* r, const float inv[3], float lim, const float pad[3],
                           float tch[8])
{
    float A[3], BN[3], BF[3];
    int neg[3];
    for (int a = 0; a < 3; ++a) {
        const float scale = wide_scale(w->exps[a]), origin = w->origin[a], ro = r->o[a];
        const float ia = fminf(fmaxf(inv[a], -ORC_INV_CLAMP), ORC_INV_CLAMP);
        neg[a] = signbit(ia) != 0;
        A[a] = scale * ia;
        const float mag = pad ? fabsf(origin) + pad[a] : fabsf(origin);
        const float e = fmaf(765.0f, scale, fmaf(4.0f, mag, 3.0f * fabsf(ro))) * (fabsf(ia) * (1.0f / 8388608.0f));
        if (pad) {
            BN[a] = ((neg[a] ? origin + pad[a] : origin - pad[a]) - ro) * ia - e;
            BF[a] = ((neg[a] ? origin - pad[a] : origin + pad[a]) - ro) * ia + e;
        } else {
            const float b = (origin - ro) * ia;
            BN[a] = b - e;
            BF[a] = b + e;
        }
    }
    const float L = fminf(r->o[3], lim);
    uint32_t hits = 0;
    for (int sl = 0; sl < 8; ++sl) {
        float tn[3], tf[3];
        for (int a = 0; a < 3; ++a) {
            const float qn = (float)(neg[a] ? w->qhi[a][sl] : w->qlo[a][sl]);
            const float qf = (float)(neg[a] ? w->qlo[a][sl] : w->qhi[a][sl]);
            tn[a] = fmaf(qn, A[a], BN[a]);
            tf[a] = fmaf(qf, A[a], BF[a]);
        }
        const float t0 = fmaxf(fmaxf(tn[0], tn[1]), fmaxf(tn[2], 0.0f));
        const float t1 = fminf(fminf(tf[0], tf[1]), fminf(tf[2], L));
        tch[sl] = t0;
        if (t0 <= t1) hits |= 1u << sl;
    }
    return hits & w->vmask & 0xFFu;
}

int orc_closest_hit_wide(const OrcScene* s, const StkRay* r, StkHit* hit, OrcStats* st)
{
    miss_hit(hit);
    if (s->nwide == 0) return 0;
    float inv[3];
    ray_invdir(r, inv);
    const uint32_t oct = (signbit(r->d[0]) ? 1u : 0u) | (signbit(r->d[1]) ? 2u : 0u) | (signbit(r->d[2]) ? 4u : 0u);
    const float tmax = r->o[3];
    int found = 0;
    float best = 1e9f;
    uint32_t bestLeaf = STK_INVALID_INDEX;
    uint64_t nvisit = 0, ntri = 0;
    uint32_t stackBase[ORC_WIDE_STACK], stackWord[ORC_WIDE_STACK];
    int sp = 0;
    uint32_t node = 0;
    for (;;) {
        const StkWideNode* w = &s->wide[node];
        ++nvisit;
        const float lim = (found ? best : 1e9f) * ORC_PRUNE_SLACK;
        float tch[8];
        const uint32_t hits = wide_slots(w, r, inv, lim, NULL, tch);
        const uint32_t vmask = w->vmask & 0xFFu, imask = w->imask, leafMask = vmask & ~imask;
        for (int sl = 0; sl < 8; ++sl) {
            if (!((hits & leafMask) >> sl & 1u)) continue;
            const StkTriangle* tr = &s->wideTris[w->triBase + popc8(leafMask & ((1u << sl) - 1u))];
            float v0[3], e0[3], e1[3], t, b[2];
            orc_get_triangle(s, tr->instId, tr->primId, v0, e0, e1);
            ++ntri;
            if (orc_ray_triangle(r->o, r->d, v0, e0, e1, &t, b) && t < tmax &&
                accept(found, t, tr->leaf, best, bestLeaf)) {
                found = 1;
                best = t;
                bestLeaf = tr->leaf;
                hit->t = t;
                hit->bary[0] = b[0];
                hit->bary[1] = b[1];
                hit->instId = tr->instId;
                hit->primId = tr->primId;
            }
        }
        uint32_t inner = hits & imask;
        if (inner) {
            int sn = 0;
            float tn = INFINITY;
            for (int sl = 0; sl < 8; ++sl)
                if ((inner >> sl & 1u) && tch[sl] < tn) {
                    tn = tch[sl];
                    sn = sl;
                }
            inner &= ~(1u << sn);
            if (inner) {
                if (sp >= ORC_WIDE_STACK) abort();
                stackBase[sp] = w->childBase;
                stackWord[sp] = imask | (xor_permute8(inner, oct) << 8);
                ++sp;
            }
            node = w->childBase + popc8(imask & ((1u << sn) - 1u));
            continue;
        }
        if (sp == 0) break;
        const uint32_t base = stackBase[sp - 1], word = stackWord[sp - 1];
        const uint32_t im = word & 0xFFu;
        uint32_t m = word >> 8;
        const int k = __builtin_ctz(m);
        m &= m - 1u;
        if (m)
            stackWord[sp - 1] = im | (m << 8);
        else
            --sp;
        node = base + popc8(im & ((1u << ((uint32_t)k ^ oct)) - 1u));
    }
    if (st) {
        st->inner += nvisit;
        st->leaf += ntri;
        st->rays += 1;
        st->hits += (uint64_t)found;
    }
    return found;
}

/* ---- two-level structure ------------------------------------------------------ */

static void world_box(const float m[16], const float lo[3], const float hi[3], float olo[3], float ohi[3])
{
    for (int r = 0; r < 3; ++r) {
        double c = m[12 + r], e = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double ck = 0.5 * ((double)lo[k] + hi[k]), ek = 0.5 * ((double)hi[k] - lo[k]);
            c += (double)m[4 * k + r] * ck;
            e += fabs((double)m[4 * k + r]) * ek;
        }
        const double l = c - e, h = c + e;
        const double mag = fabs(l) > fabs(h) ? fabs(l) : fabs(h);
        const double pad = 1e-5 * ((h - l) > mag ? (h - l) : mag) + 1e-30;
        olo[r] = nextafterf((float)(l - pad), -INFINITY);
        ohi[r] = nextafterf((float)(h + pad), INFINITY);
    }
}

typedef struct ExpandCtx {
    const StkBvhNode* top;
    const StkBvhNode* meshNodes;
    const uint32_t* meshStart;
    const uint32_t* meshCount;
    const uint32_t* instMesh;
    const StkInstance* inst;
    StkBvhNode* out;
    uint32_t capacity;
    int overflow;
} ExpandCtx;

static uint32_t expanded_size(const ExpandCtx* c, uint32_t ti)
{
    const StkBvhNode* nd = &c->top[ti];
    if (is_leaf(nd)) return c->meshCount[c->instMesh[(uint32_t)nd->instId]];
    const uint32_t left = ti + 1, right = (uint32_t)c->top[left].nodeOffset;
    return 1 + expanded_size(c, left) + expanded_size(c, right);
}

static void expand_emit(ExpandCtx* c, uint32_t ti, uint32_t at, uint32_t skip)
{
    const StkBvhNode* nd = &c->top[ti];
    if (is_leaf(nd)) {
        const uint32_t i = (uint32_t)nd->instId, m = c->instMesh[i];
        const StkBvhNode* mn = c->meshNodes + c->meshStart[m];
        const uint32_t cnt = c->meshCount[m];
        if ((uint64_t)at + cnt > c->capacity) {
            c->overflow = 1;
            return;
        }
        for (uint32_t k = 0; k < cnt; ++k) {
            StkBvhNode o = mn[k];
            if (is_leaf(&mn[k])) {
                o.instId = (int32_t)i;
            } else {
                world_box(c->inst[i].objectToWorld, mn[k].minBounds, mn[k].maxBounds, o.minBounds, o.maxBounds);
            }
            o.nodeOffset = ((uint32_t)mn[k].nodeOffset == STK_INVALID_INDEX) ? (int32_t)skip
                                                                            : (int32_t)(at + (uint32_t)mn[k].nodeOffset);
            c->out[at + k] = o;
        }
        return;
    }
    if (at >= c->capacity) {
        c->overflow = 1;
        return;
    }
    c->out[at] = *nd;
    c->out[at].nodeOffset = (int32_t)skip;
    const uint32_t left = ti + 1, right = (uint32_t)c->top[left].nodeOffset;
    const uint32_t sl = expanded_size(c, left);
    expand_emit(c, left, at + 1, at + 1 + sl);
    expand_emit(c, right, at + 1 + sl, skip);
}

uint32_t orc_expand_two_level(const StkBvhNode* top, uint32_t ntop, const StkBvhNode* meshNodes,
                              const uint32_t* meshStart, const uint32_t* meshCount, const uint32_t* instMesh,
                              const StkInstance* inst, StkBvhNode* out, uint32_t capacity)
{
    if (ntop == 0) return 0;
    ExpandCtx c = {top, meshNodes, meshStart, meshCount, instMesh, inst, out, capacity, 0};
    const uint32_t n = expanded_size(&c, 0);
    if (n > capacity) return 0;
    expand_emit(&c, 0, 0, STK_INVALID_INDEX);
    return c.overflow ? 0 : n;
}

/* The kernel's object-space ray and box widening (trace_dev.hpp object_ray).
 * The inverse direction is the IEEE quotient here; the kernel takes the
 * hardware reciprocal (within 1 ulp), which the widening covers: it can only
 * change which mesh-tree boxes a walk visits, never the hits. */
static void object_ray(const float* W, const StkRay* r, float T, StkRay* o, float inv[3], float pad[3])
{
    const float* w0 = W;
    const float* w1 = W + 4;
    const float* w2 = W + 8;
    o->o[0] = (w0[0] * r->o[0] + w0[1] * r->o[1]) + (w0[2] * r->o[2] + w0[3]);
    o->o[1] = (w1[0] * r->o[0] + w1[1] * r->o[1]) + (w1[2] * r->o[2] + w1[3]);
    o->o[2] = (w2[0] * r->o[0] + w2[1] * r->o[1]) + (w2[2] * r->o[2] + w2[3]);
    o->o[3] = r->o[3];
    o->d[0] = (w0[0] * r->d[0] + w0[1] * r->d[1]) + w0[2] * r->d[2];
    o->d[1] = (w1[0] * r->d[0] + w1[1] * r->d[1]) + w1[2] * r->d[2];
    o->d[2] = (w2[0] * r->d[0] + w2[1] * r->d[1]) + w2[2] * r->d[2];
    o->d[3] = 0.0f;
    for (int a = 0; a < 3; ++a) inv[a] = 1.0f / o->d[a];
    const float ax = fabsf(r->o[0]) + T * fabsf(r->d[0]), ay = fabsf(r->o[1]) + T * fabsf(r->d[1]),
                az = fabsf(r->o[2]) + T * fabsf(r->d[2]);
    const float k = 1.0f / 65536.0f;
    for (int a = 0; a < 3; ++a) {
        const float* w = W + 4 * a;
        pad[a] = ((fabsf(w[0]) * ax + fabsf(w[1]) * ay) + (fabsf(w[2]) * az + fabsf(w[3]))) * k + W[12];
    }
}

/* Exit distance from [lo, hi] (intersectRayBox's t1), bounded by tmax. */
static float exit_bound(const StkRay* r, const float inv[3], const float lo[3], const float hi[3])
{
    float t1 = INFINITY;
    int first = 1;
    for (int a = 0; a < 3; ++a) {
        const float f = (hi[a] - r->o[a]) * inv[a], n = (lo[a] - r->o[a]) * inv[a];
        const float m = fmaxf(f, n);
        t1 = first ? m : fminf(t1, m);
        first = 0;
    }
    return fminf(t1 * 1.001f, r->o[3]);
}

static void blas_world_triangle(const float M[16], const StkBlasTriangle* t, float v0[3], float e0[3], float e1[3])
{
    const f3 a = xform4(M, ld(t->p0), 1.0f);
    st3(v0, a);
    st3(e0, sub(xform4(M, ld(t->p1), 1.0f), a));
    st3(e1, sub(xform4(M, ld(t->p2), 1.0f), a));
}

static inline int key_less(uint32_t ikA, uint32_t leafA, uint32_t ikB, uint32_t leafB)
{
    return ikA < ikB || (ikA == ikB && leafA < leafB);
}

static uint32_t octant_of(const float d[3])
{
    return (signbit(d[0]) ? 1u : 0u) | (signbit(d[1]) ? 2u : 0u) | (signbit(d[2]) ? 4u : 0u);
}

/* Sets up the walk of instance-tree leaf record li's mesh tree (0: skipped). */
static int enter_instance(const OrcScene* s, const StkRay* r, const float inv[3], uint32_t li, int anyMode, int found,
                          uint32_t bestIk, uint32_t* inst, uint32_t* ik, StkRay* c, float cinv[3], float pad[3],
                          uint32_t* oct, uint32_t* node)
{
    const StkTlasLeaf* L = &s->tlasLeaves[li];
    if (anyMode && found && L->key > bestIk) return 0;
    *inst = L->instId;
    *ik = L->key;
    object_ray(s->instW + 16 * (size_t)L->instId, r, exit_bound(r, inv, L->lo, L->hi), c, cinv, pad);
    *oct = octant_of(c->d);
    *node = L->blasRoot;
    return 1;
}

/* The kernel's single-loop two-level walk (trace_dev.hpp two_level_walk):
 * stack entries are node groups of kind 0 (instance tree), 1 (mesh tree of
 * the current instance) or 2 (entered instances of an instance-tree node, slot
 * order); a node's instances are taken before its inner children.
 * anyMode: first hit in expanded preorder, no t pruning. */
static int two_level_walk(const OrcScene* s, const StkRay* r, StkHit* hit, OrcStats* st, int anyMode)
{
    miss_hit(hit);
    if (!s->tlas || s->ntlas == 0) return 0;
    float inv[3];
    ray_invdir(r, inv);
    const uint32_t octW = octant_of(r->d);
    int found = 0;
    float best = 1e9f;
    uint32_t bestIk = STK_INVALID_INDEX, bestLeaf = STK_INVALID_INDEX;
    uint64_t nvisit = 0, ntri = 0;
    uint32_t stFirst[ORC_WIDE_STACK], stWord[ORC_WIDE_STACK];
    int sp = 0;
    StkRay c = *r;
    float cinv[3] = {inv[0], inv[1], inv[2]}, pad[3] = {0.0f, 0.0f, 0.0f};
    uint32_t oct = octW, inst = 0, ik = 0, node = 0;
    int inMesh = 0;
    for (;;) {
        const StkWideNode* w = inMesh ? &s->blas[node] : &s->tlas[node];
        ++nvisit;
        float tch[8];
        const float lim = ((!anyMode && found) ? best : 1e9f) * ORC_PRUNE_SLACK;
        const uint32_t hits = wide_slots(w, &c, cinv, lim, pad, tch);
        const uint32_t imask = w->imask, leafMask = (w->vmask & 0xFFu) & ~imask;
        uint32_t inner = hits & imask;
        const uint32_t leafHits = hits & leafMask;
        if (inMesh) {
            const float* M = s->inst[inst].objectToWorld;
            for (uint32_t m = leafHits; m; m &= m - 1u) {
                const int ls = __builtin_ctz(m);
                const StkBlasTriangle* tr = &s->blasTris[w->triBase + popc8(leafMask & ((1u << ls) - 1u))];
                if (anyMode && found && !key_less(ik, tr->leaf, bestIk, bestLeaf)) continue;
                float v0[3], e0[3], e1[3], t, b[2];
                blas_world_triangle(M, tr, v0, e0, e1);
                ++ntri;
                if (!orc_ray_triangle(r->o, r->d, v0, e0, e1, &t, b) || !(t < r->o[3])) continue;
                const int acc = anyMode ? 1
                                        : (found ? (t < best || (t == best && key_less(ik, tr->leaf, bestIk, bestLeaf)))
                                                 : (t < 1e9f));
                if (acc) {
                    found = 1;
                    best = t;
                    bestIk = ik;
                    bestLeaf = tr->leaf;
                    hit->t = t;
                    hit->bary[0] = b[0];
                    hit->bary[1] = b[1];
                    hit->instId = inst;
                    hit->primId = tr->primId;
                }
            }
        } else if (leafHits) {
            if (sp + 2 > ORC_WIDE_STACK) abort();
            if (inner) {
                stFirst[sp] = w->childBase;
                stWord[sp] = imask | (xor_permute8(inner, oct) << 8);
                ++sp;
                inner = 0;
            }
            stFirst[sp] = w->triBase;
            stWord[sp] = leafMask | (leafHits << 8) | (2u << 16);
            ++sp;
        }
        if (inner) {
            int sn = 0;
            float tn = INFINITY;
            for (int sl = 0; sl < 8; ++sl)
                if ((inner >> sl & 1u) && tch[sl] < tn) {
                    tn = tch[sl];
                    sn = sl;
                }
            const uint32_t rest = inner & ~(1u << sn);
            if (rest) {
                if (sp >= ORC_WIDE_STACK) abort();
                stFirst[sp] = w->childBase;
                stWord[sp] = imask | (xor_permute8(rest, oct) << 8) | ((inMesh ? 1u : 0u) << 16);
                ++sp;
            }
            node = w->childBase + popc8(imask & ((1u << sn) - 1u));
            continue;
        }
        int go = 0;
        while (sp > 0) {
            const uint32_t first = stFirst[sp - 1], word = stWord[sp - 1];
            const uint32_t mask = word & 0xFFu, kind = word >> 16;
            uint32_t m = (word >> 8) & 0xFFu;
            const uint32_t k = (uint32_t)__builtin_ctz(m);
            m &= m - 1u;
            if (m)
                stWord[sp - 1] = (word & 0xFFFF00FFu) | (m << 8);
            else
                --sp;
            if (kind == 2) {
                if (!enter_instance(s, r, inv, first + popc8(mask & ((1u << k) - 1u)), anyMode, found, bestIk, &inst,
                                    &ik, &c, cinv, pad, &oct, &node))
                    continue;
                inMesh = 1;
                go = 1;
                break;
            }
            if (kind == 0 && inMesh) {
                c = *r;
                for (int a = 0; a < 3; ++a) {
                    cinv[a] = inv[a];
                    pad[a] = 0.0f;
                }
                oct = octW;
                inMesh = 0;
            }
            node = first + popc8(mask & ((1u << (k ^ oct)) - 1u));
            go = 1;
            break;
        }
        if (!go) break;
    }
    if (st) {
        st->inner += nvisit;
        st->leaf += ntri;
        st->rays += 1;
        st->hits += (uint64_t)found;
    }
    return found;
}

int orc_closest_two_level(const OrcScene* s, const StkRay* r, StkHit* hit, OrcStats* st)
{
    return two_level_walk(s, r, hit, st, 0);
}

int orc_any_two_level(const OrcScene* s, const StkRay* r, StkHit* hit, OrcStats* st)
{
    return two_level_walk(s, r, hit, st, 1);
}

static void* batch_worker(void* p)
{
    BatchArg* a = (BatchArg*)p;
    for (;;) {
        uint32_t start;
        pthread_mutex_lock(a->mu);
        start = *a->next;
        *a->next = start + ORC_CHUNK;
        pthread_mutex_unlock(a->mu);
        if (start >= a->n) break;
        uint32_t end = start + ORC_CHUNK < a->n ? start + ORC_CHUNK : a->n;
        for (uint32_t i = start; i < end; ++i) {
            switch (a->mode) {
            case ORC_MODE_REF: orc_closest_hit(a->s, &a->rays[i], &a->hits[i], &a->stats); break;
            case ORC_MODE_PRUNED: orc_closest_hit_pruned(a->s, &a->rays[i], &a->hits[i], &a->stats); break;
            case ORC_MODE_BRUTE: orc_closest_hit_brute(a->s, &a->rays[i], &a->hits[i]); break;
            case ORC_MODE_WIDE: orc_closest_hit_wide(a->s, &a->rays[i], &a->hits[i], &a->stats); break;
            case ORC_MODE_TWO_LEVEL: orc_closest_two_level(a->s, &a->rays[i], &a->hits[i], &a->stats); break;
            case ORC_MODE_ANY_TWO_LEVEL: orc_any_two_level(a->s, &a->rays[i], &a->hits[i], &a->stats); break;
            case ORC_MODE_ANY_BRUTE: orc_any_hit_brute(a->s, &a->rays[i], &a->hits[i]); break;
            default: orc_any_hit(a->s, &a->rays[i], &a->hits[i], &a->stats); break;
            }
        }
    }
    return NULL;
}

void orc_trace_batch(const OrcScene* s, const StkRay* rays, uint32_t n, StkHit* hits, int mode, int nthreads,
                     OrcStats* stats)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    volatile uint32_t next = 0;
    pthread_mutex_t mu;
    pthread_mutex_init(&mu, NULL);
    BatchArg* args = (BatchArg*)calloc((size_t)nthreads, sizeof(BatchArg));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int i = 0; i < nthreads; ++i) {
        args[i].s = s;
        args[i].rays = rays;
        args[i].hits = hits;
        args[i].n = n;
        args[i].mode = mode;
        args[i].next = &next;
        args[i].mu = &mu;
        if (nthreads == 1)
            batch_worker(&args[i]);
        else
            pthread_create(&th[i], NULL, batch_worker, &args[i]);
    }
    if (nthreads > 1)
        for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        for (int i = 0; i < nthreads; ++i) {
            stats->inner += args[i].stats.inner;
            stats->leaf += args[i].stats.leaf;
            stats->rays += args[i].stats.rays;
            stats->hits += args[i].stats.hits;
        }
    }
    pthread_mutex_destroy(&mu);
    free(args);
    free(th);
}

/* ------------------------------------------------------------------------ */
/* RNG (shaders/random.h:14-47) and CMJ helpers (shaders/helper.h:117-163)    */
/* ------------------------------------------------------------------------ */
uint32_t orc_jenkins(uint32_t x)
{
    x += x << 10;
    x ^= x >> 6;
    x += x << 3;
    x ^= x >> 11;
    x += x << 15;
    return x;
}

uint32_t orc_init_rng(uint32_t px, uint32_t py, uint32_t width, uint32_t frame)
{
    const uint32_t seed = (px * 1u + py * width) ^ orc_jenkins(frame);
    return orc_jenkins(seed);
}

uint32_t orc_xorshift(uint32_t* s)
{
    *s ^= *s << 13;
    *s ^= *s >> 17;
    *s ^= *s << 5;
    return *s;
}

float orc_rand(uint32_t* s) { return as_float(0x3f800000u | (orc_xorshift(s) >> 9)) - 1.0f; }

uint32_t orc_permute(uint32_t i, uint32_t l, uint32_t p)
{
    uint32_t w = l - 1;
    w |= w >> 1;
    w |= w >> 2;
    w |= w >> 4;
    w |= w >> 8;
    w |= w >> 16;
    do {
        i ^= p;
        i *= 0xe170893du;
        i ^= p >> 16;
        i ^= (i & w) >> 4;
        i ^= p >> 8;
        i *= 0x0929eb3fu;
        i ^= p >> 23;
        i ^= (i & w) >> 1;
        i *= 1u | p >> 27;
        i *= 0x6935fa69u;
        i ^= (i & w) >> 11;
        i *= 0x74dcb303u;
        i ^= (i & w) >> 2;
        i *= 0x9e501cc3u;
        i ^= (i & w) >> 2;
        i *= 0xc860a3dfu;
        i &= w;
        i ^= i >> 5;
    } while (i >= l);
    return (i + p) % l;
}

float orc_randfloat(uint32_t i, uint32_t p)
{
    i ^= p;
    i ^= i >> 17;
    i ^= i >> 10;
    i *= 0xb36534e5u;
    i ^= i >> 12;
    i ^= i >> 21;
    i *= 0x93fc4795u;
    i ^= 0xdf6e307fu;
    i ^= i >> 17;
    i *= 1u | p >> 18;
    return (float)i * (1.0f / 4294967808.0f);
}

/* ------------------------------------------------------------------------ */
/* Camera (shaders/pathtracerMdl.hlsl:95-202)                                  */
/* ------------------------------------------------------------------------ */
static void pixel_jitter(const StkPathTracerParam* p, uint32_t s, uint32_t* rng, float* jx, float* jy)
{
    switch (p->stratifiedSamplingType) {
    case 1:
        *jx = orc_rand(rng);
        *jy = orc_rand(rng);
        break;
    case 2: { /* stratifiedSampling1, N = 16 (:112-125) */
        const uint32_t N = 16;
        const int m = (int)sqrtf((float)N);
        const int n = ((int)N + m - 1) / m;
        const int x = (int)((s / (uint32_t)n) % (uint32_t)n);
        const int y = (int)((s % (uint32_t)m) % (uint32_t)m);
        const float r0 = orc_rand(rng);
        *jx = ((float)x + ((float)y + r0) / (float)n) / (float)m;
        const float r1 = orc_rand(rng);
        *jy = ((float)y + ((float)x + r1) / (float)m) / (float)n;
        break;
    }
    case 3: { /* stratifiedSamplingOptimized, N = 16, p = 16, a = 1 (:95-110) */
        const int N = 16, pp = 16;
        const int m = (int)sqrtf((float)N * 1.0f);
        const int n = (N + m - 1) / m;
        int si = (int)orc_permute((uint32_t)s, (uint32_t)N, (uint32_t)pp * 0x51633e2du);
        const int sx = (int)orc_permute((uint32_t)(si % m), (uint32_t)m, (uint32_t)pp * 0x68bc21ebu);
        const int sy = (int)orc_permute((uint32_t)(si / m), (uint32_t)n, (uint32_t)pp * 0x02e5be93u);
        const float fx = orc_randfloat((uint32_t)si, (uint32_t)pp * 0x967a889bu);
        const float fy = orc_randfloat((uint32_t)si, (uint32_t)pp * 0x368cc8b7u);
        *jx = ((float)sx + ((float)sy + fx) / (float)n) / (float)m;
        *jy = ((float)si + fy) / (float)N;
        break;
    }
    default:
        *jx = 0.5f;
        *jy = 0.5f;
        break;
    }
}

/* generateCameraRay (shaders/pathtracerMdl.hlsl:165-202). */
void orc_pixel_jitter(const StkPathTracerParam* p, uint32_t s, uint32_t* rng, float* jx, float* jy)
{
    pixel_jitter(p, s, rng, jx, jy);
}

void orc_camera_ray(const StkPathTracerParam* p, uint32_t px, uint32_t py, uint32_t* rng, uint32_t s, StkRay* ray)
{
    float jx, jy;
    pixel_jitter(p, s, rng, &jx, &jy);
    const float posx = (float)px + jx, posy = (float)py + jy;
    const float ndcx = (posx / (float)p->dimension[0]) * 2.0f - 1.0f;
    const float ndcy = (posy / (float)p->dimension[1]) * 2.0f - 1.0f;
    /* viewSpace = clipToView * (ndc, 1, 1) */
    const float* C = p->clipToView;
    const float vx = (C[0] * ndcx + C[4] * ndcy) + (C[8] * 1.0f + C[12] * 1.0f);
    const float vy = (C[1] * ndcx + C[5] * ndcy) + (C[9] * 1.0f + C[13] * 1.0f);
    const float vz = (C[2] * ndcx + C[6] * ndcy) + (C[10] * 1.0f + C[14] * 1.0f);
    const f3 wdir = xform4(p->viewToWorld, mk(vx, vy, vz), 0.0f);
    const f3 o = xform4(p->viewToWorld, mk(0.0f, 0.0f, 0.0f), 1.0f);
    const f3 d = nrm(wdir);
    ray->o[0] = o.x;
    ray->o[1] = o.y;
    ray->o[2] = o.z;
    ray->o[3] = 1e9f;
    ray->d[0] = d.x;
    ray->d[1] = d.y;
    ray->d[2] = d.z;
    ray->d[3] = 0.0f;
}

/* offset_ray (shaders/helper.h:12-28), Ray Tracing Gems ch. 6. */
static f3 offset_ray(f3 p, f3 n)
{
    const float origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f, int_scale = 256.0f;
    const int32_t ox = (int32_t)(int_scale * n.x), oy = (int32_t)(int_scale * n.y), oz = (int32_t)(int_scale * n.z);
    const float pix = int_as_float(as_int(p.x) + ((p.x < 0) ? -ox : ox));
    const float piy = int_as_float(as_int(p.y) + ((p.y < 0) ? -oy : oy));
    const float piz = int_as_float(as_int(p.z) + ((p.z < 0) ? -oz : oz));
    return mk(fabsf(p.x) < origin ? p.x + float_scale * n.x : pix, fabsf(p.y) < origin ? p.y + float_scale * n.y : piy,
              fabsf(p.z) < origin ? p.z + float_scale * n.z : piz);
}

void orc_offset_ray(const float p[3], const float n[3], float out[3]) { st3(out, offset_ray(ld(p), ld(n))); }

/* ------------------------------------------------------------------------ */
/* Rect light (shaders/lights.h:102-194, 230-301)                             */
/* ------------------------------------------------------------------------ */
typedef struct {
    f3 o, x, y, z;
    float z0, z0sq, x0, y0, y0sq, x1, y1, y1sq, b0, b1, b0sq, k, S;
} SphQuad;

static SphQuad sphquad_init(const StkLight* l, f3 o)
{
    SphQuad q;
    const f3 p0 = ld(l->points[0]), p1 = ld(l->points[1]), p3 = ld(l->points[3]);
    const f3 ex = sub(p1, p0), ey = sub(p3, p0);
    const f3 s = p0;
    const float exl = len3(ex), eyl = len3(ey);
    q.o = o;
    q.x = mk(ex.x / exl, ex.y / exl, ex.z / exl);
    q.y = mk(ey.x / eyl, ey.y / eyl, ey.z / eyl);
    q.z = cross(q.x, q.y);
    const f3 d = sub(s, o);
    q.z0 = dot(d, q.z);
    if (q.z0 > 0) {
        q.z = scl(q.z, -1.0f);
        q.z0 *= -1.0f;
    }
    q.z0sq = q.z0 * q.z0;
    q.x0 = dot(d, q.x);
    q.y0 = dot(d, q.y);
    q.x1 = q.x0 + exl;
    q.y1 = q.y0 + eyl;
    q.y0sq = q.y0 * q.y0;
    q.y1sq = q.y1 * q.y1;
    const f3 v00 = mk(q.x0, q.y0, q.z0), v01 = mk(q.x0, q.y1, q.z0);
    const f3 v10 = mk(q.x1, q.y0, q.z0), v11 = mk(q.x1, q.y1, q.z0);
    const f3 n0 = nrm(cross(v00, v10)), n1 = nrm(cross(v10, v11));
    const f3 n2 = nrm(cross(v11, v01)), n3 = nrm(cross(v01, v00));
    const float g0 = stk_acosf(-dot(n0, n1)), g1 = stk_acosf(-dot(n1, n2));
    const float g2 = stk_acosf(-dot(n2, n3)), g3 = stk_acosf(-dot(n3, n0));
    q.b0 = n0.z;
    q.b1 = n2.z;
    q.b0sq = q.b0 * q.b0;
    q.k = 2.0f * ORC_PI - g2 - g3;
    q.S = g0 + g1 - q.k;
    return q;
}

static f3 sphquad_sample(const SphQuad* q, float u, float v)
{
    const float au = u * q->S + q->k;
    const float fu = (stk_cosf(au) * q->b0 - q->b1) / stk_sinf(au);
    float cu = 1.0f / sqrtf(fu * fu + q->b0sq) * (fu > 0 ? 1.0f : -1.0f);
    cu = clampf(cu, -1.0f, 1.0f);
    float xu = -(cu * q->z0) / sqrtf(1.0f - cu * cu);
    xu = clampf(xu, q->x0, q->x1);
    const float d = sqrtf(xu * xu + q->z0sq);
    const float h0 = q->y0 / sqrtf(d * d + q->y0sq);
    const float h1 = q->y1 / sqrtf(d * d + q->y1sq);
    const float hv = h0 + v * (h1 - h0);
    const float hv2 = hv * hv;
    const float eps = 1e-5f;
    const float yv = (hv < 1.0f - eps) ? (hv * d) / sqrtf(1.0f - hv2) : q->y1;
    return add(add(add(q->o, scl(q->x, xu)), scl(q->y, yv)), scl(q->z, q->z0));
}

/* calcLightNormal of a rect (shaders/lights.h:230-246): the direction it emits along. */
static f3 rect_light_normal(const StkLight* l)
{
    const f3 p0 = ld(l->points[0]), p1 = ld(l->points[1]), p3 = ld(l->points[3]);
    return neg(nrm(cross(sub(p1, p0), sub(p3, p0))));
}

void orc_sample_rect_light(const StkLight* l, const float u[2], const float hitPoint[3], float point[3], float L[3],
                           float* dist, float normal[3], float* area, float* pdf, float* solidAngle)
{
    const f3 hp = ld(hitPoint);
    SphQuad q = sphquad_init(l, hp);
    const f3 pt = sphquad_sample(&q, u[0], u[1]);
    /* fillLightData (:274-282) with calcLightArea/calcLightNormal for type 0 */
    const f3 p0 = ld(l->points[0]), p1 = ld(l->points[1]), p3 = ld(l->points[3]);
    const f3 e1 = sub(p1, p0), e2 = sub(p3, p0);
    const float A = len3(cross(e1, e2));
    const f3 nl = neg(nrm(cross(e1, e2)));
    const f3 toL = sub(pt, hp);
    const float lenL = len3(toL);
    const f3 Ld = mk(toL.x / lenL, toL.y / lenL, toL.z / lenL);
    st3(point, pt);
    st3(L, Ld);
    *dist = lenL;
    st3(normal, nl);
    *area = A;
    *pdf = lenL * lenL / (-dot(Ld, nl) * A);
    if (solidAngle) *solidAngle = q.S;
}

/* ------------------------------------------------------------------------ */
/* BSDFs (non-MDL).  Lambert follows sampleMaterial                           */
/* (shaders/pathtracer.hlsl:176-234) with SampleHemisphereCosine and          */
/* GetTangentSpace (shaders/helper.h:41-50,73-79); GGX uses                   */
/* SampleGGXDistribution (helper.h:93-99), ggxNormalDistribution,             */
/* ggxVisibility and schlickFresnel (shaders/pathtracer.hlsl:108-151) and the */
/* half-vector pdf D*cos(theta_h)/(4|k1.h|) (Walter et al. 2007).  Outputs    */
/* follow Bsdf_sample_data / Bsdf_evaluate_data (shaders/mdl_types.hlsl:     */
/* 271-304): bsdf_over_pdf = f*cos/pdf, eval = f*cos, pdf in solid angle.     */
/* ------------------------------------------------------------------------ */
static void tangent_space(f3 n, f3* t, f3* b)
{
    f3 helper = mk(1.0f, 0.0f, 0.0f);
    if (fabsf(n.x) > 0.99f) helper = mk(0.0f, 0.0f, 1.0f);
    *t = nrm(cross(n, helper));
    *b = nrm(cross(n, *t));
}
/* mul(transpose(float3x3(t, b, n)), v) = t*v.x + b*v.y + n*v.z */
static inline f3 to_world(f3 t, f3 b, f3 n, f3 v) { return add(add(scl(t, v.x), scl(b, v.y)), scl(n, v.z)); }

static inline f3 sample_cosine(float u, float v)
{
    const float cosTheta = sqrtf(saturate(u));
    const float sinTheta = sqrtf(saturate(1.0f - cosTheta * cosTheta));
    const float phi = 2.0f * ORC_PI * v;
    return mk(stk_cosf(phi) * sinTheta, stk_sinf(phi) * sinTheta, cosTheta + 1e-6f);
}

static inline f3 sample_ggx(float u, float v, float alpha)
{
    const float cosTheta = sqrtf((1.0f - u) / (u * (alpha * alpha - 1.0f) + 1.0f));
    const float sinTheta = sqrtf(saturate(1.0f - cosTheta * cosTheta));
    const float phi = 2.0f * ORC_PI * v;
    return mk(stk_cosf(phi) * sinTheta, stk_sinf(phi) * sinTheta, cosTheta + 1e-6f);
}

static inline float ggx_d(float a2, f3 n, f3 m)
{
    const float c = dot(n, m);
    const float c2 = c * c;
    const float den = c2 * (a2 - 1.0f) + 1.0f;
    return c > 0.0f ? (a2 / (ORC_PI * den * den)) : 0.0f;
}

static inline float ggx_g1(float a2, f3 w, f3 n, f3 m)
{
    const float NdotW = dot(n, w);
    const float MdotW = dot(m, w);
    if (MdotW * NdotW <= 0.0f) return 0.0f;
    const float c2 = NdotW * NdotW;
    const float tan2 = (1.0f - c2) / c2;
    if (tan2 == 0.0f) return 1.0f;
    return 2.0f / (1.0f + sqrtf(1.0f + a2 * tan2));
}

static inline f3 schlick(f3 f0, float c)
{
    const float w = stk_pow5f(1.0f - c);
    return mk(f0.x + (1.0f - f0.x) * w, f0.y + (1.0f - f0.y) * w, f0.z + (1.0f - f0.z) * w);
}

static inline float ggx_alpha(const StkMaterial* m)
{
    const float a = m->roughness * m->roughness;
    return a < 1e-4f ? 1e-4f : a;
}

void orc_bsdf_sample(const StkMaterial* m, const float Np[3], const float Ngp[3], const float k1p[3], const float xi[4],
                     OrcBsdfSample* out)
{
    const f3 N = ld(Np), Ng = ld(Ngp), k1 = ld(k1p), albedo = ld(m->albedo);
    memset(out, 0, sizeof(*out));
    if (m->type == STK_MAT_MIRROR) {
        /* ideal reflection about the shading normal, Schlick(albedo, cos) */
        const float c1 = dot(N, k1);
        const f3 k2 = sub(scl(N, 2.0f * c1), k1);
        if (c1 <= 0.0f || dot(k2, Ng) <= 0.0f) return;
        st3(out->k2, k2);
        st3(out->bsdf_over_pdf, schlick(albedo, c1));
        out->event = ORC_EVENT_SPECULAR_REFLECTION;
        return;
    }
    if (m->type == STK_MAT_GLASS) {
        /* smooth dielectric; the side is the geometric normal's */
        const float ior = m->ior > 0.0f ? m->ior : 1.5f;
        const int entering = dot(k1, Ng) > 0.0f;
        const f3 n = entering ? N : neg(N), ng = entering ? Ng : neg(Ng);
        const float eta = entering ? 1.0f / ior : ior; /* n1 / n2 */
        const float c1 = dot(n, k1);
        if (c1 <= 0.0f) return;
        const float s2 = eta * eta * (1.0f - c1 * c1);
        float F = 1.0f, c2 = 0.0f;
        if (s2 < 1.0f) {
            /* Fresnel's equations, unpolarised */
            c2 = sqrtf(1.0f - s2);
            const float rs = (eta * c1 - c2) / (eta * c1 + c2);
            const float rp = (c1 - eta * c2) / (c1 + eta * c2);
            F = 0.5f * (rs * rs + rp * rp);
        }
        if (xi[2] < F) {
            const f3 k2 = sub(scl(n, 2.0f * c1), k1);
            if (dot(k2, ng) <= 0.0f) return;
            st3(out->k2, k2);
            st3(out->bsdf_over_pdf, mk(1.0f, 1.0f, 1.0f));
            out->event = ORC_EVENT_SPECULAR_REFLECTION;
        } else {
            const f3 k2 = sub(scl(n, eta * c1 - c2), scl(k1, eta)); /* Snell's law */
            if (dot(k2, ng) >= 0.0f) return;
            st3(out->k2, k2);
            st3(out->bsdf_over_pdf, albedo);
            out->event = ORC_EVENT_SPECULAR_TRANSMISSION;
        }
        return;
    }
    f3 T, B;
    tangent_space(N, &T, &B);
    if (m->type == STK_MAT_GGX) {
        const float alpha = ggx_alpha(m);
        const float a2 = alpha * alpha;
        const f3 h = nrm(to_world(T, B, N, sample_ggx(xi[0], xi[1], alpha)));
        const float k1h = dot(k1, h);
        const f3 k2 = sub(scl(h, 2.0f * k1h), k1);
        const float nk1 = dot(N, k1), nk2 = dot(N, k2), nh = dot(N, h);
        if (nk1 <= 0.0f || nk2 <= 0.0f || dot(k2, Ng) <= 0.0f || k1h <= 0.0f) {
            out->event = 0;
            return;
        }
        const float D = ggx_d(a2, N, h);
        const float G = ggx_g1(a2, k1, N, h) * ggx_g1(a2, k2, N, h);
        const f3 F = schlick(albedo, dot(k2, h));
        st3(out->k2, k2);
        out->pdf = D * nh / (4.0f * k1h);
        const float w = G * k1h / (nk1 * nh);
        st3(out->bsdf_over_pdf, scl(F, w));
        out->event = 2;
        return;
    }
    /* Lambert */
    const f3 k2 = to_world(T, B, N, sample_cosine(xi[0], xi[1]));
    if (dot(k2, Ng) <= 0.0f) {
        out->event = 0;
        return;
    }
    st3(out->k2, k2);
    out->pdf = ORC_INV_PI * (saturate(dot(N, k2)) + 1e-6f);
    st3(out->bsdf_over_pdf, albedo);
    out->event = 1;
}

void orc_bsdf_eval(const StkMaterial* m, const float Np[3], const float Ngp[3], const float k1p[3], const float k2p[3],
                   float diffuse[3], float glossy[3], float* pdf)
{
    const f3 N = ld(Np), Ng = ld(Ngp), k1 = ld(k1p), k2 = ld(k2p), albedo = ld(m->albedo);
    const f3 zero = mk(0.0f, 0.0f, 0.0f);
    st3(diffuse, zero);
    st3(glossy, zero);
    *pdf = 0.0f;
    const float nk2 = dot(N, k2);
    if (m->type == STK_MAT_MIRROR || m->type == STK_MAT_GLASS) return; /* zero off the specular direction */
    if (nk2 <= 0.0f || dot(k2, Ng) <= 0.0f) return;
    if (m->type == STK_MAT_GGX) {
        const float nk1 = dot(N, k1);
        if (nk1 <= 0.0f) return;
        const float alpha = ggx_alpha(m);
        const float a2 = alpha * alpha;
        const f3 h = nrm(add(k1, k2));
        const float k1h = dot(k1, h);
        if (k1h <= 0.0f) return;
        const float nh = dot(N, h);
        const float D = ggx_d(a2, N, h);
        const float G = ggx_g1(a2, k1, N, h) * ggx_g1(a2, k2, N, h);
        const f3 F = schlick(albedo, dot(k2, h));
        const float w = D * G / (4.0f * nk1);
        st3(glossy, scl(F, w));
        *pdf = D * nh / (4.0f * k1h);
        return;
    }
    st3(diffuse, scl(albedo, ORC_INV_PI * nk2));
    *pdf = ORC_INV_PI * (nk2 + 1e-6f);
}

/* ------------------------------------------------------------------------ */
/* Unpackers (shaders/pack.h:4-33) and attribute interpolation               */
/* (shaders/helper.h:31-39)                                                   */
/* ------------------------------------------------------------------------ */
static inline f3 unpack_normal(uint32_t v)
{
    f3 n;
    n.z = (float)((v & 0xfff00000u) >> 20) / 511.99999f * 2.0f - 1.0f;
    n.y = (float)((v & 0x000ffc00u) >> 10) / 511.99999f * 2.0f - 1.0f;
    n.x = (float)(v & 0x000003ffu) / 511.99999f * 2.0f - 1.0f;
    return n;
}

static inline f3 interp(f3 a, f3 b, f3 c, float u, float v)
{
    const float w = 1.0f - u - v;
    return add(add(scl(a, w), scl(b, u)), scl(c, v));
}

/* ------------------------------------------------------------------------ */
/* Path tracer: pathTraceCameraRays (shaders/pathtracerMdl.hlsl:204-404)      */
/* with estimateDirectLighting/sampleLights (:31-92).                         */
/* ------------------------------------------------------------------------ */
static const StkMaterial k_default_material = {STK_MAT_LAMBERT, 1.0f, 0u, 0u, {0.5f, 0.5f, 0.5f, 1.0f},
                                               {0.0f, 0.0f, 0.0f}, 0.0f, 0u, 0u, {0u, 0u}};

/* unpackUV (shaders/pack.h:15-23) */
static inline void unpack_uv(uint32_t p, float* s, float* t)
{
    *t = (float)((p & 0xffff0000u) >> 16) / 16383.99999f * 20.0f - 10.0f;
    *s = (float)(p & 0x0000ffffu) / 16383.99999f * 20.0f - 10.0f;
}

/* One axis of tex_lookup_float4_2d (shaders/mdl_runtime.hlsl:173-207): the
 * wrap of apply_wrap_and_crop (:93-120: repeat leaves the coordinate to the
 * fetch; mirrored repeat folds it into [0, 1]; every wrap but repeat clamps
 * it to the edge texels' centres), the quintic smoothing of
 * apply_smootherstep_filter (:125-140), then the bilinear filter's texel
 * coordinate x * n - 0.5. */
static float tex_axis(float c, uint32_t n, uint32_t wrap)
{
    const float fn = (float)n;
    if (wrap != STK_TEX_REPEAT) {
        if (wrap == STK_TEX_MIRRORED_REPEAT) {
            const float fl = floorf(c);
            c = ((int)fl & 1) != 0 ? 1.0f - (c - fl) : c - fl;
        }
        const float h = 0.5f / fn;
        c = fminf(fmaxf(c, h), 1.0f - h);
    }
    float r = c * fn + 0.5f;
    const float i = floorf(r);
    float f = r - i;
    f = f * f * f * (f * (f * 6.0f - 15.0f) + 10.0f);
    r = i + f;
    return ((r - 0.5f) / fn) * fn - 0.5f;
}

static int tex_wrap(int i, int n, int clampEdge)
{
    if (clampEdge) return i < 0 ? 0 : (i >= n ? n - 1 : i);
    const int m = i % n;
    return m < 0 ? m + n : m;
}

void orc_texture_lookup(const StkTexture* tx, const float* texels, float s, float t, float out[3])
{
    if (tx->wrap == STK_TEX_CLIP && (s < 0.0f || s >= 1.0f || t < 0.0f || t >= 1.0f)) {
        /* clip: zero outside the texture (:186-189) */
        out[0] = out[1] = out[2] = 0.0f;
        return;
    }
    const int ce = tx->wrap != STK_TEX_REPEAT;
    const float x = tex_axis(s, tx->width, tx->wrap), y = tex_axis(t, tx->height, tx->wrap);
    const float x0 = floorf(x), y0 = floorf(y);
    const float fx = x - x0, fy = y - y0;
    const int w = (int)tx->width, h = (int)tx->height;
    const int ix0 = tex_wrap((int)x0, w, ce), ix1 = tex_wrap((int)x0 + 1, w, ce);
    const int iy0 = tex_wrap((int)y0, h, ce), iy1 = tex_wrap((int)y0 + 1, h, ce);
    const float* a = texels + 4 * ((size_t)tx->offset + (size_t)iy0 * tx->width + ix0);
    const float* b = texels + 4 * ((size_t)tx->offset + (size_t)iy0 * tx->width + ix1);
    const float* c = texels + 4 * ((size_t)tx->offset + (size_t)iy1 * tx->width + ix0);
    const float* d = texels + 4 * ((size_t)tx->offset + (size_t)iy1 * tx->width + ix1);
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    for (int k = 0; k < 3; ++k) out[k] = (a[k] * gx + b[k] * fx) * gy + (c[k] * gx + d[k] * fx) * fy;
}

/* CalcBumpedNormal (shaders/helper.h:101-115): the tangent negated and made
 * orthogonal to the normal (Gram-Schmidt), the bitangent N x T, and the
 * map's texel (rgb = (n + 1) / 2) taken from that frame to world space:
 * normalize(mul(transpose(float3x3(T, B, N)), texel * 2 - 1)). */
void orc_bumped_normal(const float normal[3], const float tangent[3], const float texel[3], float out[3])
{
    const f3 n = nrm(ld(normal));
    f3 t = neg(nrm(ld(tangent)));
    t = nrm(sub(t, scl(n, dot(t, n))));
    const f3 b = cross(n, t);
    const f3 m = mk(texel[0] * 2.0f - 1.0f, texel[1] * 2.0f - 1.0f, texel[2] * 2.0f - 1.0f);
    st3(out, nrm(to_world(t, b, n, m)));
}

void orc_pt_sample(const OrcScene* s, const StkRenderParams* rp, uint32_t px, uint32_t py, uint32_t sampleNum,
                   float out[3])
{
    const StkPathTracerParam* p = &rp->pt;
    const uint32_t frame = p->frameNumber * (uint32_t)p->spp + sampleNum;
    uint32_t rng = orc_init_rng(px, py, (uint32_t)p->dimension[0], frame);
    StkRay ray;
    orc_camera_ray(p, px, py, &rng, frame, &ray);
    f3 color = mk(0.0f, 0.0f, 0.0f);
    f3 thr = mk(1.0f, 1.0f, 1.0f);
    const f3 sky = ld(rp->skyColor);
    int specularBounce = 0;
    const int maxDepth = (int)p->maxDepth;
    const int ref = rp->refCompat != 0;
    float prevPdf = 0.0f; /* solid-angle pdf of the BSDF sample that made `ray` (corrected estimator) */
    for (int depth = 0; depth < maxDepth; ++depth) {
        StkHit hit;
        if (!orc_closest_hit(s, &ray, &hit, NULL)) {
            color = add(color, mul(thr, sky));
            break;
        }
        const StkInstance* ic = &s->inst[hit.instId];
        if (ic->lightId != -1) {
            const StkLight* l = &s->lights[ic->lightId];
            if (ref) {
                /* emitter: counted on camera rays and after specular bounces
                 * (:229-234); a BSDF-sampled emitter hit after a diffuse or glossy
                 * bounce adds nothing and ends the path (the reference indexes
                 * material -1 there). */
                if (depth == 0 || specularBounce) color = add(color, mul(thr, ld(l->color)));
            } else if (dot(ld(ray.d), rect_light_normal(l)) < 0.0f) {
                /* corrected: the front face emits; after a diffuse or glossy
                 * bounce the hit is weighted against the previous vertex's light
                 * sample (balance heuristic, both pdfs in solid angle) */
                float w = 1.0f;
                if (depth > 0 && !specularBounce) {
                    const SphQuad q = sphquad_init(l, ld(ray.o));
                    const float lpdf = q.S > 0.0f ? (1.0f / (float)s->nlights) / q.S : 0.0f;
                    w = prevPdf / (prevPdf + lpdf);
                }
                color = add(color, scl(mul(thr, ld(l->color)), w));
            }
            break;
        }
        const uint32_t i0 = s->ib[(uint32_t)ic->indexOffset + hit.primId * 3 + 0];
        const uint32_t i1 = s->ib[(uint32_t)ic->indexOffset + hit.primId * 3 + 1];
        const uint32_t i2 = s->ib[(uint32_t)ic->indexOffset + hit.primId * 3 + 2];
        const f3 p0 = xform4(ic->objectToWorld, ld(s->vb[i0].position), 1.0f);
        const f3 p1 = xform4(ic->objectToWorld, ld(s->vb[i1].position), 1.0f);
        const f3 p2 = xform4(ic->objectToWorld, ld(s->vb[i2].position), 1.0f);
        const f3 Ng = nrm(cross(sub(p1, p0), sub(p2, p0)));
        const f3 n0 = xform3(ic->normalMatrix, unpack_normal(s->vb[i0].normal));
        const f3 n1 = xform3(ic->normalMatrix, unpack_normal(s->vb[i1].normal));
        const f3 n2 = xform3(ic->normalMatrix, unpack_normal(s->vb[i2].normal));
        const f3 P = interp(p0, p1, p2, hit.bary[0], hit.bary[1]);
        StkMaterial matv = (ic->materialId >= 0 && (uint32_t)ic->materialId < s->nmats) ? s->mats[ic->materialId]
                                                                                         : k_default_material;
        f3 N = nrm(interp(n0, n1, n2, hit.bary[0], hit.bary[1]));
        f3 emit = ld(matv.emission);
        const int albTex = matv.albedoTexture != 0 && matv.albedoTexture <= s->ntextures;
        const int roughTex = matv.roughnessTexture != 0 && matv.roughnessTexture <= s->ntextures;
        const int normTex = matv.normalTexture != 0 && matv.normalTexture <= s->ntextures;
        const int emitTex = matv.emissionTexture != 0 && matv.emissionTexture <= s->ntextures;
        if (albTex || roughTex || normTex || emitTex) {
            /* texture coordinates as shaders/pathtracerMdl.hlsl:261-265 */
            float s0, t0, s1, t1, s2, t2, tc[3];
            unpack_uv(s->vb[i0].uv, &s0, &t0);
            unpack_uv(s->vb[i1].uv, &s1, &t1);
            unpack_uv(s->vb[i2].uv, &s2, &t2);
            const float w = 1.0f - hit.bary[0] - hit.bary[1];
            const float su = (s0 * w + s1 * hit.bary[0]) + s2 * hit.bary[1];
            const float tv = (t0 * w + t1 * hit.bary[0]) + t2 * hit.bary[1];
            if (albTex) {
                orc_texture_lookup(&s->textures[matv.albedoTexture - 1], s->texels, su, tv, tc);
                for (int k = 0; k < 3; ++k) matv.albedo[k] = matv.albedo[k] * tc[k];
            }
            if (roughTex) {
                orc_texture_lookup(&s->textures[matv.roughnessTexture - 1], s->texels, su, tv, tc);
                matv.roughness = matv.roughness * tc[0];
            }
            if (emitTex) {
                orc_texture_lookup(&s->textures[matv.emissionTexture - 1], s->texels, su, tv, tc);
                emit = mul(emit, ld(tc));
            }
            if (normTex) {
                /* the world tangent as the reference forms it (:247-257) */
                const f3 ta = xform3(ic->normalMatrix, unpack_normal(s->vb[i0].tangent));
                const f3 tb = xform3(ic->normalMatrix, unpack_normal(s->vb[i1].tangent));
                const f3 tcn = xform3(ic->normalMatrix, unpack_normal(s->vb[i2].tangent));
                const f3 T = nrm(interp(ta, tb, tcn, hit.bary[0], hit.bary[1]));
                float bn[3];
                orc_texture_lookup(&s->textures[matv.normalTexture - 1], s->texels, su, tv, tc);
                orc_bumped_normal((const float*)&N, (const float*)&T, tc, bn);
                /* a tangent along the normal leaves no frame: keep the normal */
                if (!(isnan(bn[0]) || isnan(bn[1]) || isnan(bn[2]))) N = ld(bn);
            }
        }
        const StkMaterial* mat = &matv;
        const f3 k1 = neg(ld(ray.d));
        /* the EDF: the front face emits (:299-304) */
        if ((emit.x != 0.0f || emit.y != 0.0f || emit.z != 0.0f) && dot(k1, Ng) > 0.0f) color = add(color, mul(thr, emit));

        /* sampleLights + estimateDirectLighting */
        f3 toLight = mk(0.0f, 0.0f, 0.0f);
        float lightPdf = 0.0f;
        f3 radiance = mk(0.0f, 0.0f, 0.0f);
        if (s->nlights > 0) {
            const uint32_t lid = (uint32_t)((float)s->nlights * orc_rand(&rng));
            const float selPdf = 1.0f / (float)s->nlights;
            const StkLight* l = &s->lights[lid < s->nlights ? lid : s->nlights - 1];
            const float u0 = orc_rand(&rng);
            const float u1 = orc_rand(&rng);
            float uu[2] = {u0, u1}, pt[3], L[3], dist, ln[3], area, pdf, solid;
            orc_sample_rect_light(l, uu, (const float*)&P, pt, L, &dist, ln, &area, &pdf, &solid);
            toLight = ld(L);
            const f3 Li = ld(l->color);
            if (dot(N, toLight) > 0.0f && -dot(toLight, ld(ln)) > 0.0f && Li.x != 0.0f && Li.y != 0.0f &&
                Li.z != 0.0f) {
                StkRay sr;
                const f3 so = offset_ray(P, Ng);
                sr.o[0] = so.x;
                sr.o[1] = so.y;
                sr.o[2] = so.z;
                sr.o[3] = dist;
                sr.d[0] = toLight.x;
                sr.d[1] = toLight.y;
                sr.d[2] = toLight.z;
                sr.d[3] = 0.0f;
                StkHit sh;
                float vis;
                if (ref) {
                    vis = orc_any_hit(s, &sr, &sh, NULL) ? 0.0f : 1.0f;
                    if (vis == 0.0f && s->inst[sh.instId].lightId != -1) vis = 1.0f; /* Q5 */
                    /* area-measure pdf against solid-angle samples (Q6); the cosine
                     * the BSDF evaluation carries is applied again */
                    lightPdf = pdf;
                    radiance = scl(mul(scl(Li, vis), mk(1, 1, 1)), saturate(dot(N, toLight)));
                } else {
                    vis = orc_occluded(s, &sr) ? 0.0f : 1.0f;
                    lightPdf = solid > 0.0f ? 1.0f / solid : 0.0f;
                    radiance = scl(Li, vis);
                }
            }
            lightPdf *= selPdf;
        }
        if (isnan(radiance.x) || isnan(radiance.y) || isnan(radiance.z) || isnan(lightPdf)) break;
        const int nextEventValid = (dot(toLight, Ng) > 0.0f) && lightPdf != 0.0f;
        if (nextEventValid) {
            const f3 rop = mk(radiance.x / lightPdf, radiance.y / lightPdf, radiance.z / lightPdf);
            float bd[3], bg[3], bpdf;
            orc_bsdf_eval(mat, (const float*)&N, (const float*)&Ng, (const float*)&k1, (const float*)&toLight, bd, bg,
                          &bpdf);
            if (isnan(bd[0]) || isnan(bd[1]) || isnan(bd[2]) || isnan(bg[0]) || isnan(bg[1]) || isnan(bg[2])) break;
            if (bpdf > 0.0f) {
                const float mis = (lightPdf == 0.0f) ? 1.0f : lightPdf / (lightPdf + bpdf);
                const f3 w = scl(mul(thr, rop), mis);
                color = add(color, mul(w, ld(bd)));
                color = add(color, mul(w, ld(bg)));
            }
        }
        float xi[4];
        xi[0] = orc_rand(&rng);
        xi[1] = orc_rand(&rng);
        xi[2] = orc_rand(&rng);
        xi[3] = orc_rand(&rng);
        OrcBsdfSample bs;
        orc_bsdf_sample(mat, (const float*)&N, (const float*)&Ng, (const float*)&k1, xi, &bs);
        if (bs.event == 0) break;
        specularBounce = bs.event >= ORC_EVENT_SPECULAR_REFLECTION;
        /* offset to the side the new ray leaves by (:368-370) */
        const f3 nextO = offset_ray(P, dot(ld(bs.k2), Ng) > 0.0f ? Ng : neg(Ng));
        thr = mul(thr, ld(bs.bsdf_over_pdf));
        if (depth > (int)rp->rrStartDepth) {
            if (ref) {
                /* p = max(rgb), unclamped, and 1 / (p + 1e-5) (Q9) */
                const float pr = fmaxf(thr.x, fmaxf(thr.y, thr.z));
                if (orc_rand(&rng) > pr) break;
                thr = scl(thr, 1.0f / (pr + 1e-5f));
            } else {
                const float pr = fminf(fmaxf(thr.x, fmaxf(thr.y, thr.z)), 1.0f);
                if (orc_rand(&rng) > pr || !(pr > 0.0f)) break;
                thr = scl(thr, 1.0f / pr);
            }
        }
        prevPdf = bs.pdf;
        ray.o[0] = nextO.x;
        ray.o[1] = nextO.y;
        ray.o[2] = nextO.z;
        ray.o[3] = 1e9f;
        ray.d[0] = bs.k2[0];
        ray.d[1] = bs.k2[1];
        ray.d[2] = bs.k2[2];
        ray.d[3] = 0.0f;
    }
    st3(out, color);
}

typedef struct {
    const OrcScene* s;
    const StkRenderParams* rp;
    uint32_t x0, y0, w, h;
    float* samples;
    volatile uint32_t* next;
    pthread_mutex_t* mu;
} PtArg;

static void* pt_worker(void* pv)
{
    PtArg* a = (PtArg*)pv;
    const uint32_t spp = (uint32_t)a->rp->pt.spp;
    const uint32_t npix = a->w * a->h;
    for (;;) {
        uint32_t start;
        pthread_mutex_lock(a->mu);
        start = *a->next;
        *a->next = start + 64;
        pthread_mutex_unlock(a->mu);
        if (start >= npix) break;
        const uint32_t end = start + 64 < npix ? start + 64 : npix;
        for (uint32_t i = start; i < end; ++i) {
            const uint32_t px = a->x0 + i % a->w, py = a->y0 + i / a->w;
            for (uint32_t sidx = 0; sidx < spp; ++sidx)
                orc_pt_sample(a->s, a->rp, px, py, sidx, &a->samples[((size_t)i * spp + sidx) * 3]);
        }
    }
    return NULL;
}

void orc_pt_frame(const OrcScene* s, const StkRenderParams* rp, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                  float* samples, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    volatile uint32_t next = 0;
    pthread_mutex_t mu;
    pthread_mutex_init(&mu, NULL);
    PtArg a = {s, rp, x0, y0, w, h, samples, &next, &mu};
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    if (nthreads == 1)
        pt_worker(&a);
    else {
        for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, pt_worker, &a);
        for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    }
    free(th);
    pthread_mutex_destroy(&mu);
}

/* ------------------------------------------------------------------------ */
/* Post passes                                                                */
/* ------------------------------------------------------------------------ */
/* reduction.hlsl:11-32: color += invSpp * sample, in sample order. */
void orc_reduce(const float* samples, uint32_t npix, uint32_t spp, float* rgba)
{
    const float invSpp = 1.0f / (float)spp;
    for (uint32_t i = 0; i < npix; ++i) {
        float r = 0.0f, g = 0.0f, b = 0.0f;
        for (uint32_t s = 0; s < spp; ++s) {
            const float* x = &samples[((size_t)i * spp + s) * 3];
            r += invSpp * x[0];
            g += invSpp * x[1];
            b += invSpp * x[2];
        }
        rgba[i * 4 + 0] = r;
        rgba[i * 4 + 1] = g;
        rgba[i * 4 + 2] = b;
        rgba[i * 4 + 3] = 1.0f;
    }
}

/* accPt (shaders/accumulation.hlsl:86-97). */
void orc_accumulate(const float* cur, const float* hist, uint32_t npix, uint32_t iteration, float* out)
{
    const float wHistory = (float)iteration / (float)(iteration + 1);
    const float wNew = 1.0f / (float)(iteration + 1);
    for (uint32_t i = 0; i < npix; ++i) {
        for (int c = 0; c < 3; ++c) out[i * 4 + c] = hist[i * 4 + c] * wHistory + cur[i * 4 + c] * wNew;
        out[i * 4 + 3] = 1.0f;
    }
}

/* tonemap.hlsl:9-88 */
static const float kAcesIn[3][3] = {{0.59719f, 0.35458f, 0.04823f},
                                    {0.07600f, 0.90834f, 0.01566f},
                                    {0.02840f, 0.13383f, 0.83777f}};
static const float kAcesOut[3][3] = {{1.60475f, -0.53108f, -0.07367f},
                                     {-0.10208f, 1.10813f, -0.00605f},
                                     {-0.00327f, -0.07276f, 1.07602f}};

static f3 mat3_rows(const float m[3][3], f3 v)
{
    return mk(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
              m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
}

static float rrt_odt(float v)
{
    const float a = v * (v + 0.0245786f) - 0.000090537f;
    const float b = v * (0.983729f * v + 0.4329510f) + 0.238081f;
    return a / b;
}

void orc_tonemap(const float* in, uint32_t npix, uint32_t type, float* out)
{
    for (uint32_t i = 0; i < npix; ++i) {
        f3 c = ld(&in[i * 4]);
        f3 r;
        if (type == 1) {
            c = mat3_rows(kAcesIn, c);
            c = mk(rrt_odt(c.x), rrt_odt(c.y), rrt_odt(c.z));
            c = mat3_rows(kAcesOut, c);
            r = mk(saturate(c.x), saturate(c.y), saturate(c.z));
        } else if (type == 2) {
            const float a = 2.51f, b = 0.03f, cc = 2.43f, d = 0.59f, e = 0.14f;
            r.x = saturate((c.x * (a * c.x + b)) / (c.x * (cc * c.x + d) + e));
            r.y = saturate((c.y * (a * c.y + b)) / (c.y * (cc * c.y + d) + e));
            r.z = saturate((c.z * (a * c.z + b)) / (c.z * (cc * c.z + d) + e));
        } else {
            const float lum = c.x * 0.299f + c.y * 0.587f + c.z * 0.114f;
            const float rh = lum / (lum + 1.0f);
            r = scl(c, rh / lum);
        }
        out[i * 4 + 0] = r.x;
        out[i * 4 + 1] = r.y;
        out[i * 4 + 2] = r.z;
        out[i * 4 + 3] = 1.0f;
    }
}

/* ------------------------------------------------------------------------ */
/* The shared binary32 transcendentals (strelka_amd/csrc/fmath.h), exposed so */
/* the tests can pin them against binary64 libm.                              */
/* op: 0 sin, 1 cos, 2 acos, 3 pow5                                           */
/* ------------------------------------------------------------------------ */
void orc_fmath(uint32_t op, const float* in, uint32_t n, float* out)
{
    for (uint32_t i = 0; i < n; ++i) {
        switch (op) {
        case 0: out[i] = stk_sinf(in[i]); break;
        case 1: out[i] = stk_cosf(in[i]); break;
        case 2: out[i] = stk_acosf(in[i]); break;
        default: out[i] = stk_pow5f(in[i]); break;
        }
    }
}
