/*
 * oracle.h - CPU oracle for the Strelka path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * shader code (ikryukov/Strelka, shaders/*.h, shaders/*.hlsl) used by the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark to check and
 * time the HIP library.  The product library never links, loads or calls it.
 *
 * Arithmetic rules: IEEE binary32 everywhere, compiled with -ffp-contract=off
 * and without fast-math, every expression evaluated in the order written in the
 * reference (dot = (x*x' + y*y') + z*z'; mat4*vec4 in glm's pairwise order
 * (c0*x + c1*y) + (c2*z + c3*w)).  min/max are IEEE minNum/maxNum (a NaN operand
 * is ignored), which is what the GPU's v_min_f32/v_max_f32 do.
 *
 * Pinning of each part (what checks it beyond restating the formula):
 *   orc_intersect_ray_box  - the seven known-answer cases of the reference's
 *                            (disabled) tests/intersection/aabb.cpp:10-106
 *   orc_ray_triangle       - FP64 Moller-Trumbore / closed-form hits, rejection
 *                            thresholds, parallel rays
 *   orc_closest_hit        - brute force over every leaf (smallest t, then
 *                            smallest preorder leaf) on every config-1 ray
 *   orc_any_hit            - brute-force visibility on the same rays
 *   orc_jenkins/rng        - closed-form integer checks of xorshift32
 *   orc_camera_ray         - closed forms (centre pixel looks down -z)
 *   orc_sample_rect_light  - sample points lie on the rectangle; solid angle S
 *                            against numeric quadrature
 *   orc_bsdf_*             - white furnace (Lambert albedo 1), GGX D
 *                            normalisation, sample/pdf consistency
 *   orc_pt_sample          - FP64 furnace scene closed form (uniform emitter
 *                            enclosure) and per-sample GPU parity
 *   orc_reduce/accumulate/tonemap - mean identity, ACES matrices closed form
 */
#ifndef STK_ORACLE_H
#define STK_ORACLE_H

#include <stdint.h>
#include "../include/stk/stk_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrcScene {
    const StkBvhNode* nodes;
    uint32_t nnodes;
    const StkInstance* inst;
    uint32_t ninst;
    const StkVertex* vb;
    uint32_t nv;
    const uint32_t* ib;
    uint32_t ni;
    const StkLight* lights;
    uint32_t nlights;
    const StkMaterial* mats;
    uint32_t nmats;
    /* albedo textures (optional): texels RGBA32F, 4 floats each */
    const StkTexture* textures;
    uint32_t ntextures;
    const float* texels;
    /* compressed 8-wide layout of the same tree (optional: for ORC_MODE_WIDE) */
    const StkWideNode* wide;
    uint32_t nwide;
    const StkTriangle* wideTris;
    uint32_t nwideTris;
    /* two-level structure (optional: for ORC_MODE_TWO_LEVEL / ORC_MODE_ANY_TWO_LEVEL) */
    const StkWideNode* tlas;
    uint32_t ntlas;
    const StkTlasLeaf* tlasLeaves;
    const StkWideNode* blas;
    const StkBlasTriangle* blasTris;
    const uint32_t* instKey;
    const float* instW; /* 16 per instance: inverse rows (3 x 4), {pad, 0, 0, 0} */
} OrcScene;

/* Traversal counters: inner = inner nodes whose box was tested (reference
 * order) or expanded (pruned order); leaf = triangles tested. */
typedef struct OrcStats {
    uint64_t inner;
    uint64_t leaf;
    uint64_t rays;
    uint64_t hits;
} OrcStats;

/* ---- primitives ---------------------------------------------------------- */
int orc_intersect_ray_box(const StkRay* r, const float invdir[3], const float pmin[3], const float pmax[3],
                          float* t);
int orc_ray_triangle(const float o[3], const float d[3], const float v0[3], const float e0[3], const float e1[3],
                     float* t, float bary[2]);
void orc_transform_point(const float m[16], const float p[3], float out[3]);
void orc_get_triangle(const OrcScene* s, uint32_t inst, uint32_t prim, float v0[3], float e0[3], float e1[3]);

/* ---- traversal ----------------------------------------------------------- */
/* Reference order: stackless skip-link walk exactly as closestHit
 * (shaders/raytracing.h:198-247). Returns 1 on hit. */
int orc_closest_hit(const OrcScene* s, const StkRay* r, StkHit* hit, OrcStats* st);
/* Same result rule, visiting near child first with closest-t pruning; the
 * traversal order and pruning used by the GPU kernel (counts its fetches). */
int orc_closest_hit_pruned(const OrcScene* s, const StkRay* r, StkHit* hit, OrcStats* st);
/* Brute force over every leaf in preorder (no box tests). */
int orc_closest_hit_brute(const OrcScene* s, const StkRay* r, StkHit* hit);
/* anyHit (shaders/raytracing.h:249-286). Returns 1 on hit, fills inst/prim/t. */
int orc_any_hit(const OrcScene* s, const StkRay* r, StkHit* hit, OrcStats* st);
/* The first leaf in preorder (array order) whose triangle is hit with t < tmax. */
int orc_any_hit_brute(const OrcScene* s, const StkRay* r, StkHit* hit);

/* The GPU kernel's walk over the compressed 8-wide nodes (StkWideNode):
 * child boxes decoded as origin + q * 2^(e-127) (one rounding), the slab test
 * of intersectRayBox, leaf triangles re-formed by getTriangle from the leaf's
 * (instance, primitive), nearest entered inner child first and the others
 * pushed as one node group popped in octant order, the same acceptance and
 * pruning rules as orc_closest_hit_pruned.  Counts wide-node visits (inner)
 * and triangle tests (leaf). */
int orc_closest_hit_wide(const OrcScene* s, const StkRay* r, StkHit* hit, OrcStats* st);

/* Expands a two-level structure into one skip-link tree (input of the
 * reference walks): each instance leaf of the top tree `top` (leaf instId =
 * instance) becomes the mesh tree of that instance's mesh (meshNodes from
 * meshStart[m], meshCount[m] nodes; leaf primId in minBounds[0]) with leaves
 * tagged by the instance and inner boxes mapped to world boxes that hold the
 * image of the object box (double precision, widened by 1e-5 relative, rounded
 * outwards).  Skip links follow the preorder of the expanded tree.  Returns the
 * node count, 0 when it exceeds capacity. */
uint32_t orc_expand_two_level(const StkBvhNode* top, uint32_t ntop, const StkBvhNode* meshNodes,
                              const uint32_t* meshStart, const uint32_t* meshCount, const uint32_t* instMesh,
                              const StkInstance* inst, StkBvhNode* out, uint32_t capacity);

/* The GPU kernel's two-level walk (instance tree, then each entered
 * instance's mesh tree on the object-space ray with its widened boxes, the hit
 * test on the world triangle getTriangle forms), accepting by t, then the
 * expanded preorder (instance leaf, then mesh-tree leaf).  Counts wide-node
 * visits of both levels (inner) and triangle tests (leaf). */
int orc_closest_two_level(const OrcScene* s, const StkRay* r, StkHit* hit, OrcStats* st);
/* anyHit's rule on the expanded tree over the two-level walk: the hit triangle
 * first in expanded preorder. */
int orc_any_two_level(const OrcScene* s, const StkRay* r, StkHit* hit, OrcStats* st);

enum { ORC_MODE_REF = 0, ORC_MODE_PRUNED = 1, ORC_MODE_BRUTE = 2, ORC_MODE_ANY = 3, ORC_MODE_WIDE = 4,
       ORC_MODE_TWO_LEVEL = 5, ORC_MODE_ANY_TWO_LEVEL = 6, ORC_MODE_ANY_BRUTE = 7 };
/* Batched traversal over n rays using nthreads host threads (chunks of 4096
 * rays handed out dynamically).  stats may be NULL. */
void orc_trace_batch(const OrcScene* s, const StkRay* rays, uint32_t n, StkHit* hits, int mode, int nthreads,
                     OrcStats* stats);

/* ---- RNG (shaders/random.h:14-47) ---------------------------------------- */
uint32_t orc_jenkins(uint32_t x);
uint32_t orc_init_rng(uint32_t px, uint32_t py, uint32_t width, uint32_t frame);
uint32_t orc_xorshift(uint32_t* state);
float orc_rand(uint32_t* state);
uint32_t orc_permute(uint32_t i, uint32_t l, uint32_t p);
float orc_randfloat(uint32_t i, uint32_t p);

/* ---- camera / lights / BSDF ---------------------------------------------- */
void orc_camera_ray(const StkPathTracerParam* p, uint32_t px, uint32_t py, uint32_t* rng, uint32_t s, StkRay* ray);
/* The sub-pixel position of sample s of stratifiedSamplingType (0 centre,
 * 1 random, 2 stratifiedSampling1, 3 stratifiedSamplingOptimized). */
void orc_pixel_jitter(const StkPathTracerParam* p, uint32_t s, uint32_t* rng, float* jx, float* jy);
void orc_offset_ray(const float p[3], const float n[3], float out[3]);
/* Fills point, L, distToLight, normal, area, pdf of SampleRectLight
 * (shaders/lights.h:284-301) and the solid angle S of the SphQuad. */
void orc_sample_rect_light(const StkLight* l, const float u[2], const float hitPoint[3], float point[3], float L[3],
                           float* dist, float normal[3], float* area, float* pdf, float* solidAngle);

typedef struct OrcBsdfSample {
    float k2[3];
    float pdf;
    float bsdf_over_pdf[3];
    int event; /* 0 absorb, 1 diffuse reflection, 2 glossy reflection, 3 specular reflection, 4 specular
                * transmission (specular samples report pdf 0) */
} OrcBsdfSample;
enum { ORC_EVENT_SPECULAR_REFLECTION = 3, ORC_EVENT_SPECULAR_TRANSMISSION = 4 };
void orc_bsdf_sample(const StkMaterial* m, const float N[3], const float Ng[3], const float k1[3], const float xi[4],
                     OrcBsdfSample* out);
/* diffuse/glossy outputs are bsdf * dot(N, k2), pdf is the solid-angle pdf. */
void orc_bsdf_eval(const StkMaterial* m, const float N[3], const float Ng[3], const float k1[3], const float k2[3],
                   float diffuse[3], float glossy[3], float* pdf);

/* Texture lookup (rgb) at texture coordinates (s, t): the reference's
 * tex_lookup_float4_2d (shaders/mdl_runtime.hlsl:173-207) with its
 * smootherstep filter, the bilinear blend restated in binary32. */
void orc_texture_lookup(const StkTexture* tx, const float* texels, float s, float t, float out[3]);
/* The normal map's shading normal: CalcBumpedNormal (shaders/helper.h:101-115). */
void orc_bumped_normal(const float normal[3], const float tangent[3], const float texel[3], float out[3]);

/* ---- path tracer (shaders/pathtracerMdl.hlsl:204-423) -------------------- */
/* One (pixel, sample) path; seeds the RNG with frameNumber*spp + sampleNum. */
void orc_pt_sample(const OrcScene* s, const StkRenderParams* rp, uint32_t px, uint32_t py, uint32_t sampleNum,
                   float out[3]);
/* All pixels x spp of one frame into samples[(pixel*spp + s)*3 + c]; the
 * layout of the reference sampleBuffer (shaders/pathtracerMdl.hlsl:420-422).
 * Pixels are limited to the rectangle [x0,x0+w) x [y0,y0+h) of the image;
 * the output holds w*h*spp*3 floats in row-major order of that rectangle. */
void orc_pt_frame(const OrcScene* s, const StkRenderParams* rp, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                  float* samples, int nthreads);

/* ---- post passes ---------------------------------------------------------- */
void orc_reduce(const float* samples, uint32_t npix, uint32_t spp, float* rgba);
void orc_accumulate(const float* cur, const float* hist, uint32_t npix, uint32_t iteration, float* out);
void orc_tonemap(const float* in, uint32_t npix, uint32_t type, float* out);

/* The binary64 twin of orc_pt_sample (oracle_fp64.c): the same estimator
 * (refCompat 1: the reference's; 0: the corrected one) and random numbers,
 * every geometric and shading quantity in binary64. */
void orc_pt_sample_fp64(const OrcScene* s, const StkRenderParams* rp, uint32_t px, uint32_t py, uint32_t sampleNum,
                        double out[3]);
/* Per-pixel sums and sums of squares (w x h x 3) of the pixel's rp->pt.spp
 * binary64 samples, threaded over pixels. */
void orc_pt_frame_fp64(const OrcScene* s, const StkRenderParams* rp, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                       double* sum, double* sumsq, int nthreads);
/* Whether a triangle of a non-emitter instance is hit with t < tmax (the
 * corrected estimator's shadow test). */
int orc_occluded(const OrcScene* s, const StkRay* r);

/* The transcendentals both sides share (op 0 sin, 1 cos, 2 acos, 3 x^5). */
void orc_fmath(uint32_t op, const float* in, uint32_t n, float* out);

#ifdef __cplusplus
}
#endif
#endif
