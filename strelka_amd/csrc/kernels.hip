// kernels.hip - HIP kernels for gfx950: traversal, camera rays, the wavefront
// path tracer and the accumulation/tonemap pass.
//
// Wavefront organisation (replaces the per-(pixel,sample) megakernel
// computeMain/pathTraceCameraRays of shaders/pathtracerMdl.hlsl:204-423):
//   raygen -> per bounce { setup, extend (closest hit), shade (NEE sample,
//   BSDF sample, RR; compacts continuing paths and shadow rays into queues with
//   one atomic per wave), connect (any-hit, adds NEE) } -> finalize.
// The per-path sequence of random numbers and of additions to the path
// radiance is the reference's, so a path's value does not depend on the
// wavefront organisation.
#include "kernels.hpp"
#include "pt_dev.hpp"

namespace stk {
namespace {

using namespace dev;

constexpr int kBlock = 256;

__device__ __forceinline__ void store_hit(StkHit* hits, uint32_t i, const HitF& h)
{
    StkHit* o = hits + i;
    o->bary[0] = h.u;
    o->bary[1] = h.v;
    o->t = h.t;
    o->instId = h.inst;
    o->primId = h.prim;
}

__device__ __forceinline__ RayF load_ray(const StkRay* rays, uint32_t i)
{
    const float4* p = reinterpret_cast<const float4*>(rays + i);
    return make_ray(p[0], p[1]);
}

__global__ __launch_bounds__(kBlock) void k_trace_closest(BvhView bvh, const StkRay* __restrict__ rays, uint32_t n,
                                                          StkHit* __restrict__ hits)
{
    __shared__ uint32_t stack[kStackDepth * kBlock];
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const RayF r = load_ray(rays, i);
    HitF h;
    closest_packed(bvh, r, stack + threadIdx.x, kBlock, h);
    store_hit(hits, i, h);
}

__global__ __launch_bounds__(kBlock) void k_trace_closest_ref(BvhView bvh, const StkRay* __restrict__ rays,
                                                              uint32_t n, StkHit* __restrict__ hits)
{
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const RayF r = load_ray(rays, i);
    HitF h;
    closest_ref(bvh, r, h);
    store_hit(hits, i, h);
}

__global__ __launch_bounds__(kBlock) void k_trace_any(BvhView bvh, const StkRay* __restrict__ rays, uint32_t n,
                                                      StkHit* __restrict__ hits)
{
    __shared__ uint32_t stack[kStackDepth * kBlock];
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const RayF r = load_ray(rays, i);
    HitF h;
    any_packed(bvh, r, stack + threadIdx.x, kBlock, h);
    h.u = 0.0f;
    h.v = 0.0f;
    store_hit(hits, i, h);
}

__global__ __launch_bounds__(kBlock) void k_trace_any_ref(BvhView bvh, const StkRay* __restrict__ rays, uint32_t n,
                                                          StkHit* __restrict__ hits)
{
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const RayF r = load_ray(rays, i);
    HitF h;
    any_ref(bvh, r, h);
    h.u = 0.0f;
    h.v = 0.0f;
    store_hit(hits, i, h);
}

__global__ __launch_bounds__(kBlock) void k_camera_rays(StkPathTracerParam p, uint32_t sampleNum,
                                                        StkRay* __restrict__ rays)
{
    const uint32_t W = (uint32_t)p.dimension[0], H = (uint32_t)p.dimension[1];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= W * H) return;
    const uint32_t px = i % W, py = i / W;
    const uint32_t frame = p.frameNumber * (uint32_t)p.spp + sampleNum;
    uint32_t rng = init_rng(px, py, W, frame);
    float4 o, d;
    camera_ray(p, px, py, rng, frame, o, d);
    float4* out = reinterpret_cast<float4*>(rays + i);
    out[0] = o;
    out[1] = d;
}

// ---- wavefront path tracer ------------------------------------------------------

// Appends lanes with pred set to a queue: one atomic per wave, order of lanes
// kept within the wave.  Every active lane of the wave must call it.
__device__ __forceinline__ uint32_t wave_push(uint32_t* counter, bool pred)
{
    const uint64_t mask = __ballot(pred);
    if (mask == 0) return 0;
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = (uint32_t)__shfl((int)base, (int)leader);
    return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

__global__ __launch_bounds__(kBlock) void k_pt_raygen(PtLaunch lp, PathState ps)
{
    const uint32_t W = (uint32_t)lp.pt.dimension[0], H = (uint32_t)lp.pt.dimension[1];
    const uint32_t slot = blockIdx.x * kBlock + threadIdx.x;
    if (slot == 0) {
        ps.counters[0] = W * H;
        for (int k = 1; k < 8; ++k) ps.counters[k] = 0;
    }
    if (slot >= W * H) return;
    const uint32_t px = slot % W, py = slot / W;
    const uint32_t frame = lp.pt.frameNumber * (uint32_t)lp.pt.spp + lp.sampleNum;
    uint32_t rng = init_rng(px, py, W, frame);
    float4 o, d;
    camera_ray(lp.pt, px, py, rng, frame, o, d);
    ps.rayO[slot] = o;
    ps.rayD[slot] = d;
    ps.thr[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    ps.col[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    ps.rng[slot] = rng;
    ps.queue[0][slot] = slot;
    const PtCache& pc = lp.cache;
    if (pc.mode != 0 && pc.mode != 2 && slot % pc.trainStride == pc.trainOffset)
        pc.trainCount[slot / pc.trainStride] = 0;
}

__global__ void k_pt_setup(PathState ps, int depth)
{
    if (threadIdx.x == 0) {
        ps.counters[1 - (depth & 1)] = 0;
        ps.counters[2] = 0;
    }
}

__global__ __launch_bounds__(kBlock) void k_pt_extend(SceneView sc, PathState ps, int depth)
{
    __shared__ uint32_t stack[kStackDepth * kBlock];
    const int cur = depth & 1;
    const uint32_t count = ps.counters[cur];
    const uint32_t q = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (q >= count) return;
    const uint32_t slot = ps.queue[cur][q];
    const RayF r = make_ray(ps.rayO[slot], ps.rayD[slot]);
    HitF h;
    closest_packed(sc.bvh, r, stack + threadIdx.x, kBlock, h);
    ps.hit[q] = make_float4(h.u, h.v, h.t, bits_to_float(h.inst));
    ps.hitPrim[q] = h.prim;
}

__device__ __forceinline__ float4 f4(F3 v) { return make_float4(v.x, v.y, v.z, 0.0f); }
__device__ __forceinline__ F3 f3(float4 v) { return mk3(v.x, v.y, v.z); }

__global__ __launch_bounds__(kBlock) void k_pt_shade(SceneView sc, PtLaunch lp, PathState ps, int depth)
{
    const int cur = depth & 1;
    const uint32_t count = ps.counters[cur];
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    bool cont = false, shadow = false, cacheQuery = false;
    uint32_t slot = 0;
    F3 nO = mk3(0, 0, 0), nD = mk3(0, 0, 0);
    F3 sO = mk3(0, 0, 0), sD = mk3(0, 0, 0), cd = mk3(0, 0, 0), cg = mk3(0, 0, 0);
    float sT = 0.0f;
    CacheRecord qrec;
    const PtCache& pc = lp.cache;
    if (q < count) {
        slot = ps.queue[cur][q];
        const float4 h = ps.hit[q];
        const uint32_t inst = float_to_bits(h.w);
        F3 thr = f3(ps.thr[slot]);
        F3 col = f3(ps.col[slot]);
        uint32_t rng = ps.rng[slot];
        if (inst == STK_INVALID_INDEX) {
            col = col + thr * mk3(lp.sky[0], lp.sky[1], lp.sky[2]);
            ps.col[slot] = f4(col);
        } else {
            const StkInstance& ic = sc.inst[inst];
            if (ic.lightId != -1) {
                // emitter: counted on camera rays (shaders/pathtracerMdl.hlsl:229-234);
                // a BSDF-sampled emitter hit adds nothing and ends the path
                if (depth == 0 && (uint32_t)ic.lightId < sc.numLights) {
                    col = col + thr * ld3(sc.lights[ic.lightId].color);
                    ps.col[slot] = f4(col);
                }
            } else {
                const uint32_t prim = ps.hitPrim[q];
                const uint32_t base = (uint32_t)ic.indexOffset + prim * 3;
                const uint32_t i0 = sc.ib[base], i1 = sc.ib[base + 1], i2 = sc.ib[base + 2];
                const StkVertex &v0 = sc.vb[i0], &v1 = sc.vb[i1], &v2 = sc.vb[i2];
                const F3 p0 = xform4(ic.objectToWorld, ld3(v0.position), 1.0f);
                const F3 p1 = xform4(ic.objectToWorld, ld3(v1.position), 1.0f);
                const F3 p2 = xform4(ic.objectToWorld, ld3(v2.position), 1.0f);
                const F3 Ng = nrm(cross(p1 - p0, p2 - p0));
                const F3 n0 = xform3(ic.normalMatrix, unpack_normal(v0.normal));
                const F3 n1 = xform3(ic.normalMatrix, unpack_normal(v1.normal));
                const F3 n2 = xform3(ic.normalMatrix, unpack_normal(v2.normal));
                const F3 P = interp(p0, p1, p2, h.x, h.y);
                const F3 N = nrm(interp(n0, n1, n2, h.x, h.y));
                StkMaterial mat;
                if (ic.materialId >= 0 && (uint32_t)ic.materialId < sc.numMats) {
                    mat = sc.mats[ic.materialId];
                } else {
                    mat.type = STK_MAT_LAMBERT;
                    mat.roughness = 1.0f;
                    mat.albedo[0] = mat.albedo[1] = mat.albedo[2] = 0.5f;
                    mat.albedo[3] = 1.0f;
                }
                const F3 k1 = -f3(ps.rayD[slot]);
                const float rough = mat.type == STK_MAT_GGX ? mat.roughness : 1.0f;
                const bool isTrain = pc.mode != 0 && pc.mode != 2 && slot % pc.trainStride == pc.trainOffset;
                if ((pc.mode == 1 || pc.mode == 2) && !isTrain && depth == (int)pc.queryDepth) {
                    // the path ends here: its remaining radiance comes from the cache
                    cacheQuery = true;
                    qrec.posRough = make_float4(P.x, P.y, P.z, rough);
                    qrec.normal = make_float4(N.x, N.y, N.z, bits_to_float(slot));
                    qrec.dir = make_float4(k1.x, k1.y, k1.z, 0.0f);
                    qrec.target = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                } else if (isTrain) {
                    float4* v = pc.trainVerts + ((size_t)(slot / pc.trainStride) * lp.pt.maxDepth + depth) * 5;
                    v[0] = make_float4(P.x, P.y, P.z, rough);
                    v[1] = make_float4(N.x, N.y, N.z, 0.0f);
                    v[2] = make_float4(k1.x, k1.y, k1.z, 0.0f);
                    v[3] = f4(col);
                    v[4] = f4(thr);
                    pc.trainCount[slot / pc.trainStride] = (uint32_t)depth + 1;
                }
                // sampleLights / estimateDirectLighting (shaders/pathtracerMdl.hlsl:31-92)
                F3 toLight = mk3(0, 0, 0), radiance = mk3(0, 0, 0);
                float lightPdf = 0.0f, dist = 0.0f;
                if (sc.numLights > 0 && !cacheQuery) {
                    const uint32_t lid = (uint32_t)((float)sc.numLights * rnd(rng));
                    const float selPdf = 1.0f / (float)sc.numLights;
                    const StkLight& l = sc.lights[lid < sc.numLights ? lid : sc.numLights - 1];
                    const float u0 = rnd(rng);
                    const float u1 = rnd(rng);
                    const LightSample ls = sample_rect_light(l, u0, u1, P);
                    toLight = ls.L;
                    const F3 Li = ld3(l.color);
                    if (dot(N, toLight) > 0.0f && -dot(toLight, ls.normal) > 0.0f && Li.x != 0.0f && Li.y != 0.0f &&
                        Li.z != 0.0f) {
                        lightPdf = ls.pdf;
                        dist = ls.dist;
                        radiance = (Li * 1.0f) * saturate(dot(N, toLight)); // visibility applied by connect
                    }
                    lightPdf *= selPdf;
                }
                bool alive = !(isnan(radiance.x) || isnan(radiance.y) || isnan(radiance.z) || isnan(lightPdf));
                if (alive && (dot(toLight, Ng) > 0.0f) && lightPdf != 0.0f) {
                    const F3 rop = mk3(radiance.x / lightPdf, radiance.y / lightPdf, radiance.z / lightPdf);
                    F3 bd, bg;
                    float bpdf;
                    bsdf_eval(mat, N, Ng, k1, toLight, bd, bg, bpdf);
                    if (isnan(bd.x) || isnan(bd.y) || isnan(bd.z) || isnan(bg.x) || isnan(bg.y) || isnan(bg.z)) {
                        alive = false;
                    } else if (bpdf > 0.0f) {
                        const float mis = (lightPdf == 0.0f) ? 1.0f : lightPdf / (lightPdf + bpdf);
                        const F3 w = (thr * rop) * mis;
                        cd = w * bd;
                        cg = w * bg;
                        shadow = true;
                        sO = offset_ray(P, Ng);
                        sD = toLight;
                        sT = dist;
                    }
                }
                if (alive && !cacheQuery) {
                    const float x0 = rnd(rng);
                    const float x1 = rnd(rng);
                    (void)rnd(rng);
                    (void)rnd(rng);
                    const BsdfSample bs = bsdf_sample(mat, N, Ng, k1, x0, x1);
                    if (bs.event != 0) {
                        thr = thr * bs.weight;
                        bool survive = true;
                        if (depth > (int)lp.rrStartDepth) {
                            const float pr = fmaxf(thr.x, fmaxf(thr.y, thr.z));
                            if (rnd(rng) > pr)
                                survive = false;
                            else
                                thr = thr * (1.0f / (pr + 1e-5f));
                        }
                        if (survive) {
                            cont = true;
                            nO = offset_ray(P, Ng);
                            nD = bs.k2;
                        }
                    }
                }
            }
        }
        ps.rng[slot] = rng;
        if (cont) {
            ps.thr[slot] = f4(thr);
            ps.rayO[slot] = make_float4(nO.x, nO.y, nO.z, 1e9f);
            ps.rayD[slot] = make_float4(nD.x, nD.y, nD.z, 0.0f);
        }
    }
    const uint32_t qi = wave_push(&ps.counters[1 - cur], cont);
    if (cont) ps.queue[1 - cur][qi] = slot;
    const uint32_t si = wave_push(&ps.counters[2], shadow);
    if (shadow) {
        ps.shadowQ[si] = slot;
        ps.shO[si] = make_float4(sO.x, sO.y, sO.z, sT);
        ps.shD[si] = make_float4(sD.x, sD.y, sD.z, 0.0f);
        ps.shCd[si] = f4(cd);
        ps.shCg[si] = f4(cg);
    }
    if (pc.mode == 1 || pc.mode == 2) {
        const uint32_t ci = wave_push(&ps.counters[3], cacheQuery);
        if (cacheQuery) pc.queryRecs[ci] = qrec;
    }
}

// Adds the cache's answer to the paths that ended at a query:
// radiance += throughput * max(prediction, 0).
__global__ __launch_bounds__(kBlock) void k_cache_apply(PtLaunch lp, PathState ps)
{
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= ps.counters[3]) return;
    const uint32_t slot = float_to_bits(lp.cache.queryRecs[q].normal.w);
    const float4 o = lp.cache.queryOut[q];
    const F3 thr = f3(ps.thr[slot]);
    const F3 col = f3(ps.col[slot]) + thr * mk3(fmaxf(o.x, 0.0f), fmaxf(o.y, 0.0f), fmaxf(o.z, 0.0f));
    ps.col[slot] = f4(col);
}

// Training targets: the radiance a training path gathered from vertex d on,
// divided by the throughput at d - a one-sample estimate of the outgoing
// radiance at d under the same estimator as the plain path tracer.
__global__ __launch_bounds__(kBlock) void k_cache_targets(PtLaunch lp, PathState ps)
{
    const PtCache& pc = lp.cache;
    const uint32_t maxDepth = lp.pt.maxDepth;
    const uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t tidx = id / maxDepth, d = id % maxDepth;
    bool emit = false;
    CacheRecord rec;
    if (tidx < pc.numTrainPaths && d < pc.trainCount[tidx]) {
        const uint32_t slot = tidx * pc.trainStride + pc.trainOffset;
        const float4* v = pc.trainVerts + ((size_t)tidx * maxDepth + d) * 5;
        const float4 before = v[3], thr = v[4], fin = ps.col[slot];
        if (thr.x > 1e-8f && thr.y > 1e-8f && thr.z > 1e-8f) {
            rec.posRough = v[0];
            rec.normal = v[1];
            rec.dir = v[2];
            // one-sample targets are clamped to [0, kTargetMax]: Russian-roulette
            // survivors and near-light NEE samples can be arbitrarily large, and
            // a few such fireflies would dominate the loss
            const float kTargetMax = 100.0f;
            rec.target = make_float4(fminf(fmaxf((fin.x - before.x) / thr.x, 0.0f), kTargetMax),
                                     fminf(fmaxf((fin.y - before.y) / thr.y, 0.0f), kTargetMax),
                                     fminf(fmaxf((fin.z - before.z) / thr.z, 0.0f), kTargetMax), 0.0f);
            emit = isfinite(fin.x) && isfinite(fin.y) && isfinite(fin.z);
        }
    }
    const uint32_t i = wave_push(&ps.counters[4], emit);
    if (emit && i < pc.trainCap) pc.trainBatch[i] = rec;
}

// Shadow rays use the stackless skip-link walk: for these long incoherent
// segments it beats the packed left-first walk (no LDS stack, so 8 waves per
// SIMD instead of 5) while returning the same first-in-preorder hit.
__global__ __launch_bounds__(kBlock) void k_pt_connect(SceneView sc, PathState ps)
{
    const uint32_t count = ps.counters[2];
    const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (s >= count) return;
    const RayF r = make_ray(ps.shO[s], ps.shD[s]);
    HitF h;
    bool vis = !any_ref(sc.bvh, r, h);
    if (!vis && sc.inst[h.inst].lightId != -1) vis = true; // a light hit counts as visible
    if (vis) {
        const uint32_t slot = ps.shadowQ[s];
        F3 col = f3(ps.col[slot]);
        col = col + f3(ps.shCd[s]);
        col = col + f3(ps.shCg[s]);
        ps.col[slot] = f4(col);
    }
}

__global__ __launch_bounds__(kBlock) void k_pt_finalize(PtLaunch lp, PathState ps, float4* __restrict__ accum,
                                                        float* __restrict__ samples, uint32_t npix)
{
    const uint32_t slot = blockIdx.x * kBlock + threadIdx.x;
    if (slot >= npix) return;
    const float4 c = ps.col[slot];
    float4 a = accum[slot];
    // reduction.hlsl:11-32: color += invSpp * sample, in sample order
    a.x = a.x + lp.invSpp * c.x;
    a.y = a.y + lp.invSpp * c.y;
    a.z = a.z + lp.invSpp * c.z;
    a.w = 1.0f;
    accum[slot] = a;
    if (samples) {
        float* o = samples + ((size_t)slot * (uint32_t)lp.pt.spp + lp.sampleNum) * 3;
        o[0] = c.x;
        o[1] = c.y;
        o[2] = c.z;
    }
}

__global__ __launch_bounds__(kBlock) void k_copy(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n)
{
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__global__ __launch_bounds__(kBlock) void k_zero(float4* p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// tonemap.hlsl:9-88
__device__ __forceinline__ float rrt_odt(float v)
{
    const float a = v * (v + 0.0245786f) - 0.000090537f;
    const float b = v * (0.983729f * v + 0.4329510f) + 0.238081f;
    return a / b;
}

__device__ __forceinline__ F3 tonemap(F3 c, uint32_t type)
{
    if (type == 1) {
        F3 x = mk3(0.59719f * c.x + 0.35458f * c.y + 0.04823f * c.z, 0.07600f * c.x + 0.90834f * c.y + 0.01566f * c.z,
                   0.02840f * c.x + 0.13383f * c.y + 0.83777f * c.z);
        x = mk3(rrt_odt(x.x), rrt_odt(x.y), rrt_odt(x.z));
        x = mk3(1.60475f * x.x + -0.53108f * x.y + -0.07367f * x.z, -0.10208f * x.x + 1.10813f * x.y + -0.00605f * x.z,
                -0.00327f * x.x + -0.07276f * x.y + 1.07602f * x.z);
        return mk3(saturate(x.x), saturate(x.y), saturate(x.z));
    }
    if (type == 2) {
        const float a = 2.51f, b = 0.03f, cc = 2.43f, d = 0.59f, e = 0.14f;
        return mk3(saturate((c.x * (a * c.x + b)) / (c.x * (cc * c.x + d) + e)),
                   saturate((c.y * (a * c.y + b)) / (c.y * (cc * c.y + d) + e)),
                   saturate((c.z * (a * c.z + b)) / (c.z * (cc * c.z + d) + e)));
    }
    const float lum = c.x * 0.299f + c.y * 0.587f + c.z * 0.114f;
    const float rh = lum / (lum + 1.0f);
    return c * (rh / lum);
}

__global__ __launch_bounds__(kBlock) void k_accumulate_tonemap(const float4* __restrict__ cur,
                                                               float4* __restrict__ history, float4* __restrict__ out,
                                                               uint32_t npix, uint32_t iteration, int accumulate,
                                                               int doTonemap, uint32_t type)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= npix) return;
    F3 res = f3(cur[i]);
    if (accumulate) {
        // accPt (shaders/accumulation.hlsl:86-97)
        const float wHistory = (float)iteration / (float)(iteration + 1);
        const float wNew = 1.0f / (float)(iteration + 1);
        const F3 prev = f3(history[i]);
        res = prev * wHistory + res * wNew;
        history[i] = make_float4(res.x, res.y, res.z, 1.0f);
    }
    const F3 o = doTonemap ? tonemap(res, type) : res;
    out[i] = make_float4(o.x, o.y, o.z, 1.0f);
}

inline uint32_t blocks(size_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

} // namespace

hipError_t launchTraceClosest(const BvhView& bvh, const StkRay* rays, uint32_t n, StkHit* hits, bool refOrder,
                              hipStream_t s)
{
    if (n == 0) return hipSuccess;
    if (refOrder)
        hipLaunchKernelGGL(k_trace_closest_ref, dim3(blocks(n)), dim3(kBlock), 0, s, bvh, rays, n, hits);
    else
        hipLaunchKernelGGL(k_trace_closest, dim3(blocks(n)), dim3(kBlock), 0, s, bvh, rays, n, hits);
    return hipGetLastError();
}

hipError_t launchTraceAny(const BvhView& bvh, const StkRay* rays, uint32_t n, StkHit* hits, bool refOrder,
                          hipStream_t s)
{
    if (n == 0) return hipSuccess;
    if (refOrder)
        hipLaunchKernelGGL(k_trace_any_ref, dim3(blocks(n)), dim3(kBlock), 0, s, bvh, rays, n, hits);
    else
        hipLaunchKernelGGL(k_trace_any, dim3(blocks(n)), dim3(kBlock), 0, s, bvh, rays, n, hits);
    return hipGetLastError();
}

hipError_t launchCameraRays(const StkPathTracerParam& p, uint32_t sampleNum, StkRay* rays, hipStream_t s)
{
    const size_t n = (size_t)p.dimension[0] * (size_t)p.dimension[1];
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_camera_rays, dim3(blocks(n)), dim3(kBlock), 0, s, p, sampleNum, rays);
    return hipGetLastError();
}

hipError_t launchPathTracePass(const SceneView& sc, const PtLaunch& lp, PathState& ps, float4* accum,
                               float* samples, CacheBuffers* cb, hipStream_t s)
{
    const uint32_t npix = (uint32_t)lp.pt.dimension[0] * (uint32_t)lp.pt.dimension[1];
    if (npix == 0) return hipSuccess;
    const dim3 g(blocks(npix)), b(kBlock);
    const PtCache& pc = lp.cache;
    const bool query = pc.mode == 1 || pc.mode == 2, train = pc.mode == 1 || pc.mode == 3;
    // with a cache query at depth q, non-training paths end at q; training
    // paths still run to maxDepth
    hipLaunchKernelGGL(k_pt_raygen, g, b, 0, s, lp, ps);
    for (int depth = 0; depth < (int)lp.pt.maxDepth; ++depth) {
        hipLaunchKernelGGL(k_pt_setup, dim3(1), dim3(64), 0, s, ps, depth);
        hipLaunchKernelGGL(k_pt_extend, g, b, 0, s, sc, ps, depth);
        hipLaunchKernelGGL(k_pt_shade, g, b, 0, s, sc, lp, ps, depth);
        hipLaunchKernelGGL(k_pt_connect, g, b, 0, s, sc, ps);
    }
    if (query && cb) {
        hipError_t e = cacheInfer(*cb, pc.queryRecs, npix, ps.counters + 3, pc.queryOut, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_cache_apply, g, b, 0, s, lp, ps);
    }
    hipLaunchKernelGGL(k_pt_finalize, g, b, 0, s, lp, ps, accum, samples, npix);
    if (train && cb && pc.numTrainPaths > 0) {
        hipLaunchKernelGGL(k_cache_targets, dim3(blocks((size_t)pc.numTrainPaths * lp.pt.maxDepth)), b, 0, s, lp,
                           ps);
        hipError_t e = cacheBackward(*cb, pc.trainBatch, pc.trainCap, ps.counters + 4, s);
        if (e != hipSuccess) return e;
        e = cacheAdam(*cb, 1.0f, ps.counters + 4, pc.trainCap, s);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t launchCopy(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    const size_t n = bytes / 16;
    if (n == 0) return hipSuccess;
    const uint32_t g = (uint32_t)std::min<size_t>(blocks(n), 8192);
    hipLaunchKernelGGL(k_copy, dim3(g), dim3(kBlock), 0, s, static_cast<uint4*>(dst), static_cast<const uint4*>(src),
                       n);
    return hipGetLastError();
}

hipError_t launchZero(float4* p, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero, dim3(blocks(n)), dim3(kBlock), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launchAccumulateTonemap(const float4* cur, float4* history, float4* out, uint32_t npix,
                                   uint32_t iteration, int accumulate, int doTonemap, uint32_t tonemapType,
                                   hipStream_t s)
{
    if (npix == 0) return hipSuccess;
    hipLaunchKernelGGL(k_accumulate_tonemap, dim3(blocks(npix)), dim3(kBlock), 0, s, cur, history, out, npix,
                       iteration, accumulate, doTonemap, tonemapType);
    return hipGetLastError();
}

} // namespace stk
