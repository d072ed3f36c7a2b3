// context.hip - the C ABI (include/stk/stk.h): context, device buffers, BVH
// upload, traversal and render entry points.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include <cmath>

#include "../../include/stk/stk.h"
#include "bvh.hpp"
#include "cache.hpp"
#include "kernels.hpp"

namespace {

// A device allocation owned by a context.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

bool failAllocForTest()
{
    const char* e = std::getenv("STK_TEST_FAIL_DEVICE_ALLOC");
    return e && e[0] == '1';
}

} // namespace

struct stk_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    hipEvent_t events[16] = {};

    std::vector<StkVertex> hvb;
    std::vector<uint32_t> hib;
    std::vector<StkInstance> hinst;
    uint32_t numLights = 0, numMats = 0;
    bool haveGeometry = false;

    stk::Bvh bvh;
    bool haveBvh = false;

    DevBuf vb, ib, inst, lights, mats, packed, tris, refNodes;
    DevBuf rayScratch, hitScratch;

    // render state
    stk::PathState ps;
    std::vector<DevBuf> psBufs;
    DevBuf accum, history, out, samples;
    uint32_t width = 0, height = 0, iteration = 0;
    size_t samplesCount = 0;

    // radiance cache
    stk::CacheBuffers cache;
    bool haveCache = false;
    StkCacheConfig cacheCfg{};
    DevBuf cGridH, cGridP, cGridM, cGridV, cGridG, cMlpP, cMlpM, cMlpV, cMlpG, cFwd, cBwd, cXT, cDyT, cLoss;
    DevBuf cGridE, cGridHE, cMlpE, cFwdE, cBwdE;
    DevBuf queryRecs, queryOut, trainVerts, trainCount, trainBatch;

    std::vector<DevBuf*> owned()
    {
        return {&vb, &ib, &inst, &lights, &mats, &packed, &tris, &refNodes, &rayScratch, &hitScratch,
                &accum, &history, &out, &samples, &cGridH, &cGridP, &cGridM, &cGridV, &cGridG, &cMlpP,
                &cMlpM, &cMlpV, &cMlpG, &cFwd, &cBwd, &cXT, &cDyT, &cLoss, &cGridE, &cGridHE, &cMlpE,
                &cFwdE, &cBwdE, &queryRecs, &queryOut, &trainVerts, &trainCount, &trainBatch};
    }
};

namespace {

stk_status fail(stk_ctx* c, stk_status st, const std::string& msg)
{
    if (c) c->err = msg;
    return st;
}

stk_status hipFail(stk_ctx* c, hipError_t e, const char* what)
{
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    if (c) c->err = m;
    return (e == hipErrorOutOfMemory) ? STK_E_OOM : STK_E_DEVICE;
}

#define STK_HIP(c, call)                                                                                              \
    do {                                                                                                              \
        hipError_t e_ = (call);                                                                                       \
        if (e_ != hipSuccess) return hipFail((c), e_, #call);                                                         \
    } while (0)

stk_status devRelease(stk_ctx* c, DevBuf& b)
{
    if (b.ptr) {
        hipError_t e = hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
        if (e != hipSuccess) return hipFail(c, e, "hipFree");
    }
    return STK_OK;
}

stk_status devReserve(stk_ctx* c, DevBuf& b, size_t bytes)
{
    if (b.bytes >= bytes && b.ptr) return STK_OK;
    stk_status st = devRelease(c, b);
    if (st != STK_OK) return st;
    if (bytes == 0) return STK_OK;
    if (failAllocForTest()) return fail(c, STK_E_OOM, "device allocation failed (test hook)");
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        (void)hipGetLastError();
        return hipFail(c, e, "hipMalloc");
    }
    b.bytes = bytes;
    return STK_OK;
}

stk_status devUpload(stk_ctx* c, DevBuf& b, const void* src, size_t bytes)
{
    stk_status st = devReserve(c, b, bytes);
    if (st != STK_OK || bytes == 0) return st;
    STK_HIP(c, hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    return STK_OK;
}

stk::BvhView bvhView(stk_ctx* c)
{
    stk::BvhView v;
    v.packed = static_cast<const float4*>(c->packed.ptr);
    v.tris = static_cast<const float4*>(c->tris.ptr);
    v.refNodes = static_cast<const StkBvhNode*>(c->refNodes.ptr);
    v.instances = static_cast<const StkInstance*>(c->inst.ptr);
    v.hdr = c->bvh.header;
    return v;
}

stk_status uploadBvh(stk_ctx* c)
{
    stk::Bvh& b = c->bvh;
    // device copy of the skip-link array: leaves carry their triangle index in maxBounds[0]
    std::vector<StkBvhNode> dev(b.nodes);
    for (size_t i = 0; i < dev.size(); ++i)
        if ((uint32_t)dev[i].instId != STK_INVALID_INDEX) dev[i].maxBounds[0] = stk::bits_to_float(b.leafToTri[i]);
    stk_status st;
    if ((st = devUpload(c, c->refNodes, dev.data(), dev.size() * sizeof(StkBvhNode))) != STK_OK) return st;
    if ((st = devUpload(c, c->packed, b.packed.data(), b.packed.size() * sizeof(StkPackedNode))) != STK_OK) return st;
    if ((st = devUpload(c, c->tris, b.tris.data(), b.tris.size() * sizeof(StkTriangle))) != STK_OK) return st;
    c->haveBvh = true;
    return STK_OK;
}

bool validCtx(stk_ctx* c) { return c != nullptr; }

} // namespace

extern "C" {

STK_API const char* stk_version(void) { return "strelka_amd 0.1 (gfx950)"; }

STK_API stk_status stk_create(int hipDevice, stk_ctx** out)
{
    if (!out) return STK_E_INVALID;
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return STK_E_DEVICE;
    }
    if (hipDevice < 0 || hipDevice >= n) return STK_E_INVALID;
    std::unique_ptr<stk_ctx> c(new (std::nothrow) stk_ctx());
    if (!c) return STK_E_OOM;
    c->device = hipDevice;
    if (hipSetDevice(hipDevice) != hipSuccess) return STK_E_DEVICE;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return STK_E_DEVICE;
    for (auto& ev : c->events)
        if (hipEventCreate(&ev) != hipSuccess) return STK_E_DEVICE;
    *out = c.release();
    return STK_OK;
}

STK_API void stk_destroy(stk_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (DevBuf* b : c->owned()) devRelease(c, *b);
    for (DevBuf& b : c->psBufs) devRelease(c, b);
    for (auto& ev : c->events)
        if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

STK_API const char* stk_last_error(const stk_ctx* c) { return c ? c->err.c_str() : "null context"; }

STK_API stk_status stk_sync(stk_ctx* c)
{
    if (!validCtx(c)) return STK_E_INVALID;
    STK_HIP(c, hipStreamSynchronize(c->stream));
    return STK_OK;
}

STK_API void* stk_stream(stk_ctx* c) { return c ? (void*)c->stream : nullptr; }

STK_API stk_status stk_set_geometry(stk_ctx* c, const StkVertex* vb, uint32_t nv, const uint32_t* ib, uint32_t ni,
                                    const StkInstance* inst, uint32_t ninst)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if ((nv && !vb) || (ni && !ib) || (ninst && !inst) || ni % 3 != 0)
        return fail(c, STK_E_INVALID, "stk_set_geometry: null buffer or index count not a multiple of 3");
    for (uint32_t i = 0; i < ninst; ++i) {
        const StkInstance& ic = inst[i];
        if (ic.indexOffset < 0 || ic.indexCount < 0 || ic.indexCount % 3 ||
            (uint64_t)ic.indexOffset + (uint64_t)ic.indexCount > ni)
            return fail(c, STK_E_INVALID, "stk_set_geometry: instance index range outside the index buffer");
    }
    for (uint32_t i = 0; i < ni; ++i)
        if (ib[i] >= nv) return fail(c, STK_E_INVALID, "stk_set_geometry: index outside the vertex buffer");
    try {
        c->hvb.assign(vb, vb + nv);
        c->hib.assign(ib, ib + ni);
        c->hinst.assign(inst, inst + ninst);
    } catch (const std::bad_alloc&) {
        return fail(c, STK_E_OOM, "host allocation failed");
    }
    stk_status st;
    (void)hipSetDevice(c->device);
    if ((st = devUpload(c, c->vb, vb, (size_t)nv * sizeof(StkVertex))) != STK_OK) return st;
    if ((st = devUpload(c, c->ib, ib, (size_t)ni * sizeof(uint32_t))) != STK_OK) return st;
    if ((st = devUpload(c, c->inst, inst, (size_t)ninst * sizeof(StkInstance))) != STK_OK) return st;
    c->haveGeometry = true;
    c->haveBvh = false;
    return STK_OK;
}

STK_API stk_status stk_set_lights(stk_ctx* c, const StkLight* l, uint32_t n)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (n && !l) return fail(c, STK_E_INVALID, "stk_set_lights: null buffer");
    for (uint32_t i = 0; i < n; ++i)
        if (l[i].type != 0) return fail(c, STK_E_INVALID, "stk_set_lights: only rect lights (type 0) are sampled");
    (void)hipSetDevice(c->device);
    stk_status st = devUpload(c, c->lights, l, (size_t)n * sizeof(StkLight));
    if (st == STK_OK) c->numLights = n;
    return st;
}

STK_API stk_status stk_set_materials(stk_ctx* c, const StkMaterial* m, uint32_t n)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (n && !m) return fail(c, STK_E_INVALID, "stk_set_materials: null buffer");
    (void)hipSetDevice(c->device);
    stk_status st = devUpload(c, c->mats, m, (size_t)n * sizeof(StkMaterial));
    if (st == STK_OK) c->numMats = n;
    return st;
}

STK_API stk_status stk_build_bvh(stk_ctx* c, const StkBvhOptions* opts)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveGeometry) return fail(c, STK_E_STATE, "stk_build_bvh: no geometry set");
    std::vector<stk::BvhInputTri> tris;
    std::string err;
    const StkBvhOptions o = opts ? *opts : stk::defaultBvhOptions();
    try {
        if (!stk::flattenInstances(c->hvb.data(), (uint32_t)c->hvb.size(), c->hib.data(), (uint32_t)c->hib.size(),
                                   c->hinst.data(), (uint32_t)c->hinst.size(), tris, err) ||
            !stk::buildBvh(tris, o, c->bvh, err) ||
            !stk::packBvh(c->hvb.data(), (uint32_t)c->hvb.size(), c->hib.data(), (uint32_t)c->hib.size(),
                          c->hinst.data(), (uint32_t)c->hinst.size(), c->bvh, err))
            return fail(c, STK_E_INVALID, "stk_build_bvh: " + err);
    } catch (const std::bad_alloc&) {
        return fail(c, STK_E_OOM, "stk_build_bvh: host allocation failed");
    }
    if (c->bvh.depth > (uint32_t)stk::dev::kStackDepth)
        return fail(c, STK_E_INVALID, "stk_build_bvh: tree deeper than the traversal stack (32)");
    (void)hipSetDevice(c->device);
    return uploadBvh(c);
}

STK_API stk_status stk_set_bvh(stk_ctx* c, const StkBvhNode* nodes, uint32_t n)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveGeometry) return fail(c, STK_E_STATE, "stk_set_bvh: no geometry set");
    if (n && !nodes) return fail(c, STK_E_INVALID, "stk_set_bvh: null buffer");
    std::string err;
    try {
        c->bvh = stk::Bvh{};
        c->bvh.nodes.assign(nodes, nodes + n);
        if (!stk::packBvh(c->hvb.data(), (uint32_t)c->hvb.size(), c->hib.data(), (uint32_t)c->hib.size(),
                          c->hinst.data(), (uint32_t)c->hinst.size(), c->bvh, err))
            return fail(c, STK_E_INVALID, "stk_set_bvh: " + err);
    } catch (const std::bad_alloc&) {
        return fail(c, STK_E_OOM, "stk_set_bvh: host allocation failed");
    }
    if (c->bvh.depth > (uint32_t)stk::dev::kStackDepth)
        return fail(c, STK_E_INVALID, "stk_set_bvh: tree deeper than the traversal stack (32)");
    (void)hipSetDevice(c->device);
    return uploadBvh(c);
}

STK_API stk_status stk_get_bvh(stk_ctx* c, StkBvhNode* nodes, uint32_t cap, uint32_t* n)
{
    if (!validCtx(c) || !n) return STK_E_INVALID;
    if (!c->haveBvh) return fail(c, STK_E_STATE, "stk_get_bvh: no BVH");
    *n = (uint32_t)c->bvh.nodes.size();
    if (!nodes) return STK_OK;
    if (cap < *n) return fail(c, STK_E_INVALID, "stk_get_bvh: capacity too small");
    memcpy(nodes, c->bvh.nodes.data(), c->bvh.nodes.size() * sizeof(StkBvhNode));
    return STK_OK;
}

STK_API stk_status stk_trace_closest_device(stk_ctx* c, const StkRay* rays, uint32_t n, StkHit* hits, uint32_t flags)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveBvh) return fail(c, STK_E_STATE, "stk_trace_closest: no BVH");
    if (n && (!rays || !hits)) return fail(c, STK_E_INVALID, "stk_trace_closest: null buffer");
    (void)hipSetDevice(c->device);
    STK_HIP(c, stk::launchTraceClosest(bvhView(c), rays, n, hits, (flags & STK_TRACE_REFERENCE_ORDER) != 0,
                                       c->stream));
    return STK_OK;
}

STK_API stk_status stk_trace_any_device(stk_ctx* c, const StkRay* rays, uint32_t n, StkHit* hits, uint32_t flags)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveBvh) return fail(c, STK_E_STATE, "stk_trace_any: no BVH");
    if (n && (!rays || !hits)) return fail(c, STK_E_INVALID, "stk_trace_any: null buffer");
    (void)hipSetDevice(c->device);
    STK_HIP(c, stk::launchTraceAny(bvhView(c), rays, n, hits, (flags & STK_TRACE_REFERENCE_ORDER) != 0, c->stream));
    return STK_OK;
}

static stk_status traceHost(stk_ctx* c, const StkRay* rays, uint32_t n, StkHit* hits, int any, uint32_t flags)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveBvh) return fail(c, STK_E_STATE, "trace: no BVH");
    if (n == 0) return STK_OK;
    if (!rays || !hits) return fail(c, STK_E_INVALID, "trace: null buffer");
    (void)hipSetDevice(c->device);
    stk_status st;
    if ((st = devReserve(c, c->rayScratch, (size_t)n * sizeof(StkRay))) != STK_OK) return st;
    if ((st = devReserve(c, c->hitScratch, (size_t)n * sizeof(StkHit))) != STK_OK) return st;
    STK_HIP(c, hipMemcpyAsync(c->rayScratch.ptr, rays, (size_t)n * sizeof(StkRay), hipMemcpyHostToDevice, c->stream));
    StkRay* dr = static_cast<StkRay*>(c->rayScratch.ptr);
    StkHit* dh = static_cast<StkHit*>(c->hitScratch.ptr);
    if (any)
        STK_HIP(c, stk::launchTraceAny(bvhView(c), dr, n, dh, (flags & STK_TRACE_REFERENCE_ORDER) != 0, c->stream));
    else
        STK_HIP(c, stk::launchTraceClosest(bvhView(c), dr, n, dh, (flags & STK_TRACE_REFERENCE_ORDER) != 0,
                                           c->stream));
    STK_HIP(c, hipMemcpyAsync(hits, dh, (size_t)n * sizeof(StkHit), hipMemcpyDeviceToHost, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    return STK_OK;
}

STK_API stk_status stk_trace_closest(stk_ctx* c, const StkRay* rays, uint32_t n, StkHit* hits, uint32_t flags)
{
    return traceHost(c, rays, n, hits, 0, flags);
}

STK_API stk_status stk_trace_any(stk_ctx* c, const StkRay* rays, uint32_t n, StkHit* hits, uint32_t flags)
{
    return traceHost(c, rays, n, hits, 1, flags);
}

STK_API stk_status stk_generate_camera_rays_device(stk_ctx* c, const StkPathTracerParam* p, uint32_t sampleNum,
                                                   StkRay* rays)
{
    if (!validCtx(c) || !p) return STK_E_INVALID;
    if (p->dimension[0] <= 0 || p->dimension[1] <= 0 || p->spp <= 0 || !rays)
        return fail(c, STK_E_INVALID, "stk_generate_camera_rays_device: bad dimension, spp or buffer");
    (void)hipSetDevice(c->device);
    STK_HIP(c, stk::launchCameraRays(*p, sampleNum, rays, c->stream));
    return STK_OK;
}

static stk_status ensureRenderBuffers(stk_ctx* c, uint32_t W, uint32_t H, size_t nsamples)
{
    const size_t npix = (size_t)W * H;
    stk_status st;
    if (c->width != W || c->height != H) {
        c->iteration = 0;
        c->width = W;
        c->height = H;
        if ((st = devReserve(c, c->history, npix * sizeof(float4))) != STK_OK) return st;
        STK_HIP(c, stk::launchZero(static_cast<float4*>(c->history.ptr), npix, c->stream));
    }
    if ((st = devReserve(c, c->accum, npix * sizeof(float4))) != STK_OK) return st;
    if ((st = devReserve(c, c->out, npix * sizeof(float4))) != STK_OK) return st;
    if (nsamples) {
        if ((st = devReserve(c, c->samples, nsamples * sizeof(float))) != STK_OK) return st;
    }
    c->samplesCount = nsamples;
    if (c->ps.capacity < npix) {
        for (DevBuf& b : c->psBufs) devRelease(c, b);
        c->psBufs.assign(17, DevBuf{});
        stk::PathState& ps = c->ps;
        const size_t f4 = npix * sizeof(float4), u4 = npix * sizeof(uint32_t);
        void** ptrs[] = {(void**)&ps.rayO, (void**)&ps.rayD, (void**)&ps.thr, (void**)&ps.col, (void**)&ps.rng,
                         (void**)&ps.hit, (void**)&ps.hitPrim, (void**)&ps.queue[0], (void**)&ps.queue[1],
                         (void**)&ps.shadowQ, (void**)&ps.shO, (void**)&ps.shD, (void**)&ps.shCd, (void**)&ps.shCg,
                         (void**)&ps.counters};
        const size_t sizes[] = {f4, f4, f4, f4, u4, f4, u4, u4, u4, u4, f4, f4, f4, f4, 64};
        for (size_t i = 0; i < sizeof(sizes) / sizeof(sizes[0]); ++i) {
            if ((st = devReserve(c, c->psBufs[i], sizes[i])) != STK_OK) {
                ps.capacity = 0;
                return st;
            }
            *ptrs[i] = c->psBufs[i].ptr;
        }
        ps.capacity = (uint32_t)npix;
    }
    return STK_OK;
}

STK_API stk_status stk_render(stk_ctx* c, const StkRenderParams* p, float* outRgba, int outIsDevice)
{
    if (!validCtx(c) || !p) return STK_E_INVALID;
    if (!c->haveBvh) return fail(c, STK_E_STATE, "stk_render: no BVH");
    const int32_t W = p->pt.dimension[0], H = p->pt.dimension[1];
    if (W <= 0 || H <= 0 || p->pt.spp <= 0 || p->pt.maxDepth > 64)
        return fail(c, STK_E_INVALID, "stk_render: bad dimension, spp or maxDepth");
    if ((uint64_t)W * (uint64_t)H > (1ull << 28)) return fail(c, STK_E_INVALID, "stk_render: image too large");
    (void)hipSetDevice(c->device);
    const size_t npix = (size_t)W * H;
    const size_t nsamples = p->keepSamples ? npix * (size_t)p->pt.spp * 3 : 0;
    stk_status st = ensureRenderBuffers(c, (uint32_t)W, (uint32_t)H, nsamples);
    if (st != STK_OK) return st;
    stk::SceneView sv;
    sv.bvh = bvhView(c);
    sv.vb = static_cast<const StkVertex*>(c->vb.ptr);
    sv.ib = static_cast<const uint32_t*>(c->ib.ptr);
    sv.inst = static_cast<const StkInstance*>(c->inst.ptr);
    sv.lights = static_cast<const StkLight*>(c->lights.ptr);
    sv.numLights = c->numLights;
    sv.mats = static_cast<const StkMaterial*>(c->mats.ptr);
    sv.numMats = c->numMats;
    stk::PtLaunch lp;
    lp.pt = p->pt;
    for (int i = 0; i < 4; ++i) lp.sky[i] = p->skyColor[i];
    lp.rrStartDepth = p->rrStartDepth;
    lp.refCompat = p->refCompat;
    lp.invSpp = 1.0f / (float)p->pt.spp;
    stk::PtCache& pc = lp.cache;
    pc = stk::PtCache{};
    if (p->cacheMode != 0) {
        if (!c->haveCache) return fail(c, STK_E_STATE, "stk_render: cacheMode set but no cache (stk_cache_create)");
        if (p->cacheMode > 3) return fail(c, STK_E_INVALID, "stk_render: cacheMode must be 0..3");
        pc.mode = p->cacheMode;
        pc.queryDepth = p->cacheQueryDepth;
        const uint32_t paths = p->cacheTrainPaths ? p->cacheTrainPaths : 65536u;
        pc.trainStride = std::max<uint32_t>(1u, (uint32_t)(npix / paths));
        const size_t maxTrain = (npix + pc.trainStride - 1) / pc.trainStride;
        pc.trainCap = (uint32_t)std::min<size_t>(c->cache.nCap, maxTrain * p->pt.maxDepth);
        if ((st = devReserve(c, c->queryRecs, npix * sizeof(stk::CacheRecord))) != STK_OK) return st;
        if ((st = devReserve(c, c->queryOut, npix * sizeof(float4))) != STK_OK) return st;
        if ((st = devReserve(c, c->trainVerts, maxTrain * p->pt.maxDepth * 5 * sizeof(float4))) != STK_OK) return st;
        if ((st = devReserve(c, c->trainCount, maxTrain * sizeof(uint32_t))) != STK_OK) return st;
        if ((st = devReserve(c, c->trainBatch, (size_t)pc.trainCap * sizeof(stk::CacheRecord))) != STK_OK) return st;
        pc.queryRecs = static_cast<stk::CacheRecord*>(c->queryRecs.ptr);
        pc.queryOut = static_cast<float4*>(c->queryOut.ptr);
        pc.trainVerts = static_cast<float4*>(c->trainVerts.ptr);
        pc.trainCount = static_cast<uint32_t*>(c->trainCount.ptr);
        pc.trainBatch = static_cast<stk::CacheRecord*>(c->trainBatch.ptr);
    }
    float4* accum = static_cast<float4*>(c->accum.ptr);
    STK_HIP(c, stk::launchZero(accum, npix, c->stream));
    float* samples = nsamples ? static_cast<float*>(c->samples.ptr) : nullptr;
    for (int s = 0; s < p->pt.spp; ++s) {
        lp.sampleNum = (uint32_t)s;
        if (pc.mode != 0) {
            // a different subset of pixels trains each pass
            uint32_t h = p->pt.frameNumber * (uint32_t)p->pt.spp + (uint32_t)s;
            h ^= h >> 16;
            h *= 0x7feb352du;
            h ^= h >> 15;
            pc.trainOffset = h % pc.trainStride;
            pc.numTrainPaths = (uint32_t)((npix - pc.trainOffset + pc.trainStride - 1) / pc.trainStride);
        }
        STK_HIP(c, stk::launchPathTracePass(sv, lp, c->ps, accum, samples, c->haveCache ? &c->cache : nullptr,
                                            c->stream));
    }
    const int acc = p->enableAccumulation != 0;
    STK_HIP(c, stk::launchAccumulateTonemap(accum, static_cast<float4*>(c->history.ptr),
                                             outIsDevice && outRgba ? reinterpret_cast<float4*>(outRgba)
                                                                    : static_cast<float4*>(c->out.ptr),
                                             (uint32_t)npix, c->iteration, acc, p->enableTonemap != 0,
                                             p->tonemapperType, c->stream));
    if (acc) ++c->iteration;
    if (outRgba && !outIsDevice) {
        STK_HIP(c, hipMemcpyAsync(outRgba, c->out.ptr, npix * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
        STK_HIP(c, hipStreamSynchronize(c->stream));
    }
    return STK_OK;
}

STK_API stk_status stk_reset_accumulation(stk_ctx* c)
{
    if (!validCtx(c)) return STK_E_INVALID;
    c->iteration = 0;
    return STK_OK;
}

STK_API stk_status stk_get_samples(stk_ctx* c, float* out, size_t count)
{
    if (!validCtx(c) || !out) return STK_E_INVALID;
    if (count > c->samplesCount || !c->samples.ptr)
        return fail(c, STK_E_STATE, "stk_get_samples: no samples kept (set keepSamples) or count too large");
    STK_HIP(c, hipMemcpyAsync(out, c->samples.ptr, count * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    return STK_OK;
}

STK_API stk_status stk_get_accumulation(stk_ctx* c, float* out, size_t count)
{
    if (!validCtx(c) || !out) return STK_E_INVALID;
    if (count > (size_t)c->width * c->height * 4 || !c->history.ptr)
        return fail(c, STK_E_STATE, "stk_get_accumulation: nothing rendered or count too large");
    STK_HIP(c, hipMemcpyAsync(out, c->history.ptr, count * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    return STK_OK;
}

// ---- radiance cache -----------------------------------------------------------

static StkCacheConfig defaultCacheConfig()
{
    StkCacheConfig c{};
    c.log2TableSize = 19;
    c.baseResolution = 16;
    c.maxResolution = 2048;
    c.seed = 1;
    c.learningRate = 1e-2f;
    c.beta1 = 0.9f;
    c.beta2 = 0.99f;
    c.epsilon = 1e-15f;
    c.maxBatch = 1u << 18;
    c.emaDecay = 0.99f;
    return c;
}

STK_API stk_status stk_cache_destroy(stk_ctx* c)
{
    if (!validCtx(c)) return STK_E_INVALID;
    DevBuf* bufs[] = {&c->cGridH, &c->cGridP, &c->cGridM, &c->cGridV, &c->cGridG, &c->cMlpP, &c->cMlpM,
                      &c->cMlpV, &c->cMlpG, &c->cFwd, &c->cBwd, &c->cXT, &c->cDyT, &c->cLoss,
                      &c->cGridE, &c->cGridHE, &c->cMlpE, &c->cFwdE, &c->cBwdE};
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (DevBuf* b : bufs) devRelease(c, *b);
    c->cache = stk::CacheBuffers{};
    c->haveCache = false;
    return STK_OK;
}

STK_API stk_status stk_cache_create(stk_ctx* c, const StkCacheConfig* cfgIn)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveBvh) return fail(c, STK_E_STATE, "stk_cache_create: the scene bounds come from the BVH; set one first");
    StkCacheConfig cfg = cfgIn ? *cfgIn : defaultCacheConfig();
    if (cfg.log2TableSize < 8 || cfg.log2TableSize > 24 || cfg.baseResolution < 2 ||
        cfg.maxResolution < cfg.baseResolution || cfg.maxResolution > (1u << 20) || cfg.maxBatch == 0 ||
        cfg.maxBatch > (1u << 24) || !(cfg.learningRate > 0.0f) || !(cfg.emaDecay >= 0.0f && cfg.emaDecay < 1.0f))
        return fail(c, STK_E_INVALID, "stk_cache_create: configuration out of range");
    stk_cache_destroy(c);
    stk::CacheBuffers& cb = c->cache;
    stk::GridDesc& gd = cb.grid;
    const StkBvhHeader& h = c->bvh.header;
    float ext = 0.0f;
    for (int a = 0; a < 3; ++a) {
        gd.lo[a] = (h.root >= 0) ? h.lo[a] : 0.0f;
        if (h.root >= 0) ext = std::max(ext, h.hi[a] - h.lo[a]);
    }
    gd.scale = ext > 0.0f ? 1.0f / ext : 1.0f;
    const uint64_t T = 1ull << cfg.log2TableSize;
    gd.tableMask = (uint32_t)(T - 1);
    gd.log2Table = cfg.log2TableSize;
    const double growth = std::exp((std::log((double)cfg.maxResolution) - std::log((double)cfg.baseResolution)) /
                                   (double)(stk::kCacheLevels - 1));
    for (int l = 0; l < stk::kCacheLevels; ++l) {
        const uint32_t res = (uint32_t)std::floor((double)cfg.baseResolution * std::pow(growth, (double)l) + 1e-9);
        gd.res[l] = res;
        const uint64_t s = (uint64_t)res + 1;
        gd.dense[l] = (s * s * s <= T) ? 1u : 0u;
    }
    cb.gridParams = (size_t)stk::kCacheLevels * T * stk::kCacheFeatures;
    cb.nCap = (cfg.maxBatch + 1023u) & ~1023u;
    cb.lr = cfg.learningRate;
    cb.beta1 = cfg.beta1;
    cb.beta2 = cfg.beta2;
    cb.eps = cfg.epsilon;
    cb.step = 0;
    const size_t gp = cb.gridParams, mp = stk::kCacheMlpParams;
    struct {
        DevBuf* b;
        size_t bytes;
    } allocs[] = {{&c->cGridH, gp * 2},  {&c->cGridP, gp * 4}, {&c->cGridM, gp * 4},   {&c->cGridV, gp * 4},
                  {&c->cGridG, gp * 4},  {&c->cMlpP, mp * 4},  {&c->cMlpM, mp * 4},    {&c->cMlpV, mp * 4},
                  {&c->cMlpG, mp * 4},   {&c->cFwd, 68 * 256 * 2}, {&c->cBwd, 68 * 256 * 2},
                  {&c->cXT, (size_t)5 * 64 * cb.nCap * 2}, {&c->cDyT, (size_t)272 * cb.nCap * 2},
                  {&c->cLoss, 64}};
    (void)hipSetDevice(c->device);
    for (auto& a : allocs) {
        stk_status st = devReserve(c, *a.b, a.bytes);
        if (st != STK_OK) {
            stk_cache_destroy(c);
            return st;
        }
    }
    cb.emaDecay = cfg.emaDecay;
    if (cb.emaDecay > 0.0f) {
        struct {
            DevBuf* b;
            size_t bytes;
        } ema[] = {{&c->cGridE, gp * 4}, {&c->cGridHE, gp * 2}, {&c->cMlpE, mp * 4}, {&c->cFwdE, 68 * 256 * 2},
                   {&c->cBwdE, 68 * 256 * 2}};
        for (auto& a : ema) {
            stk_status st = devReserve(c, *a.b, a.bytes);
            if (st != STK_OK) {
                stk_cache_destroy(c);
                return st;
            }
        }
        cb.gridE = static_cast<float*>(c->cGridE.ptr);
        cb.gridHE = static_cast<__half*>(c->cGridHE.ptr);
        cb.mlpE = static_cast<float*>(c->cMlpE.ptr);
        cb.wFwdE = static_cast<__half*>(c->cFwdE.ptr);
        cb.wBwdE = static_cast<__half*>(c->cBwdE.ptr);
    }
    cb.gridH = static_cast<__half*>(c->cGridH.ptr);
    cb.gridP = static_cast<float*>(c->cGridP.ptr);
    cb.gridM = static_cast<float*>(c->cGridM.ptr);
    cb.gridV = static_cast<float*>(c->cGridV.ptr);
    cb.gridG = static_cast<float*>(c->cGridG.ptr);
    cb.mlpP = static_cast<float*>(c->cMlpP.ptr);
    cb.mlpM = static_cast<float*>(c->cMlpM.ptr);
    cb.mlpV = static_cast<float*>(c->cMlpV.ptr);
    cb.mlpG = static_cast<float*>(c->cMlpG.ptr);
    cb.wFwd = static_cast<__half*>(c->cFwd.ptr);
    cb.wBwd = static_cast<__half*>(c->cBwd.ptr);
    cb.xT = static_cast<__half*>(c->cXT.ptr);
    cb.dyT = static_cast<__half*>(c->cDyT.ptr);
    cb.loss = static_cast<float*>(c->cLoss.ptr);
    // initialisation: grid U(-1e-4, 1e-4); weights He-uniform U(-sqrt(6/64), sqrt(6/64))
    std::vector<float> grid(gp), mlp(mp);
    uint64_t s = 0x9E3779B97F4A7C15ull ^ ((uint64_t)cfg.seed << 1);
    auto uni = [&s]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (float)((s >> 40) & 0xFFFFFF) * (1.0f / 16777216.0f);
    };
    for (float& x : grid) x = (uni() * 2.0f - 1.0f) * 1e-4f;
    const float a = std::sqrt(6.0f / 64.0f);
    for (float& x : mlp) x = (uni() * 2.0f - 1.0f) * a;
    for (size_t in = 0; in < 64; ++in)
        for (size_t o = 3; o < 16; ++o) mlp[16384 + in * 16 + o] = 0.0f; // unused output columns
    STK_HIP(c, hipMemcpyAsync(cb.gridP, grid.data(), gp * 4, hipMemcpyHostToDevice, c->stream));
    STK_HIP(c, hipMemcpyAsync(cb.mlpP, mlp.data(), mp * 4, hipMemcpyHostToDevice, c->stream));
    STK_HIP(c, hipMemsetAsync(cb.gridM, 0, gp * 4, c->stream));
    STK_HIP(c, hipMemsetAsync(cb.gridV, 0, gp * 4, c->stream));
    STK_HIP(c, hipMemsetAsync(cb.gridG, 0, gp * 4, c->stream));
    STK_HIP(c, hipMemsetAsync(cb.mlpM, 0, mp * 4, c->stream));
    STK_HIP(c, hipMemsetAsync(cb.mlpV, 0, mp * 4, c->stream));
    STK_HIP(c, hipMemsetAsync(cb.mlpG, 0, mp * 4, c->stream));
    STK_HIP(c, stk::cachePack(cb, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    c->cacheCfg = cfg;
    c->haveCache = true;
    return STK_OK;
}

STK_API stk_status stk_cache_info(stk_ctx* c, StkCacheInfo* out)
{
    if (!validCtx(c) || !out) return STK_E_INVALID;
    if (!c->haveCache) return fail(c, STK_E_STATE, "stk_cache_info: no cache");
    const stk::GridDesc& gd = c->cache.grid;
    for (int a = 0; a < 3; ++a) out->lo[a] = gd.lo[a];
    out->scale = gd.scale;
    for (int l = 0; l < 16; ++l) {
        out->resolution[l] = gd.res[l];
        out->dense[l] = gd.dense[l];
    }
    out->log2TableSize = gd.log2Table;
    out->gridParams = (uint32_t)c->cache.gridParams;
    out->mlpParams = stk::kCacheMlpParams;
    out->step = c->cache.step;
    return STK_OK;
}

STK_API stk_status stk_cache_infer(stk_ctx* c, const StkCacheRecord* recs, uint32_t n, float* out)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveCache) return fail(c, STK_E_STATE, "stk_cache_infer: no cache");
    if (n && (!recs || !out)) return fail(c, STK_E_INVALID, "stk_cache_infer: null buffer");
    (void)hipSetDevice(c->device);
    STK_HIP(c, stk::cacheInfer(c->cache, reinterpret_cast<const stk::CacheRecord*>(recs), n, nullptr,
                               reinterpret_cast<float4*>(out), c->stream));
    return STK_OK;
}

STK_API stk_status stk_cache_train_step(stk_ctx* c, const StkCacheRecord* recs, uint32_t n, float* lossOut)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveCache) return fail(c, STK_E_STATE, "stk_cache_train_step: no cache");
    if (n == 0 || !recs) return fail(c, STK_E_INVALID, "stk_cache_train_step: empty batch");
    if (n > c->cache.nCap) return fail(c, STK_E_INVALID, "stk_cache_train_step: batch above maxBatch");
    (void)hipSetDevice(c->device);
    STK_HIP(c, stk::cacheBackward(c->cache, reinterpret_cast<const stk::CacheRecord*>(recs), n, nullptr, c->stream));
    float loss = 0.0f;
    STK_HIP(c, hipMemcpyAsync(&loss, c->cache.loss, 4, hipMemcpyDeviceToHost, c->stream));
    STK_HIP(c, stk::cacheAdam(c->cache, 1.0f / (float)n, nullptr, n, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    if (lossOut) *lossOut = loss / (float)n;
    return STK_OK;
}

STK_API stk_status stk_cache_gradients(stk_ctx* c, const StkCacheRecord* recs, uint32_t n, float* gridGrad,
                                       float* mlpGrad, float* lossSum)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveCache) return fail(c, STK_E_STATE, "stk_cache_gradients: no cache");
    if (n == 0 || !recs || n > c->cache.nCap) return fail(c, STK_E_INVALID, "stk_cache_gradients: bad batch");
    (void)hipSetDevice(c->device);
    const stk::CacheBuffers& cb = c->cache;
    STK_HIP(c, stk::cacheZeroGrads(cb, c->stream));
    STK_HIP(c, stk::cacheBackward(cb, reinterpret_cast<const stk::CacheRecord*>(recs), n, nullptr, c->stream));
    if (gridGrad)
        STK_HIP(c, hipMemcpyAsync(gridGrad, cb.gridG, cb.gridParams * 4, hipMemcpyDeviceToHost, c->stream));
    if (mlpGrad)
        STK_HIP(c, hipMemcpyAsync(mlpGrad, cb.mlpG, stk::kCacheMlpParams * 4, hipMemcpyDeviceToHost, c->stream));
    if (lossSum) STK_HIP(c, hipMemcpyAsync(lossSum, cb.loss, 4, hipMemcpyDeviceToHost, c->stream));
    STK_HIP(c, stk::cacheZeroGrads(cb, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    return STK_OK;
}

STK_API stk_status stk_cache_get_params(stk_ctx* c, float* grid, float* mlp)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveCache) return fail(c, STK_E_STATE, "stk_cache_get_params: no cache");
    if (grid)
        STK_HIP(c, hipMemcpyAsync(grid, c->cache.gridP, c->cache.gridParams * 4, hipMemcpyDeviceToHost, c->stream));
    if (mlp)
        STK_HIP(c, hipMemcpyAsync(mlp, c->cache.mlpP, stk::kCacheMlpParams * 4, hipMemcpyDeviceToHost, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    return STK_OK;
}

STK_API stk_status stk_cache_set_params(stk_ctx* c, const float* grid, const float* mlp)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveCache) return fail(c, STK_E_STATE, "stk_cache_set_params: no cache");
    if (grid)
        STK_HIP(c, hipMemcpyAsync(c->cache.gridP, grid, c->cache.gridParams * 4, hipMemcpyHostToDevice, c->stream));
    if (mlp)
        STK_HIP(c, hipMemcpyAsync(c->cache.mlpP, mlp, stk::kCacheMlpParams * 4, hipMemcpyHostToDevice, c->stream));
    STK_HIP(c, stk::cachePack(c->cache, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    return STK_OK;
}

STK_API stk_status stk_get_counters(stk_ctx* c, uint32_t out[8])
{
    if (!validCtx(c) || !out) return STK_E_INVALID;
    if (!c->ps.counters) return fail(c, STK_E_STATE, "stk_get_counters: nothing rendered");
    STK_HIP(c, hipMemcpyAsync(out, c->ps.counters, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    return STK_OK;
}

STK_API stk_status stk_device_alloc(stk_ctx* c, size_t bytes, void** out)
{
    if (!validCtx(c) || !out) return STK_E_INVALID;
    *out = nullptr;
    if (failAllocForTest()) return fail(c, STK_E_OOM, "device allocation failed (test hook)");
    (void)hipSetDevice(c->device);
    STK_HIP(c, hipMalloc(out, bytes));
    return STK_OK;
}

STK_API stk_status stk_device_free(stk_ctx* c, void* p)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (p) STK_HIP(c, hipFree(p));
    return STK_OK;
}

STK_API stk_status stk_copy_to_device(stk_ctx* c, void* dst, const void* src, size_t bytes)
{
    if (!validCtx(c) || (bytes && (!dst || !src))) return STK_E_INVALID;
    STK_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    return STK_OK;
}

STK_API stk_status stk_copy_to_host(stk_ctx* c, void* dst, const void* src, size_t bytes)
{
    if (!validCtx(c) || (bytes && (!dst || !src))) return STK_E_INVALID;
    STK_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    return STK_OK;
}

STK_API stk_status stk_device_copy(stk_ctx* c, void* dst, const void* src, size_t bytes)
{
    if (!validCtx(c) || (bytes && (!dst || !src)) || bytes % 16) return STK_E_INVALID;
    (void)hipSetDevice(c->device);
    STK_HIP(c, stk::launchCopy(dst, src, bytes, c->stream));
    return STK_OK;
}

STK_API stk_status stk_event_record(stk_ctx* c, int slot)
{
    if (!validCtx(c) || slot < 0 || slot >= 16) return STK_E_INVALID;
    STK_HIP(c, hipEventRecord(c->events[slot], c->stream));
    return STK_OK;
}

STK_API stk_status stk_event_elapsed(stk_ctx* c, int a, int b, float* ms)
{
    if (!validCtx(c) || a < 0 || a >= 16 || b < 0 || b >= 16 || !ms) return STK_E_INVALID;
    STK_HIP(c, hipEventSynchronize(c->events[b]));
    STK_HIP(c, hipEventElapsedTime(ms, c->events[a], c->events[b]));
    return STK_OK;
}

} // extern "C"
