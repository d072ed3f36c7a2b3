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

#include "../../include/stk/stk.h"
#include "bvh.hpp"
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

    std::vector<DevBuf*> owned()
    {
        return {&vb, &ib, &inst, &lights, &mats, &packed, &tris, &refNodes, &rayScratch, &hitScratch,
                &accum, &history, &out, &samples};
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
    float4* accum = static_cast<float4*>(c->accum.ptr);
    STK_HIP(c, stk::launchZero(accum, npix, c->stream));
    float* samples = nsamples ? static_cast<float*>(c->samples.ptr) : nullptr;
    for (int s = 0; s < p->pt.spp; ++s) {
        lp.sampleNum = (uint32_t)s;
        STK_HIP(c, stk::launchPathTracePass(sv, lp, c->ps, accum, samples, c->stream));
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
