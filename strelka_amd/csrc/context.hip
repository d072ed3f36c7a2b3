// context.hip - the C ABI (include/stk/stk.h): context, device buffers, BVH
// upload, traversal and render entry points.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>
#include <new>
#include <string>
#include <vector>

#include <cmath>

#include "../../include/stk/stk.h"
#include "../../include/stk/stk_scene.h"
#include "bvh.hpp"
#include "cache.hpp"
#include "gpubuild.hpp"
#include "kernels.hpp"

namespace {

// Most streams a render call spreads its passes over (StkLaunchConfig::passStreams).
constexpr int kMaxPassStreams = 4;

// A device allocation owned by a context.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// One scene staged for a batched render (stk_batch_add_scene) and its BVH
// (built by stk_batch_commit).
struct BatchScene {
    std::vector<StkVertex> vb;
    std::vector<uint32_t> ib;
    std::vector<StkInstance> inst;
    std::vector<StkLight> lights;
    std::vector<StkMaterial> mats;
    StkPathTracerParam cam;
    float sky[4];
    stk::Bvh bvh;
    std::string err;
};

// ---- static staging for stk_trace_closest_noalloc ----------------------------
// The one entry point that allocates nothing stages its rays and hits through
// this file-scope device array instead of the context's allocations: rays in
// the first 32 MiB, hits in the next 20 MiB.  Every GPU the library runs on
// holds its own copy (52 MiB) until the process ends.  In host code `scratch`
// names a host-side stand-in, never the device copy: host code uses the
// per-device address looked up by setupScratch.
constexpr uint32_t kNoAllocMaxRays = 1u << 20;
constexpr size_t kNoAllocRayBytes = (size_t)kNoAllocMaxRays * sizeof(StkRay);
constexpr size_t kNoAllocBytes = kNoAllocRayBytes + (size_t)kNoAllocMaxRays * sizeof(StkHit);
static_assert(kNoAllocRayBytes % 256 == 0, "the hit records start 256-B aligned");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratchMu;
std::atomic<void*> g_scratchPtr[kMaxScratchDevices];
hipError_t g_scratchErr[kMaxScratchDevices];

// With `device` current: looks up that device's copy of `scratch`, once per
// device (a failed lookup is retried by the next context created there).
void setupScratch(int device)
{
    if (device < 0 || device >= kMaxScratchDevices) return;
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (g_scratchPtr[device].load() != nullptr) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratchErr[device] = e;
    if (e != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        return;
    }
    g_scratchPtr[device].store(p);
}

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
    // two-level structure (stk_build_two_level): replaces the flat tree
    stk::TwoLevel tl;
    bool twoLevel = false;
    // the flat tree's device arrays as the walks see them: counts of the wide
    // nodes and emitter triangles, and whether the tree was built on the GPU
    // (then the host-side arrays of bvh are empty and stk_get_bvh downloads)
    uint32_t numNodes = 0, numWide = 0, numEmit = 0;
    bool gpuBuilt = false;

    DevBuf vb, ib, inst, lights, mats, packed, tris, refNodes, wideNodes, wideTris, emitters, textures, texels;
    uint32_t numTextures = 0;
    DevBuf tlasWide, tlasLeaves, blasWide, blasTris, instKey, instW, tlEmit;
    // the two-level structure's world triangles (stk_two_level_world_triangles)
    DevBuf tlWorldTris, tlTriOff;
    bool tlWorld = false;
    DevBuf rayScratch, hitScratch;

    // render state.  One pass stream (passStreams = 1): every pass runs on
    // `stream` with `ps`.  n > 1 pass streams: the passes run on xStream[i]
    // with xPs[i], in turn across render calls, and `stream` only takes each
    // call's accumulate / tonemap step, after joining the pass streams, so a
    // call's passes overlap the previous call's last ones.  Calls alternate
    // between two accumulation buffers (accum, accum2); accumFree[l] marks
    // the last read of buffer l.  finEv[i]: finalize of the last pass on
    // xStream[i]; joinEv[i]: joins.
    stk::PathState ps;
    std::vector<DevBuf> psBufs;
    stk::PathState xPs[kMaxPassStreams];
    std::vector<DevBuf> xPsBufs[kMaxPassStreams];
    hipStream_t xStream[kMaxPassStreams] = {};
    hipEvent_t finEv[kMaxPassStreams] = {}, joinEv[kMaxPassStreams] = {}, accumFree[2] = {};
    // radiance cache with several pass streams: each stream's query and
    // training-vertex buffers, and cacheEv[i], the end of the cache steps of
    // the last pass on xStream[i] (the next pass's inference waits for it)
    DevBuf xQueryRecs[kMaxPassStreams], xQueryOut[kMaxPassStreams], xTrainVerts[kMaxPassStreams],
        xTrainCount[kMaxPassStreams];
    hipEvent_t cacheEv[kMaxPassStreams] = {};
    bool accumFreeSet[2] = {false, false};
    uint32_t renderCalls = 0, passTurn = 0; // multi-stream calls so far, next pass stream
    uint32_t lastStreams = 1;    // pass streams of the last render call
    uint32_t lastPassStream = 0; // stream of its last pass
    std::vector<stk::SceneDesc> descsHost; // what descs holds

    stk::PathState& passState(uint32_t i) { return lastStreams > 1 ? xPs[i] : ps; }
    DevBuf accum, accum2, history, out, samples, descs, staging;
    uint32_t iteration = 0;
    uint64_t imageKey = 0;
    size_t imagePixels = 0, samplesCount = 0;
    stk::LaunchConfig launch;
    bool useGraph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    std::vector<uint64_t> graphKey;

    // batched scenes: staged host scenes and the merged device forest
    std::vector<BatchScene> batch;
    bool batchReady = false;
    std::vector<stk::SceneDesc> batchDescs;
    DevBuf fVb, fIb, fInst, fLights, fMats, fPacked, fTris, fRefNodes, fWide, fWideTris, fEmit;
    uint32_t fNumMats = 0;

    // radiance cache
    stk::CacheBuffers cache;
    bool haveCache = false;
    StkCacheConfig cacheCfg{};
    DevBuf cGridH, cGridP, cGridM, cGridV, cGridG, cMlpP, cMlpM, cMlpV, cMlpG, cFwd, cBwd, cXT, cDyT, cLoss;
    DevBuf cGridE, cGridEW, cGridHE, cMlpE, cFwdE, cBwdE, cGridGD, cSortKeys, cSortVals, cSortTemp;
    DevBuf queryRecs, queryOut, trainVerts, trainCount, trainBatch;
    uint32_t lastTrainCap = 0; // training records the last cached render could hold

    // watchdog (stk_set_watchdog / stk_abort): the abort word in mapped
    // pinned host memory (host and device addresses), the device word the
    // accumulation step reads (abortAny), and a thread that sets the host word
    // when an armed deadline passes.  The state of the call before the last
    // one lets a later call undo a frame that was abandoned after its call
    // had returned (device output, no watchdog armed).
    uint32_t* abortHost = nullptr;
    uint32_t* abortHostDev = nullptr;
    void* abortAny = nullptr;
    std::atomic<uint32_t> watchdogMs{0};
    std::mutex wdMutex;
    std::condition_variable wdCv;
    std::thread wdThread;
    bool wdStop = false, wdArmed = false;
    std::chrono::steady_clock::time_point wdDeadline;
    uint32_t lastIterBefore = 0, lastStepBefore = 0, lastEmaBefore = 0;
    bool lastSynced = true; // the last render call waited for its frame

    std::vector<DevBuf*> owned()
    {
        return {&vb, &ib, &inst, &lights, &mats, &textures, &texels, &packed, &tris, &refNodes, &wideNodes, &wideTris, &emitters,
                &tlasWide, &tlasLeaves, &blasWide, &blasTris, &instKey, &instW, &tlEmit, &tlWorldTris, &tlTriOff,
                &rayScratch, &hitScratch,
                &accum, &accum2, &history, &out, &samples, &descs, &staging, &fVb, &fIb, &fInst, &fLights, &fMats, &fPacked,
                &fTris, &fRefNodes, &fWide, &fWideTris, &fEmit, &cGridH, &cGridP, &cGridM, &cGridV, &cGridG, &cMlpP,
                &cMlpM, &cMlpV, &cMlpG, &cFwd, &cBwd, &cXT, &cDyT, &cLoss, &cGridE, &cGridEW, &cGridHE, &cMlpE,
                &cFwdE, &cBwdE, &cGridGD, &cSortKeys, &cSortVals, &cSortTemp, &queryRecs, &queryOut, &trainVerts, &trainCount, &trainBatch,
                &xQueryRecs[0], &xQueryRecs[1], &xQueryRecs[2], &xQueryRecs[3], &xQueryOut[0], &xQueryOut[1],
                &xQueryOut[2], &xQueryOut[3], &xTrainVerts[0], &xTrainVerts[1], &xTrainVerts[2], &xTrainVerts[3],
                &xTrainCount[0], &xTrainCount[1], &xTrainCount[2], &xTrainCount[3]};
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
        // the pass streams may still read any buffer
        for (hipStream_t s : c->xStream)
            if (s) (void)hipStreamSynchronize(s);
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

// The wide walk needs one node-group stack entry per wide level.
bool wideUsable(const stk::Bvh& b) { return !b.wide.empty() && b.wideDepth <= (uint32_t)stk::dev::kWideStack; }

// The same for the context's flat tree (host-built or GPU-built).
bool ctxWideUsable(const stk_ctx* c) { return c->numWide > 0 && c->bvh.wideDepth <= (uint32_t)stk::dev::kWideStack; }

stk::BvhView bvhView(stk_ctx* c)
{
    stk::BvhView v;
    v.packed = static_cast<const float4*>(c->packed.ptr);
    v.tris = static_cast<const float4*>(c->tris.ptr);
    v.refNodes = static_cast<const StkBvhNode*>(c->refNodes.ptr);
    v.instances = static_cast<const StkInstance*>(c->inst.ptr);
    v.hdr = c->bvh.header;
    v.wide = static_cast<const float4*>(c->wideNodes.ptr);
    v.wideTris = static_cast<const float4*>(c->wideTris.ptr);
    v.emitters = static_cast<const float4*>(c->emitters.ptr);
    v.wideRoot = ctxWideUsable(c) ? 0u : STK_INVALID_INDEX;
    v.emitCount = c->numEmit;
    v.packedOk = c->bvh.depth <= (uint32_t)stk::dev::kStackDepth ? 1u : 0u;
    v.wideDepth = c->bvh.wideDepth;
    v.tlas = v.tlasLeaves = v.blas = v.blasTris = v.instW = v.worldTris = nullptr;
    v.instKey = v.instTriOff = nullptr;
    if (c->twoLevel) {
        v.tlas = static_cast<const float4*>(c->tlasWide.ptr);
        v.tlasLeaves = static_cast<const float4*>(c->tlasLeaves.ptr);
        v.blas = static_cast<const float4*>(c->blasWide.ptr);
        v.blasTris = static_cast<const float4*>(c->blasTris.ptr);
        v.instKey = static_cast<const uint32_t*>(c->instKey.ptr);
        v.instW = static_cast<const float4*>(c->instW.ptr);
        v.emitters = static_cast<const float4*>(c->tlEmit.ptr);
        v.emitCount = (uint32_t)c->tl.emitters.size();
        v.wideRoot = STK_INVALID_INDEX;
        if (c->tlWorld) {
            v.worldTris = static_cast<const float4*>(c->tlWorldTris.ptr);
            v.instTriOff = static_cast<const uint32_t*>(c->tlTriOff.ptr);
        }
    }
    return v;
}

// Uploads a validated host tree and makes it the context's flat tree.  The
// context's tree is replaced only once every device array holds the new one;
// a failed upload leaves the context without a tree (the device arrays are
// then a mix of old and new), never with a host description of one tree over
// the device arrays of another.
stk_status uploadBvh(stk_ctx* c, stk::Bvh&& b)
{
    c->haveBvh = false;
    c->twoLevel = false;
    c->numWide = 0;
    c->numNodes = 0;
    c->numEmit = 0;
    // device copy of the skip-link array: leaves carry their triangle index in maxBounds[0]
    std::vector<StkBvhNode> dev(b.nodes);
    for (size_t i = 0; i < dev.size(); ++i)
        if ((uint32_t)dev[i].instId != STK_INVALID_INDEX) dev[i].maxBounds[0] = stk::bits_to_float(b.leafToTri[i]);
    stk_status st;
    if ((st = devUpload(c, c->refNodes, dev.data(), dev.size() * sizeof(StkBvhNode))) != STK_OK) return st;
    if ((st = devUpload(c, c->packed, b.packed.data(), b.packed.size() * sizeof(StkPackedNode))) != STK_OK) return st;
    if ((st = devUpload(c, c->tris, b.tris.data(), b.tris.size() * sizeof(StkTriangle))) != STK_OK) return st;
    if ((st = devUpload(c, c->wideNodes, b.wide.data(), b.wide.size() * sizeof(StkWideNode))) != STK_OK) return st;
    if ((st = devUpload(c, c->wideTris, b.wideTris.data(), b.wideTris.size() * sizeof(StkTriangle))) != STK_OK)
        return st;
    if ((st = devUpload(c, c->emitters, b.emitters.data(), b.emitters.size() * sizeof(StkTriangle))) != STK_OK)
        return st;
    c->bvh = std::move(b);
    c->numNodes = (uint32_t)c->bvh.nodes.size();
    c->numWide = (uint32_t)c->bvh.wide.size();
    c->numEmit = (uint32_t)c->bvh.emitters.size();
    c->gpuBuilt = false;
    c->haveBvh = true;
    return STK_OK;
}

bool validCtx(stk_ctx* c) { return c != nullptr; }

void invalidateGraph(stk_ctx* c)
{
    if (c->graphExec) (void)hipGraphExecDestroy(c->graphExec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graphExec = nullptr;
    c->graph = nullptr;
    c->graphKey.clear();
}

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
    void* h = nullptr;
    if (hipHostMalloc(&h, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) return STK_E_DEVICE;
    std::memset(h, 0, 64);
    c->abortHost = static_cast<uint32_t*>(h);
    void* hd = nullptr;
    if (hipHostGetDevicePointer(&hd, h, 0) != hipSuccess || hipMalloc(&c->abortAny, 256) != hipSuccess ||
        hipMemset(c->abortAny, 0, 256) != hipSuccess) {
        if (c->abortAny) (void)hipFree(c->abortAny);
        (void)hipHostFree(h);
        return STK_E_DEVICE;
    }
    c->abortHostDev = static_cast<uint32_t*>(hd);
    for (stk::PathState* ps : {&c->ps, &c->xPs[0], &c->xPs[1], &c->xPs[2], &c->xPs[3]}) {
        ps->abortHost = c->abortHostDev;
        ps->abortAny = static_cast<uint32_t*>(c->abortAny);
    }
    setupScratch(hipDevice);
    *out = c.release();
    return STK_OK;
}

STK_API void stk_destroy(stk_ctx* c)
{
    if (!c) return;
    if (c->wdThread.joinable()) {
        {
            std::lock_guard<std::mutex> lk(c->wdMutex);
            c->wdStop = true;
        }
        c->wdCv.notify_all();
        c->wdThread.join();
    }
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipStream_t s : c->xStream)
        if (s) (void)hipStreamSynchronize(s);
    for (DevBuf* b : c->owned()) devRelease(c, *b);
    for (DevBuf& b : c->psBufs) devRelease(c, b);
    for (auto& bufs : c->xPsBufs)
        for (DevBuf& b : bufs) devRelease(c, b);
    invalidateGraph(c);
    for (auto& ev : c->events)
        if (ev) (void)hipEventDestroy(ev);
    for (int i = 0; i < kMaxPassStreams; ++i) {
        if (c->finEv[i]) (void)hipEventDestroy(c->finEv[i]);
        if (c->joinEv[i]) (void)hipEventDestroy(c->joinEv[i]);
        if (c->cacheEv[i]) (void)hipEventDestroy(c->cacheEv[i]);
    }
    for (auto& ev : c->accumFree)
        if (ev) (void)hipEventDestroy(ev);
    for (hipStream_t s : c->xStream)
        if (s) (void)hipStreamDestroy(s);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->abortAny) (void)hipFree(c->abortAny);
    if (c->abortHost) (void)hipHostFree(c->abortHost);
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
    c->twoLevel = false;
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
    for (uint32_t i = 0; i < n; ++i)
        if (m[i].type > STK_MAT_GLASS || !(m[i].ior >= 0.0f))
            return fail(c, STK_E_INVALID, "stk_set_materials: material " + std::to_string(i) +
                                              " has an unknown type or a negative index of refraction");
    (void)hipSetDevice(c->device);
    stk_status st = devUpload(c, c->mats, m, (size_t)n * sizeof(StkMaterial));
    if (st == STK_OK) c->numMats = n;
    return st;
}

STK_API stk_status stk_set_textures(stk_ctx* c, const StkTexture* t, uint32_t n, const float* texels,
                                    uint64_t numTexels)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if ((n && !t) || (numTexels && !texels)) return fail(c, STK_E_INVALID, "stk_set_textures: null buffer");
    if (numTexels > (1ull << 31)) return fail(c, STK_E_INVALID, "stk_set_textures: too many texels");
    for (uint32_t i = 0; i < n; ++i) {
        const StkTexture& x = t[i];
        if (x.width == 0 || x.height == 0 || x.width > 16384 || x.height > 16384 || x.wrap > STK_TEX_CLIP ||
            (uint64_t)x.offset + (uint64_t)x.width * x.height > numTexels)
            return fail(c, STK_E_INVALID, "stk_set_textures: texture " + std::to_string(i) +
                                              " empty, too large, with an unknown wrap or outside the texels");
    }
    (void)hipSetDevice(c->device);
    invalidateGraph(c);
    c->numTextures = 0;
    stk_status st;
    if ((st = devUpload(c, c->textures, t, (size_t)n * sizeof(StkTexture))) != STK_OK) return st;
    if ((st = devUpload(c, c->texels, texels, (size_t)numTexels * 4 * sizeof(float))) != STK_OK) return st;
    c->numTextures = n;
    return STK_OK;
}

STK_API stk_status stk_build_bvh(stk_ctx* c, const StkBvhOptions* opts)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveGeometry) return fail(c, STK_E_STATE, "stk_build_bvh: no geometry set");
    std::vector<stk::BvhInputTri> tris;
    std::string err;
    const StkBvhOptions o = opts ? *opts : stk::defaultBvhOptions();
    // built aside: a failed build leaves the context's tree as it was
    stk::Bvh b;
    try {
        if (!stk::flattenInstances(c->hvb.data(), (uint32_t)c->hvb.size(), c->hib.data(), (uint32_t)c->hib.size(),
                                   c->hinst.data(), (uint32_t)c->hinst.size(), tris, err) ||
            !stk::buildBvh(tris, o, b, err) ||
            !stk::packBvh(c->hvb.data(), (uint32_t)c->hvb.size(), c->hib.data(), (uint32_t)c->hib.size(),
                          c->hinst.data(), (uint32_t)c->hinst.size(), b, err) ||
            !stk::buildWide(c->hinst.data(), (uint32_t)c->hinst.size(), b, err))
            return fail(c, STK_E_INVALID, "stk_build_bvh: " + err);
    } catch (const std::bad_alloc&) {
        return fail(c, STK_E_OOM, "stk_build_bvh: host allocation failed");
    }
    if (b.depth > (uint32_t)stk::dev::kStackDepth)
        return fail(c, STK_E_INVALID, "stk_build_bvh: tree deeper than the traversal stack (32)");
    (void)hipSetDevice(c->device);
    invalidateGraph(c);
    return uploadBvh(c, std::move(b));
}

// The renderer's one-time scene read (src/render/ptrender.cpp:656-667 and
// createInstanceBuffer, 560-601): the container's global buffers, instance
// records, lights and materials go to the device and the flat BVH is built
// over them.  The scene may be freed or changed afterwards; a later change
// takes effect at the next stk_set_scene.
STK_API stk_status stk_set_scene(stk_ctx* c, const stk_scene* s, const StkBvhOptions* opts)
{
    if (!validCtx(c)) return STK_E_INVALID;
    StkSceneArrays a;
    stk_status st = stk_scene_get(s, &a);
    if (st != STK_OK) return fail(c, st, "stk_set_scene: null scene or host allocation failed");
    for (uint32_t i = 0; i < a.numInstances; ++i)
        if (a.inst[i].materialId >= (int32_t)a.numMaterials)
            return fail(c, STK_E_INVALID, "stk_set_scene: instance " + std::to_string(i) +
                                              " names material " + std::to_string(a.inst[i].materialId) +
                                              " of " + std::to_string(a.numMaterials));
    if ((st = stk_set_geometry(c, a.vb, a.numVertices, a.ib, a.numIndices, a.inst, a.numInstances)) != STK_OK)
        return st;
    if ((st = stk_set_lights(c, a.lights, a.numLights)) != STK_OK) return st;
    if ((st = stk_set_materials(c, a.mats, a.numMaterials)) != STK_OK) return st;
    return stk_build_bvh(c, opts);
}

STK_API stk_status stk_build_two_level(stk_ctx* c, const StkBvhOptions* opts)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveGeometry) return fail(c, STK_E_STATE, "stk_build_two_level: no geometry set");
    std::string err;
    const StkBvhOptions o = opts ? *opts : stk::defaultBvhOptions();
    // built aside: a failed build leaves the context's structure as it was
    stk::TwoLevel t;
    try {
        if (!stk::buildTwoLevel(c->hvb.data(), (uint32_t)c->hvb.size(), c->hib.data(), (uint32_t)c->hib.size(),
                                c->hinst.data(), (uint32_t)c->hinst.size(), o, t, err))
            return fail(c, STK_E_INVALID, "stk_build_two_level: " + err);
    } catch (const std::bad_alloc&) {
        return fail(c, STK_E_OOM, "stk_build_two_level: host allocation failed");
    }
    if (t.tlasWide.empty()) return fail(c, STK_E_INVALID, "stk_build_two_level: no instance has triangles");
    if (t.tlasDepth + t.blasDepth > (uint32_t)stk::dev::kWideStack)
        return fail(c, STK_E_INVALID, "stk_build_two_level: instance and mesh trees deeper than the walk's stack (16)");
    (void)hipSetDevice(c->device);
    // from here on the device arrays change: without a tree until all are written
    c->haveBvh = false;
    c->twoLevel = false;
    c->tlWorld = false; // the world triangles were the previous structure's
    c->numWide = 0;
    stk_status st;
    if ((st = devUpload(c, c->tlasWide, t.tlasWide.data(), t.tlasWide.size() * sizeof(StkWideNode))) != STK_OK)
        return st;
    if ((st = devUpload(c, c->tlasLeaves, t.tlasLeaves.data(), t.tlasLeaves.size() * sizeof(StkTlasLeaf))) != STK_OK)
        return st;
    if ((st = devUpload(c, c->blasWide, t.blasWide.data(), t.blasWide.size() * sizeof(StkWideNode))) != STK_OK)
        return st;
    if ((st = devUpload(c, c->blasTris, t.blasTris.data(), t.blasTris.size() * sizeof(StkBlasTriangle))) != STK_OK)
        return st;
    if ((st = devUpload(c, c->instKey, t.instKey.data(), t.instKey.size() * sizeof(uint32_t))) != STK_OK) return st;
    if ((st = devUpload(c, c->instW, t.instW.data(), t.instW.size() * sizeof(float))) != STK_OK) return st;
    if ((st = devUpload(c, c->tlEmit, t.emitters.data(), t.emitters.size() * sizeof(StkTriangle))) != STK_OK)
        return st;
    c->tl = std::move(t);
    // the flat tree goes; its header keeps the scene box (the cache's domain)
    c->bvh = stk::Bvh{};
    c->bvh.header.root = INT32_MIN;
    for (int a = 0; a < 3; ++a) {
        c->bvh.header.lo[a] = c->tl.lo[a];
        c->bvh.header.hi[a] = c->tl.hi[a];
    }
    c->gpuBuilt = false;
    c->twoLevel = true;
    c->haveBvh = true;
    invalidateGraph(c);
    return STK_OK;
}

STK_API stk_status stk_two_level_world_triangles(stk_ctx* c, int enable)
{
    if (!validCtx(c)) return STK_E_INVALID;
    (void)hipSetDevice(c->device);
    c->tlWorld = false;
    invalidateGraph(c);
    stk_status st;
    if (!enable) {
        if ((st = devRelease(c, c->tlWorldTris)) != STK_OK) return st;
        return devRelease(c, c->tlTriOff);
    }
    if (!c->twoLevel) return fail(c, STK_E_STATE, "stk_two_level_world_triangles: no two-level structure");
    const stk::TwoLevel& t = c->tl;
    const uint32_t ninst = (uint32_t)t.instMesh.size();
    std::vector<uint32_t> meshBase(ninst, 0u), count(ninst, 0u), off(ninst, 0u);
    std::vector<uint64_t> worldBase(ninst, 0u);
    uint64_t total = 0;
    for (uint32_t i = 0; i < ninst; ++i) {
        const uint32_t m = t.instMesh[i];
        if (m == STK_INVALID_INDEX) continue;
        meshBase[i] = t.meshTriBase[m];
        count[i] = t.meshTriCount[m];
        worldBase[i] = total;
        off[i] = (uint32_t)(total - meshBase[i]); // modulo 2^32: the walk adds it to a blasTris index
        total += count[i];
    }
    // the walk indexes the array with 32-bit triangle numbers
    if (total >= (1ull << 31))
        return fail(c, STK_E_INVALID, "stk_two_level_world_triangles: more than 2^31 instanced triangles");
    DevBuf dMeshBase, dCount, dWorldBase;
    auto release = [&]() {
        (void)devRelease(c, dMeshBase);
        (void)devRelease(c, dCount);
        (void)devRelease(c, dWorldBase);
    };
    if ((st = devReserve(c, c->tlWorldTris, (size_t)total * sizeof(StkTriangle))) != STK_OK ||
        (st = devUpload(c, c->tlTriOff, off.data(), off.size() * sizeof(uint32_t))) != STK_OK ||
        (st = devUpload(c, dMeshBase, meshBase.data(), meshBase.size() * sizeof(uint32_t))) != STK_OK ||
        (st = devUpload(c, dCount, count.data(), count.size() * sizeof(uint32_t))) != STK_OK ||
        (st = devUpload(c, dWorldBase, worldBase.data(), worldBase.size() * sizeof(uint64_t))) != STK_OK) {
        release();
        return st;
    }
    hipError_t e = stk::launchWorldTriangles(static_cast<const StkInstance*>(c->inst.ptr),
                                             static_cast<const float4*>(c->blasTris.ptr),
                                             static_cast<const uint32_t*>(dMeshBase.ptr),
                                             static_cast<const uint32_t*>(dCount.ptr),
                                             static_cast<const uint64_t*>(dWorldBase.ptr), ninst,
                                             static_cast<float4*>(c->tlWorldTris.ptr), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    release();
    if (e != hipSuccess) return hipFail(c, e, "stk_two_level_world_triangles");
    c->tlWorld = true;
    return STK_OK;
}

STK_API stk_status stk_set_bvh(stk_ctx* c, const StkBvhNode* nodes, uint32_t n)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveGeometry) return fail(c, STK_E_STATE, "stk_set_bvh: no geometry set");
    if (n && !nodes) return fail(c, STK_E_INVALID, "stk_set_bvh: null buffer");
    std::string err;
    // validated aside: a rejected tree leaves the context's tree as it was
    stk::Bvh b;
    try {
        b.nodes.assign(nodes, nodes + n);
        if (!stk::packBvh(c->hvb.data(), (uint32_t)c->hvb.size(), c->hib.data(), (uint32_t)c->hib.size(),
                          c->hinst.data(), (uint32_t)c->hinst.size(), b, err) ||
            !stk::buildWide(c->hinst.data(), (uint32_t)c->hinst.size(), b, err))
            return fail(c, STK_E_INVALID, "stk_set_bvh: " + err);
    } catch (const std::bad_alloc&) {
        return fail(c, STK_E_OOM, "stk_set_bvh: host allocation failed");
    }
    if (b.depth > (uint32_t)stk::dev::kStackDepth)
        return fail(c, STK_E_INVALID, "stk_set_bvh: tree deeper than the traversal stack (32)");
    (void)hipSetDevice(c->device);
    invalidateGraph(c);
    return uploadBvh(c, std::move(b));
}

STK_API stk_status stk_get_bvh(stk_ctx* c, StkBvhNode* nodes, uint32_t cap, uint32_t* n)
{
    if (!validCtx(c) || !n) return STK_E_INVALID;
    if (!c->haveBvh) return fail(c, STK_E_STATE, "stk_get_bvh: no BVH");
    if (c->twoLevel) return fail(c, STK_E_STATE, "stk_get_bvh: the context holds a two-level structure");
    if (c->gpuBuilt) {
        // the device copy: leaves carry their triangle index in maxBounds[0],
        // which the reference layout leaves zero
        *n = c->numNodes;
        if (!nodes) return STK_OK;
        if (cap < *n) return fail(c, STK_E_INVALID, "stk_get_bvh: capacity too small");
        (void)hipSetDevice(c->device);
        STK_HIP(c, hipMemcpyAsync(nodes, c->refNodes.ptr, (size_t)*n * sizeof(StkBvhNode), hipMemcpyDeviceToHost,
                                  c->stream));
        STK_HIP(c, hipStreamSynchronize(c->stream));
        for (uint32_t i = 0; i < *n; ++i)
            if ((uint32_t)nodes[i].instId != STK_INVALID_INDEX) nodes[i].maxBounds[0] = 0.0f;
        return STK_OK;
    }
    *n = (uint32_t)c->bvh.nodes.size();
    if (!nodes) return STK_OK;
    if (cap < *n) return fail(c, STK_E_INVALID, "stk_get_bvh: capacity too small");
    memcpy(nodes, c->bvh.nodes.data(), c->bvh.nodes.size() * sizeof(StkBvhNode));
    return STK_OK;
}

STK_API stk_status stk_get_bvh_info(stk_ctx* c, StkBvhInfo* out)
{
    if (!validCtx(c) || !out) return STK_E_INVALID;
    if (!c->haveBvh) return fail(c, STK_E_STATE, "stk_get_bvh_info: no BVH");
    memset(out, 0, sizeof(*out));
    out->header = c->bvh.header;
    out->depth = c->bvh.depth;
    out->wideDepth = c->twoLevel ? c->tl.tlasDepth + c->tl.blasDepth : c->bvh.wideDepth;
    out->numNodes = c->numNodes;
    out->numWide = c->twoLevel ? (uint32_t)(c->tl.tlasWide.size() + c->tl.blasWide.size()) : c->numWide;
    out->numEmitters = c->twoLevel ? (uint32_t)c->tl.emitters.size() : c->numEmit;
    out->twoLevel = c->twoLevel ? 1u : 0u;
    out->gpuBuilt = c->gpuBuilt ? 1u : 0u;
    out->walkStack = stk::closestWalkStack(bvhView(c));
    return STK_OK;
}

STK_API stk_status stk_build_bvh_gpu(stk_ctx* c) { return stk_build_bvh_gpu_method(c, STK_GPU_BUILD_LBVH); }

STK_API stk_status stk_build_bvh_gpu_method(stk_ctx* c, uint32_t method)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (method != STK_GPU_BUILD_LBVH && method != STK_GPU_BUILD_PLOC)
        return fail(c, STK_E_INVALID, "stk_build_bvh_gpu: unknown method");
    if (!c->haveGeometry) return fail(c, STK_E_STATE, "stk_build_bvh_gpu: no geometry set");
    const uint32_t ninst = (uint32_t)c->hinst.size();
    std::vector<uint32_t> start(ninst + 1, 0);
    uint64_t total = 0;
    for (uint32_t i = 0; i < ninst; ++i) {
        start[i] = (uint32_t)total;
        total += (uint64_t)(c->hinst[i].indexCount / 3);
    }
    if (total >= (1ull << 30)) return fail(c, STK_E_INVALID, "stk_build_bvh_gpu: too many triangles");
    start[ninst] = (uint32_t)total;
    (void)hipSetDevice(c->device);
    std::vector<DevBuf> temps;
    auto alloc = [&](int slot, size_t bytes) -> void* {
        DevBuf* b = nullptr;
        switch (slot) {
        case stk::kGbRefNodes: b = &c->refNodes; break;
        case stk::kGbPacked: b = &c->packed; break;
        case stk::kGbTris: b = &c->tris; break;
        case stk::kGbWide: b = &c->wideNodes; break;
        case stk::kGbWideTris: b = &c->wideTris; break;
        case stk::kGbEmitters: b = &c->emitters; break;
        default: {
            const size_t k = (size_t)(slot - stk::kGbFirstTemp);
            if (temps.size() <= k) temps.resize(k + 1);
            b = &temps[k];
        }
        }
        if (devReserve(c, *b, std::max<size_t>(bytes, 16)) != STK_OK) return nullptr;
        return b->ptr;
    };
    stk::GpuBvhResult r;
    std::string err;
    c->haveBvh = false;
    const hipError_t e = stk::gpuBuildBvh(static_cast<const StkVertex*>(c->vb.ptr), static_cast<const uint32_t*>(c->ib.ptr),
                                          static_cast<const StkInstance*>(c->inst.ptr), start.data(), ninst,
                                          (uint32_t)total, method, alloc, r, err, c->stream);
    (void)hipStreamSynchronize(c->stream);
    for (DevBuf& b : temps) (void)devRelease(c, b);
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) return fail(c, STK_E_OOM, "stk_build_bvh_gpu: " + err);
        if (!err.empty()) return fail(c, STK_E_INVALID, "stk_build_bvh_gpu: " + err);
        return hipFail(c, e, "stk_build_bvh_gpu");
    }
    if (total > 0 && r.wideDepth > (uint32_t)stk::dev::kWideStack)
        return fail(c, STK_E_INVALID, "stk_build_bvh_gpu: wide tree deeper than the walk's stack (16)");
    c->bvh = stk::Bvh{};
    c->bvh.header = r.header;
    c->bvh.depth = r.depth;
    c->bvh.wideDepth = r.wideDepth;
    c->numNodes = r.numNodes;
    c->numWide = r.numWide;
    c->numEmit = r.numEmit;
    c->gpuBuilt = true;
    c->twoLevel = false;
    c->haveBvh = true;
    invalidateGraph(c);
    return STK_OK;
}

// The reference-order and binary walks need the flat tree.
static stk_status flatOnlyFlags(stk_ctx* c, uint32_t flags)
{
    if (c->twoLevel && (flags & (STK_TRACE_REFERENCE_ORDER | STK_TRACE_BVH2)))
        return fail(c, STK_E_STATE, "trace: the reference-order and binary walks need the flat BVH (stk_build_bvh)");
    if (!c->twoLevel && c->bvh.depth > (uint32_t)stk::dev::kStackDepth && (flags & STK_TRACE_BVH2))
        return fail(c, STK_E_STATE, "trace: the tree is deeper than the binary walk's stack (32)");
    return STK_OK;
}

STK_API stk_status stk_trace_closest_device(stk_ctx* c, const StkRay* rays, uint32_t n, StkHit* hits, uint32_t flags)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveBvh) return fail(c, STK_E_STATE, "stk_trace_closest: no BVH");
    if (n && (!rays || !hits)) return fail(c, STK_E_INVALID, "stk_trace_closest: null buffer");
    if (flatOnlyFlags(c, flags) != STK_OK) return STK_E_STATE;
    (void)hipSetDevice(c->device);
    STK_HIP(c, stk::launchTraceClosest(bvhView(c), rays, n, hits, flags, c->stream));
    return STK_OK;
}

STK_API stk_status stk_trace_closest_image_device(stk_ctx* c, const StkRay* rays, uint32_t width, uint32_t height,
                                                  StkHit* hits, uint32_t flags)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveBvh) return fail(c, STK_E_STATE, "stk_trace_closest_image: no BVH");
    const uint64_t n = (uint64_t)width * height;
    if (n > 0xFFFFFFFFull) return fail(c, STK_E_INVALID, "stk_trace_closest_image: image too large");
    if (n && (!rays || !hits)) return fail(c, STK_E_INVALID, "stk_trace_closest_image: null buffer");
    if (flatOnlyFlags(c, flags) != STK_OK) return STK_E_STATE;
    (void)hipSetDevice(c->device);
    STK_HIP(c, stk::launchTraceClosest(bvhView(c), rays, (uint32_t)n, hits, flags, c->stream, width));
    return STK_OK;
}

STK_API stk_status stk_trace_any_device(stk_ctx* c, const StkRay* rays, uint32_t n, StkHit* hits, uint32_t flags)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveBvh) return fail(c, STK_E_STATE, "stk_trace_any: no BVH");
    if (n && (!rays || !hits)) return fail(c, STK_E_INVALID, "stk_trace_any: null buffer");
    if (flatOnlyFlags(c, flags) != STK_OK) return STK_E_STATE;
    (void)hipSetDevice(c->device);
    STK_HIP(c, stk::launchTraceAny(bvhView(c), rays, n, hits, flags, c->stream));
    return STK_OK;
}

static stk_status traceHost(stk_ctx* c, const StkRay* rays, uint32_t n, StkHit* hits, int any, uint32_t flags)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveBvh) return fail(c, STK_E_STATE, "trace: no BVH");
    if (flatOnlyFlags(c, flags) != STK_OK) return STK_E_STATE;
    if (n == 0) return STK_OK;
    if (!rays || !hits) return fail(c, STK_E_INVALID, "trace: null buffer");
    (void)hipSetDevice(c->device);
    stk_status st;
    if ((st = devReserve(c, c->rayScratch, (size_t)n * sizeof(StkRay))) != STK_OK) return st;
    if ((st = devReserve(c, c->hitScratch, (size_t)n * sizeof(StkHit))) != STK_OK) return st;
    STK_HIP(c, hipMemcpyAsync(c->rayScratch.ptr, rays, (size_t)n * sizeof(StkRay), hipMemcpyHostToDevice, c->stream));
    StkRay* dr = static_cast<StkRay*>(c->rayScratch.ptr);
    StkHit* dh = static_cast<StkHit*>(c->hitScratch.ptr);
    if (any) {
        STK_HIP(c, stk::launchTraceAny(bvhView(c), dr, n, dh, flags, c->stream));
    } else {
        STK_HIP(c, stk::launchTraceClosest(bvhView(c), dr, n, dh, flags, c->stream));
    }
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

STK_API stk_status stk_trace_closest_noalloc(stk_ctx* c, const StkRay* rays, uint32_t n, StkHit* hits, uint32_t flags)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (!c->haveBvh) return fail(c, STK_E_STATE, "stk_trace_closest_noalloc: no BVH");
    if (flatOnlyFlags(c, flags) != STK_OK) return STK_E_STATE;
    if (n > kNoAllocMaxRays)
        return fail(c, STK_E_INVALID, "stk_trace_closest_noalloc: more rays than the static staging buffer holds "
                                      "(1048576); split the batch or use stk_trace_closest");
    if (n == 0) return STK_OK;
    if (!rays || !hits) return fail(c, STK_E_INVALID, "stk_trace_closest_noalloc: null buffer");
    void* base = (c->device >= 0 && c->device < kMaxScratchDevices) ? g_scratchPtr[c->device].load() : nullptr;
    if (!base)
        return fail(c, STK_E_DEVICE, std::string("stk_trace_closest_noalloc: no static staging buffer on this device: ") +
                                         hipGetErrorString(c->device < kMaxScratchDevices ? g_scratchErr[c->device]
                                                                                          : hipErrorInvalidDevice));
    (void)hipSetDevice(c->device);
    // every ray is copied in and every hit written by the kernel: nothing of a
    // previous call's contents is read
    StkRay* dr = static_cast<StkRay*>(base);
    StkHit* dh = reinterpret_cast<StkHit*>(static_cast<unsigned char*>(base) + kNoAllocRayBytes);
    STK_HIP(c, hipMemcpyAsync(dr, rays, (size_t)n * sizeof(StkRay), hipMemcpyHostToDevice, c->stream));
    STK_HIP(c, stk::launchTraceClosest(bvhView(c), dr, n, dh, flags, c->stream));
    STK_HIP(c, hipMemcpyAsync(hits, dh, (size_t)n * sizeof(StkHit), hipMemcpyDeviceToHost, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    return STK_OK;
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

// ---- rendering ------------------------------------------------------------------

static stk_status ensurePathState(stk_ctx* c, stk::PathState& ps, std::vector<DevBuf>& bufs, size_t slots)
{
    const size_t qcap = (size_t)stk::kStripes * stk::stripeSegment(slots);
    if (ps.capacity >= slots && ps.queueCapacity >= qcap) return STK_OK;
    invalidateGraph(c);
    stk_status st;
    for (hipStream_t s : c->xStream)
        if (s) (void)hipStreamSynchronize(s);
    for (DevBuf& b : bufs) devRelease(c, b);
    bufs.assign(20, DevBuf{});
    ps.capacity = 0;
    ps.queueCapacity = 0;
    const size_t f4 = slots * sizeof(float4), u4 = slots * sizeof(uint32_t);
    const size_t qf4 = qcap * sizeof(float4), qu4 = qcap * sizeof(uint32_t);
    void** ptrs[] = {(void**)&ps.rayO,     (void**)&ps.rayD,     (void**)&ps.thr,      (void**)&ps.col,
                     (void**)&ps.rng,      (void**)&ps.scene,    (void**)&ps.pixel,    (void**)&ps.hit,
                     (void**)&ps.hitPrim,  (void**)&ps.queue[0], (void**)&ps.queue[1], (void**)&ps.shadowQ,
                     (void**)&ps.shO,      (void**)&ps.shD,      (void**)&ps.shCd,     (void**)&ps.shCg,
                     (void**)&ps.counters, (void**)&ps.pass,     (void**)&ps.stripes,  (void**)&ps.work};
    const size_t sizes[] = {f4, f4, f4, f4, u4, u4, u4, qf4, qu4, qu4, qu4, qu4, qf4, qf4, qf4, qf4, 64, 64,
                            stk::kQueueKinds * stk::kStripes * sizeof(uint32_t), stk::kStripes * sizeof(uint32_t)};
    static_assert(sizeof(sizes) / sizeof(sizes[0]) == sizeof(ptrs) / sizeof(ptrs[0]), "path state buffers");
    for (size_t i = 0; i < sizeof(sizes) / sizeof(sizes[0]); ++i) {
        if ((st = devReserve(c, bufs[i], sizes[i])) != STK_OK) return st;
        *ptrs[i] = bufs[i].ptr;
    }
    ps.capacity = (uint32_t)slots;
    ps.queueCapacity = qcap;
    return STK_OK;
}

// Image buffers; the accumulation restarts when the image layout changes.
static stk_status ensureImage(stk_ctx* c, uint64_t key, size_t npix, size_t nsamples)
{
    stk_status st;
    const size_t before[4] = {(size_t)c->history.ptr, (size_t)c->accum.ptr, (size_t)c->out.ptr,
                              (size_t)c->samples.ptr};
    if (c->imageKey != key) {
        c->iteration = 0;
        c->imageKey = key;
        if ((st = devReserve(c, c->history, npix * sizeof(float4))) != STK_OK) return st;
        STK_HIP(c, stk::launchZero(static_cast<float4*>(c->history.ptr), npix, c->stream));
    }
    if ((st = devReserve(c, c->accum, npix * sizeof(float4))) != STK_OK) return st;
    if ((st = devReserve(c, c->out, npix * sizeof(float4))) != STK_OK) return st;
    if (nsamples && (st = devReserve(c, c->samples, nsamples * sizeof(float))) != STK_OK) return st;
    c->samplesCount = nsamples;
    c->imagePixels = npix;
    const size_t after[4] = {(size_t)c->history.ptr, (size_t)c->accum.ptr, (size_t)c->out.ptr,
                             (size_t)c->samples.ptr};
    if (memcmp(before, after, sizeof(before)) != 0) invalidateGraph(c);
    return STK_OK;
}

// Renders p->pt.spp samples of every view in `descs` (one view for
// stk_render, many for stk_batch_render) over the forest in `base`.
static stk_status renderForestRun(stk_ctx* c, const stk::SceneView& base, std::vector<stk::SceneDesc>& descs,
                               const StkRenderParams* p, float* outRgba, int outIsDevice, bool allowCache,
                               uint64_t imageKey)
{
    stk_status st;
    if (p->pt.spp <= 0 || p->pt.maxDepth > 64) return fail(c, STK_E_INVALID, "render: bad spp or maxDepth");
    const uint32_t T = std::max<uint32_t>(1u, c->launch.tile);
    size_t slots = 0, npix = 0;
    for (stk::SceneDesc& d : descs) {
        if (d.width == 0 || d.height == 0) return fail(c, STK_E_INVALID, "render: empty view");
        d.tilesX = (d.width + T - 1) / T;
        const size_t tilesY = (d.height + T - 1) / T;
        d.slotBase = (uint32_t)slots;
        d.pixelBase = (uint32_t)npix;
        d.cam.dimension[0] = (int32_t)d.width;
        d.cam.dimension[1] = (int32_t)d.height;
        d.cam.spp = p->pt.spp;
        d.cam.stratifiedSamplingType = p->pt.stratifiedSamplingType;
        slots += (size_t)d.tilesX * T * tilesY * T;
        npix += (size_t)d.width * d.height;
    }
    if (slots > (1ull << 30) || npix > (1ull << 29)) return fail(c, STK_E_INVALID, "render: image too large");
    (void)hipSetDevice(c->device);
    const size_t nsamples = p->keepSamples ? npix * (size_t)p->pt.spp * 3 : 0;
    // paths in flight per pass: a frame's slots are traced in passes of at most
    // Q slots (whole waves); the cache trains on whole frames, so it takes one pass
    size_t Q = slots;
    if (c->launch.pathsInFlight && p->cacheMode == 0)
        Q = std::min(slots, ((size_t)c->launch.pathsInFlight + 63) / 64 * 64);
    // one persistent launch per frame (plain rendering): its workgroups hold
    // 256 slots each of the path state, on the context stream
    const bool frameKernel = c->launch.frameBlocks != 0 && p->cacheMode == 0;
    if (frameKernel) Q = (size_t)c->launch.frameBlocks * (c->launch.frameChunk ? c->launch.frameChunk : 256u);
    // several pass streams (with the cache, a pass's bounces overlap the
    // previous pass's inference and training, which its own wait for)
    const uint32_t nstreams = frameKernel ? 1u : std::max<uint32_t>(1u, c->launch.passStreams);
    const bool multi = nstreams > 1;
    if (!multi && (st = ensurePathState(c, c->ps, c->psBufs, Q)) != STK_OK) return st;
    for (uint32_t i = 0; multi && i < nstreams; ++i) {
        if ((st = ensurePathState(c, c->xPs[i], c->xPsBufs[i], Q)) != STK_OK) return st;
        if (!c->xStream[i]) STK_HIP(c, hipStreamCreateWithFlags(&c->xStream[i], hipStreamNonBlocking));
    }
    if (multi && !c->finEv[0]) {
        for (int i = 0; i < kMaxPassStreams; ++i) {
            STK_HIP(c, hipEventCreateWithFlags(&c->finEv[i], hipEventDisableTiming));
            STK_HIP(c, hipEventCreateWithFlags(&c->joinEv[i], hipEventDisableTiming));
            STK_HIP(c, hipEventCreateWithFlags(&c->cacheEv[i], hipEventDisableTiming));
        }
        for (auto& ev : c->accumFree) STK_HIP(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    // multi-stream calls alternate between the two accumulation buffers
    const uint32_t lane = multi ? (c->renderCalls & 1u) : 0u;
    DevBuf& accBuf = lane ? c->accum2 : c->accum;
    if ((st = ensureImage(c, imageKey, npix, nsamples)) != STK_OK) return st;
    if ((st = devReserve(c, accBuf, npix * sizeof(float4))) != STK_OK) return st;
    // the views' records; the kernels take the frame number from the pass
    // state, so the records of a progressive render do not change per frame
    // and are uploaded (a synchronising copy) only when they do
    for (stk::SceneDesc& d : descs) d.cam.frameNumber = 0;
    const size_t descBytes = descs.size() * sizeof(stk::SceneDesc);
    if (c->descs.bytes < descBytes) invalidateGraph(c);
    if (!c->descs.ptr || c->descsHost.size() != descs.size() ||
        memcmp(c->descsHost.data(), descs.data(), descBytes) != 0) {
        c->descsHost.clear();
        if ((st = devUpload(c, c->descs, descs.data(), descBytes)) != STK_OK) return st;
        c->descsHost = descs;
    }
    stk::SceneView sv = base;
    sv.scenes = static_cast<const stk::SceneDesc*>(c->descs.ptr);
    sv.numScenes = (uint32_t)descs.size();
    sv.totalSlots = (uint32_t)slots;
    sv.tile = T;
    stk::PtLaunch lp{};
    lp.maxDepth = p->pt.maxDepth;
    lp.spp = (uint32_t)p->pt.spp;
    lp.samplingType = p->pt.stratifiedSamplingType;
    lp.rrStartDepth = p->rrStartDepth;
    lp.refCompat = p->refCompat;
    lp.invSpp = 1.0f / (float)p->pt.spp;
    lp.passSlots = (uint32_t)Q;
    c->ps.segCap = stk::stripeSegment(Q);
    for (stk::PathState& x : c->xPs) x.segCap = c->ps.segCap;
    stk::PtCache& pc = lp.cache;
    if (p->cacheMode != 0) {
        if (!allowCache) return fail(c, STK_E_INVALID, "render: the radiance cache serves single-scene renders");
        if (!c->haveCache) return fail(c, STK_E_STATE, "render: cacheMode set but no cache (stk_cache_create)");
        if (p->cacheMode > 4) return fail(c, STK_E_INVALID, "render: cacheMode must be 0..4");
        pc.mode = p->cacheMode;
        pc.queryDepth = p->cacheQueryDepth;
        if (!(p->cacheMinRoughness >= 0.0f && p->cacheMinRoughness <= 1.0f))
            return fail(c, STK_E_INVALID, "render: cacheMinRoughness must be in [0, 1]");
        pc.queryMinRough = p->cacheMinRoughness;
        if (p->cacheTrainSteps > 64) return fail(c, STK_E_INVALID, "render: cacheTrainSteps must be at most 64");
        pc.trainSteps = p->cacheTrainSteps ? p->cacheTrainSteps : 1u;
        if (p->cacheTrainSuffix > 64) return fail(c, STK_E_INVALID, "render: cacheTrainSuffix must be at most 64");
        pc.trainSuffix = p->cacheMode == 1 ? p->cacheTrainSuffix : 0u;
        if (p->cacheQueryDirect > 1) return fail(c, STK_E_INVALID, "render: cacheQueryDirect must be 0 or 1");
        pc.queryDirect = p->cacheQueryDirect;
        const uint32_t paths = p->cacheTrainPaths ? p->cacheTrainPaths : 65536u;
        pc.trainStride = std::max<uint32_t>(1u, (uint32_t)(slots / paths));
        pc.maxTrainPaths = (uint32_t)((slots + pc.trainStride - 1) / pc.trainStride);
        pc.trainCap = (uint32_t)std::min<size_t>(c->cache.nCap, (size_t)pc.maxTrainPaths * p->pt.maxDepth);
        // per pass stream (one set on a single stream): queries, training vertices
        for (uint32_t i = 0; i < nstreams; ++i) {
            DevBuf& qr = multi ? c->xQueryRecs[i] : c->queryRecs;
            DevBuf& qo = multi ? c->xQueryOut[i] : c->queryOut;
            DevBuf& tv = multi ? c->xTrainVerts[i] : c->trainVerts;
            DevBuf& tc = multi ? c->xTrainCount[i] : c->trainCount;
            if ((st = devReserve(c, qr, slots * sizeof(stk::CacheRecord))) != STK_OK) return st;
            if ((st = devReserve(c, qo, slots * sizeof(float4))) != STK_OK) return st;
            if ((st = devReserve(c, tv, (size_t)pc.maxTrainPaths * p->pt.maxDepth * 6 * sizeof(float4))) != STK_OK)
                return st;
            if ((st = devReserve(c, tc, (size_t)pc.maxTrainPaths * sizeof(uint32_t))) != STK_OK) return st;
        }
        if ((st = devReserve(c, c->trainBatch, (size_t)pc.trainCap * sizeof(stk::CacheRecord))) != STK_OK) return st;
        c->lastTrainCap = pc.trainCap;
        pc.queryRecs = static_cast<stk::CacheRecord*>((multi ? c->xQueryRecs[0] : c->queryRecs).ptr);
        pc.queryOut = static_cast<float4*>((multi ? c->xQueryOut[0] : c->queryOut).ptr);
        pc.trainVerts = static_cast<float4*>((multi ? c->xTrainVerts[0] : c->trainVerts).ptr);
        pc.trainCount = static_cast<uint32_t*>((multi ? c->xTrainCount[0] : c->trainCount).ptr);
        pc.trainBatch = static_cast<stk::CacheRecord*>(c->trainBatch.ptr);
    }
    float4* accum = static_cast<float4*>(accBuf.ptr);
    const uint32_t first = c->passTurn % std::max(1u, nstreams); // pass stream of this call's first pass
    if (!multi) {
        STK_HIP(c, stk::launchZero(accum, npix, c->stream));
        STK_HIP(c, stk::launchClearRayTotals(c->ps, c->stream));
    } else {
        // After a single-stream call the pass streams wait for the context
        // stream; after a multi-stream call they go on at once (what a pass
        // reads is uploaded with a synchronising copy; what it writes is its
        // stream's path state, this call's accumulation buffer and the
        // samples, ordered below).
        if (c->lastStreams == 1) {
            STK_HIP(c, hipEventRecord(c->joinEv[0], c->stream));
            for (uint32_t i = 0; i < nstreams; ++i) STK_HIP(c, hipStreamWaitEvent(c->xStream[i], c->joinEv[0], 0));
        }
        for (uint32_t i = 0; i < nstreams; ++i) STK_HIP(c, stk::launchClearRayTotals(c->xPs[i], c->xStream[i]));
        // the buffer is free once the call before last has tonemapped it
        if (c->accumFreeSet[lane]) STK_HIP(c, hipStreamWaitEvent(c->xStream[first], c->accumFree[lane], 0));
        STK_HIP(c, stk::launchZero(accum, npix, c->xStream[first]));
    }
    float* samples = nsamples ? static_cast<float*>(c->samples.ptr) : nullptr;
    // One sample pass is captured once as a graph and replayed per sample (the
    // pass number and frame come from device memory).  Training passes of the
    // cache keep host-side Adam state and run eagerly.
    const bool graph = c->useGraph && pc.mode == 0 && !multi && !frameKernel;
    if (graph) {
        // everything the captured kernels take by value
        std::vector<uint64_t> key = {(uint64_t)sv.bvh.packed, (uint64_t)sv.bvh.tris, (uint64_t)sv.bvh.refNodes,
                                     (uint64_t)sv.bvh.wide, (uint64_t)sv.bvh.wideTris, (uint64_t)sv.bvh.emitters,
                                     (uint64_t)sv.bvh.instances, (uint64_t)sv.bvh.tlas, (uint64_t)sv.bvh.tlasLeaves,
                                     (uint64_t)sv.bvh.blas, (uint64_t)sv.bvh.blasTris, (uint64_t)sv.bvh.instKey,
                                     (uint64_t)sv.bvh.instW, (uint64_t)sv.vb, (uint64_t)sv.ib,
                                     (uint64_t)sv.inst, (uint64_t)sv.lights, (uint64_t)sv.mats, sv.numMats,
                                     (uint64_t)sv.textures, sv.numTextures, (uint64_t)sv.texels,
                                     (uint64_t)sv.scenes, sv.numScenes, sv.totalSlots, sv.tile, lp.maxDepth,
                                     lp.spp, lp.samplingType, lp.rrStartDepth, lp.refCompat, lp.passSlots,
                                     c->launch.blockSize, c->launch.tile, c->launch.persistentBlocks, c->ps.segCap,
                                     c->launch.binaryWalk, c->launch.stripeThreads,
                                     (uint64_t)accum, (uint64_t)samples, (uint64_t)c->ps.rayO, c->ps.capacity};
        if (!c->graphExec || key != c->graphKey) {
            invalidateGraph(c);
            STK_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = stk::launchPathTracePass(sv, lp, c->launch, c->ps, accum, samples, nullptr, c->stream);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(c->stream, &g);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return hipFail(c, e != hipSuccess ? e : e2, "graph capture of a path-tracing pass");
            }
            c->graph = g;
            STK_HIP(c, hipGraphInstantiate(&c->graphExec, c->graph, nullptr, nullptr, 0));
            c->graphKey = key;
        }
    }
    uint32_t k = 0;              // pass number of this call
    hipError_t issued = hipSuccess; // first failure of the multi-stream passes
    if (frameKernel)
        STK_HIP(c, stk::launchPathTraceFrame(sv, lp, c->launch, c->ps, accum, samples, p->pt.frameNumber,
                                             (uint32_t)slots, c->stream));
    for (int s = 0; !frameKernel && s < p->pt.spp && issued == hipSuccess; ++s) {
        for (size_t begin = 0; begin < slots && issued == hipSuccess; begin += Q, ++k) {
            stk::PassState pass{};
            pass.frameNumber = p->pt.frameNumber;
            pass.sampleNum = (uint32_t)s;
            pass.slotBegin = (uint32_t)begin;
            pass.slotCount = (uint32_t)std::min(Q, slots - begin);
            if (pc.mode != 0) {
                // a different subset of paths trains each pass
                uint32_t h = p->pt.frameNumber * (uint32_t)p->pt.spp + (uint32_t)s;
                h ^= h >> 16;
                h *= 0x7feb352du;
                h ^= h >> 15;
                pass.trainOffset = h % pc.trainStride;
                pass.numTrainPaths = (uint32_t)((slots - pass.trainOffset + pc.trainStride - 1) / pc.trainStride);
            }
            if (multi) {
                // the pass streams in turn; each finalize waits for the
                // previous one (of this call or, after a multi-stream call,
                // of that call: the samples buffer is shared), so every pixel
                // sums its samples in order and no sample is overwritten late
                const uint32_t i = (first + k) % nstreams;
                const bool chained = k > 0 || c->lastStreams > 1;
                stk::PtLaunch lpi = lp;
                if (pc.mode != 0) {
                    lpi.cache.queryRecs = static_cast<stk::CacheRecord*>(c->xQueryRecs[i].ptr);
                    lpi.cache.queryOut = static_cast<float4*>(c->xQueryOut[i].ptr);
                    lpi.cache.trainVerts = static_cast<float4*>(c->xTrainVerts[i].ptr);
                    lpi.cache.trainCount = static_cast<uint32_t*>(c->xTrainCount[i].ptr);
                }
                issued = stk::launchSetPass(c->xPs[i], pass, c->xStream[i]);
                if (issued == hipSuccess)
                    issued = stk::launchPathTracePass(
                        sv, lpi, c->launch, c->xPs[i], accum, samples, c->haveCache ? &c->cache : nullptr,
                        c->xStream[i], chained ? c->finEv[c->lastPassStream] : nullptr, c->finEv[i],
                        chained ? c->cacheEv[c->lastPassStream] : nullptr, c->cacheEv[i]);
                c->lastPassStream = i;
                c->lastStreams = nstreams;
                continue;
            }
            STK_HIP(c, stk::launchSetPass(c->ps, pass, c->stream));
            if (graph)
                STK_HIP(c, hipGraphLaunch(c->graphExec, c->stream));
            else
                STK_HIP(c, stk::launchPathTracePass(sv, lp, c->launch, c->ps, accum, samples,
                                                    c->haveCache ? &c->cache : nullptr, c->stream));
        }
    }
    if (multi) {
        // join the pass streams into the context stream; after a failed
        // launch also wait for them, so nothing of this call is left running
        hipError_t joined = hipSuccess;
        for (uint32_t i = 0; i < nstreams; ++i) {
            hipError_t e = hipEventRecord(c->joinEv[i], c->xStream[i]);
            if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, c->joinEv[i], 0);
            if (e != hipSuccess && joined == hipSuccess) joined = e;
        }
        c->passTurn += k;
        ++c->renderCalls;
        if (issued != hipSuccess || joined != hipSuccess) {
            for (uint32_t i = 0; i < nstreams; ++i) (void)hipStreamSynchronize(c->xStream[i]);
            (void)hipStreamSynchronize(c->stream);
            return hipFail(c, issued != hipSuccess ? issued : joined, "render: a path-tracing pass");
        }
    } else {
        c->lastStreams = 1;
        c->lastPassStream = 0;
    }
    const int acc = p->enableAccumulation != 0;
    STK_HIP(c, stk::launchAccumulateTonemap(accum, static_cast<float4*>(c->history.ptr),
                                             outIsDevice && outRgba ? reinterpret_cast<float4*>(outRgba)
                                                                    : static_cast<float4*>(c->out.ptr),
                                             (uint32_t)npix, c->iteration, acc, p->enableTonemap != 0,
                                             p->tonemapperType, static_cast<const uint32_t*>(c->abortAny),
                                             c->stream));
    if (acc) ++c->iteration;
    if (c->accumFree[lane]) {
        STK_HIP(c, hipEventRecord(c->accumFree[lane], c->stream));
        c->accumFreeSet[lane] = true;
    }
    if (outRgba && !outIsDevice) {
        STK_HIP(c, hipMemcpyAsync(outRgba, c->out.ptr, npix * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
        STK_HIP(c, hipStreamSynchronize(c->stream));
    }
    return STK_OK;
}

// ---- watchdog -------------------------------------------------------------------

static void watchdogLoop(stk_ctx* c)
{
    std::unique_lock<std::mutex> lk(c->wdMutex);
    while (!c->wdStop) {
        if (!c->wdArmed) {
            c->wdCv.wait(lk);
            continue;
        }
        const auto deadline = c->wdDeadline;
        c->wdCv.wait_until(lk, deadline);
        if (c->wdArmed && !c->wdStop && c->wdDeadline == deadline && std::chrono::steady_clock::now() >= deadline) {
            __atomic_store_n(c->abortHost, 1u, __ATOMIC_RELEASE);
            c->wdArmed = false;
        }
    }
}

static void armWatchdog(stk_ctx* c, uint32_t ms)
{
    std::lock_guard<std::mutex> lk(c->wdMutex);
    c->wdDeadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(ms);
    c->wdArmed = true;
    c->wdCv.notify_all();
}

static void disarmWatchdog(stk_ctx* c)
{
    std::lock_guard<std::mutex> lk(c->wdMutex);
    c->wdArmed = false;
    c->wdCv.notify_all();
}

// After every stream of the frame has finished: if the abort word was set,
// learns whether a pass was abandoned (abortAny), clears both words and, for
// an abandoned frame, puts back the iteration count and the cache's step
// counters of before the frame and rebuilds the cache's derived copies (its
// optimiser steps returned at entry; their gradients are dropped).
static stk_status settleAbort(stk_ctx* c, const char* what, bool restart = false)
{
    if (__atomic_load_n(c->abortHost, __ATOMIC_ACQUIRE) == 0u) return STK_OK;
    uint32_t any = 0;
    STK_HIP(c, hipStreamSynchronize(c->stream));
    for (hipStream_t s : c->xStream)
        if (s) STK_HIP(c, hipStreamSynchronize(s));
    STK_HIP(c, hipMemcpyAsync(&any, c->abortAny, 4, hipMemcpyDeviceToHost, c->stream));
    STK_HIP(c, hipMemsetAsync(c->abortAny, 0, 4, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    __atomic_store_n(c->abortHost, 0u, __ATOMIC_RELEASE);
    if (!any) return STK_OK;
    // restart: the abandoned passes may include an earlier call's (several
    // pass streams overlap consecutive calls), so the image restarts
    c->iteration = restart ? 0u : c->lastIterBefore;
    if (c->haveCache) {
        // the optimiser steps that ran to completion stay (they used the step
        // numbers they were enqueued with, consecutive from lastStepBefore)
        uint32_t done = c->lastStepBefore;
        STK_HIP(c, hipMemcpyAsync(&done, c->cache.stepsDone, 4, hipMemcpyDeviceToHost, c->stream));
        STK_HIP(c, hipStreamSynchronize(c->stream));
        const uint32_t kept = done >= c->lastStepBefore ? done - c->lastStepBefore : 0u;
        c->cache.step = c->lastStepBefore + kept;
        c->cache.emaSteps = c->cache.emaDecay > 0.0f ? c->lastEmaBefore + kept : c->lastEmaBefore;
        STK_HIP(c, stk::cacheRestore(c->cache, c->stream));
        STK_HIP(c, hipStreamSynchronize(c->stream));
    }
    return fail(c, STK_E_TIMEOUT, what);
}

// renderForestRun under the watchdog: with a deadline armed the call waits
// for its frame, and a frame abandoned by the deadline or by stk_abort
// returns STK_E_TIMEOUT with the accumulation as it was before the call.
static stk_status renderForest(stk_ctx* c, const stk::SceneView& base, std::vector<stk::SceneDesc>& descs,
                               const StkRenderParams* p, float* outRgba, int outIsDevice, bool allowCache,
                               uint64_t imageKey)
{
    stk_status st;
    if (!c->lastSynced && __atomic_load_n(c->abortHost, __ATOMIC_ACQUIRE) != 0u) {
        // an abort that reached the previous call's frame after that call
        // had returned: its accumulation step is undone here
        c->lastSynced = true;
        if ((st = settleAbort(c, "render: the previous call's frame was abandoned after the call returned; its "
                                 "accumulation step was undone")) != STK_OK)
            return st;
    }
    const bool prevSynced = c->lastSynced;
    c->lastIterBefore = c->iteration;
    c->lastStepBefore = c->cache.step;
    c->lastEmaBefore = c->cache.emaSteps;
    const uint32_t ms = c->watchdogMs.load();
    if (ms) armWatchdog(c, ms);
    st = renderForestRun(c, base, descs, p, outRgba, outIsDevice, allowCache, imageKey);
    if (ms) {
        // the deadline covers the frame on the GPU, not only the enqueueing
        hipError_t e = hipStreamSynchronize(c->stream);
        disarmWatchdog(c);
        if (e != hipSuccess && st == STK_OK) st = hipFail(c, e, "render: waiting for the frame");
    }
    c->lastSynced = ms != 0 || (outRgba && !outIsDevice);
    const bool requested = __atomic_load_n(c->abortHost, __ATOMIC_ACQUIRE) != 0u;
    if (requested && (c->lastSynced || st != STK_OK)) {
        c->lastSynced = true;
        const stk_status a = settleAbort(c,
                                         ms ? "render: frame abandoned by the watchdog deadline"
                                            : "render: frame abandoned by stk_abort",
                                         !prevSynced && c->lastStreams > 1);
        if (st == STK_OK) st = a;
    }
    return st;
}

STK_API stk_status stk_set_watchdog(stk_ctx* c, uint32_t milliseconds)
{
    if (!validCtx(c)) return STK_E_INVALID;
    c->watchdogMs.store(milliseconds);
    if (milliseconds && !c->wdThread.joinable()) {
        try {
            c->wdThread = std::thread(watchdogLoop, c);
        } catch (const std::system_error&) {
            c->watchdogMs.store(0);
            return fail(c, STK_E_STATE, "stk_set_watchdog: cannot start the watchdog thread");
        }
    }
    return STK_OK;
}

STK_API stk_status stk_abort(stk_ctx* c)
{
    if (!validCtx(c)) return STK_E_INVALID;
    __atomic_store_n(c->abortHost, 1u, __ATOMIC_RELEASE);
    return STK_OK;
}

static stk::SceneView singleSceneView(stk_ctx* c)
{
    stk::SceneView sv{};
    sv.bvh = bvhView(c);
    sv.vb = static_cast<const StkVertex*>(c->vb.ptr);
    sv.ib = static_cast<const uint32_t*>(c->ib.ptr);
    sv.inst = static_cast<const StkInstance*>(c->inst.ptr);
    sv.lights = static_cast<const StkLight*>(c->lights.ptr);
    sv.mats = static_cast<const StkMaterial*>(c->mats.ptr);
    sv.numMats = c->numMats;
    sv.textures = static_cast<const StkTexture*>(c->textures.ptr);
    sv.numTextures = c->numTextures;
    sv.texels = static_cast<const float4*>(c->texels.ptr);
    return sv;
}

STK_API stk_status stk_render(stk_ctx* c, const StkRenderParams* p, float* outRgba, int outIsDevice)
{
    if (!validCtx(c) || !p) return STK_E_INVALID;
    if (!c->haveBvh) return fail(c, STK_E_STATE, "stk_render: no BVH");
    const int32_t W = p->pt.dimension[0], H = p->pt.dimension[1];
    if (W <= 0 || H <= 0 || p->pt.spp <= 0 || p->pt.maxDepth > 64)
        return fail(c, STK_E_INVALID, "stk_render: bad dimension, spp or maxDepth");
    if ((uint64_t)W * (uint64_t)H > (1ull << 28)) return fail(c, STK_E_INVALID, "stk_render: image too large");
    std::vector<stk::SceneDesc> descs(1);
    stk::SceneDesc& d = descs[0];
    memset(&d, 0, sizeof(d));
    d.hdr = c->bvh.header;
    d.refRoot = c->twoLevel ? STK_INVALID_INDEX : 0u;
    d.wideRoot = (!c->twoLevel && ctxWideUsable(c)) ? 0u : STK_INVALID_INDEX;
    d.emitBase = 0;
    d.emitCount = c->twoLevel ? (uint32_t)c->tl.emitters.size() : c->numEmit;
    d.lightBase = 0;
    d.lightCount = c->numLights;
    d.width = (uint32_t)W;
    d.height = (uint32_t)H;
    for (int i = 0; i < 4; ++i) d.sky[i] = p->skyColor[i];
    d.cam = p->pt;
    const uint64_t key = ((uint64_t)W << 32) ^ (uint64_t)H ^ (1ull << 63);
    return renderForest(c, singleSceneView(c), descs, p, outRgba, outIsDevice, true, key);
}

// ---- batched scenes -----------------------------------------------------------

STK_API stk_status stk_batch_clear(stk_ctx* c)
{
    if (!validCtx(c)) return STK_E_INVALID;
    c->batch.clear();
    c->batchDescs.clear();
    c->batchReady = false;
    return STK_OK;
}

STK_API stk_status stk_batch_add_scene(stk_ctx* c, const StkBatchScene* s, uint32_t* index)
{
    if (!validCtx(c) || !s) return STK_E_INVALID;
    if ((s->numVertices && !s->vb) || (s->numIndices && !s->ib) || (s->numInstances && !s->inst) ||
        (s->numLights && !s->lights) || (s->numMaterials && !s->mats) || s->numIndices % 3 != 0)
        return fail(c, STK_E_INVALID, "stk_batch_add_scene: null buffer or index count not a multiple of 3");
    if (s->camera.dimension[0] <= 0 || s->camera.dimension[1] <= 0 ||
        (uint64_t)s->camera.dimension[0] * (uint64_t)s->camera.dimension[1] > (1ull << 26))
        return fail(c, STK_E_INVALID, "stk_batch_add_scene: bad view size");
    if (c->batch.size() >= 4096) return fail(c, STK_E_INVALID, "stk_batch_add_scene: at most 4096 scenes");
    for (uint32_t i = 0; i < s->numMaterials; ++i) {
        const StkMaterial& m = s->mats[i];
        if (m.albedoTexture != 0 || m.roughnessTexture != 0 || m.normalTexture != 0 || m.emissionTexture != 0)
            return fail(c, STK_E_INVALID, "stk_batch_add_scene: textured materials serve single-scene renders");
        if (m.type > STK_MAT_GLASS || !(m.ior >= 0.0f))
            return fail(c, STK_E_INVALID, "stk_batch_add_scene: unknown material type or negative index of refraction");
    }
    for (uint32_t i = 0; i < s->numInstances; ++i) {
        const StkInstance& ic = s->inst[i];
        if (ic.indexOffset < 0 || ic.indexCount < 0 || ic.indexCount % 3 ||
            (uint64_t)ic.indexOffset + (uint64_t)ic.indexCount > s->numIndices)
            return fail(c, STK_E_INVALID, "stk_batch_add_scene: instance index range outside the index buffer");
        if (ic.lightId >= 0 && (uint32_t)ic.lightId >= s->numLights)
            return fail(c, STK_E_INVALID, "stk_batch_add_scene: instance lightId outside the scene's lights");
    }
    for (uint32_t i = 0; i < s->numIndices; ++i)
        if (s->ib[i] >= s->numVertices) return fail(c, STK_E_INVALID, "stk_batch_add_scene: index out of range");
    for (uint32_t i = 0; i < s->numLights; ++i)
        if (s->lights[i].type != 0) return fail(c, STK_E_INVALID, "stk_batch_add_scene: only rect lights (type 0)");
    try {
        BatchScene b;
        b.vb.assign(s->vb, s->vb + s->numVertices);
        b.ib.assign(s->ib, s->ib + s->numIndices);
        b.inst.assign(s->inst, s->inst + s->numInstances);
        b.lights.assign(s->lights, s->lights + s->numLights);
        b.mats.assign(s->mats, s->mats + s->numMaterials);
        b.cam = s->camera;
        memcpy(b.sky, s->sky, sizeof(b.sky));
        c->batch.push_back(std::move(b));
    } catch (const std::bad_alloc&) {
        return fail(c, STK_E_OOM, "stk_batch_add_scene: host allocation failed");
    }
    c->batchReady = false;
    if (index) *index = (uint32_t)c->batch.size() - 1;
    return STK_OK;
}

// Builds every staged scene's BVH (scenes in parallel, largest first) and
// uploads the merged forest: each scene's vertices, indices, instances,
// lights, materials, nodes and triangles are appended to one array per kind
// with every cross reference rebased by the offsets of the scenes before it.
STK_API stk_status stk_batch_commit(stk_ctx* c, const StkBvhOptions* opts)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (c->batch.empty()) return fail(c, STK_E_STATE, "stk_batch_commit: no scenes added");
    const StkBvhOptions o = opts ? *opts : stk::defaultBvhOptions();
    const size_t n = c->batch.size();
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(),
              [&](size_t a, size_t b) { return c->batch[a].ib.size() > c->batch[b].ib.size(); });
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t k; (k = next.fetch_add(1)) < n;) {
            BatchScene& b = c->batch[order[k]];
            try {
                std::vector<stk::BvhInputTri> tris;
                b.bvh = stk::Bvh{};
                if (!stk::flattenInstances(b.vb.data(), (uint32_t)b.vb.size(), b.ib.data(), (uint32_t)b.ib.size(),
                                           b.inst.data(), (uint32_t)b.inst.size(), tris, b.err) ||
                    !stk::buildBvh(tris, o, b.bvh, b.err) ||
                    !stk::packBvh(b.vb.data(), (uint32_t)b.vb.size(), b.ib.data(), (uint32_t)b.ib.size(),
                                  b.inst.data(), (uint32_t)b.inst.size(), b.bvh, b.err) ||
                    !stk::buildWide(b.inst.data(), (uint32_t)b.inst.size(), b.bvh, b.err))
                    continue;
                if (b.bvh.depth > (uint32_t)stk::dev::kStackDepth) b.err = "tree deeper than the traversal stack";
            } catch (const std::bad_alloc&) {
                b.err = "host allocation failed";
            }
        }
    };
    const size_t nth = std::min<size_t>(n, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nth; ++t) pool.emplace_back(work);
    work();
    for (std::thread& t : pool) t.join();
    for (size_t i = 0; i < n; ++i)
        if (!c->batch[i].err.empty())
            return fail(c, STK_E_INVALID, "stk_batch_commit: scene " + std::to_string(i) + ": " + c->batch[i].err);

    size_t nv = 0, ni = 0, nin = 0, nl = 0, nm = 0, nn = 0, np = 0, nt = 0, nw = 0, ne = 0;
    for (const BatchScene& b : c->batch) {
        nv += b.vb.size();
        ni += b.ib.size();
        nin += b.inst.size();
        nl += b.lights.size();
        nm += b.mats.size();
        nn += b.bvh.nodes.size();
        np += b.bvh.packed.size();
        nt += b.bvh.tris.size();
        nw += b.bvh.wide.size();
        ne += b.bvh.emitters.size();
    }
    if (nv >= (1ull << 31) || ni >= (1ull << 31) || nn >= (1ull << 31) || np >= (1ull << 31) || nt >= (1ull << 31))
        return fail(c, STK_E_INVALID, "stk_batch_commit: merged scene exceeds 2^31 elements");
    std::vector<StkVertex> vb;
    std::vector<uint32_t> ib;
    std::vector<StkInstance> inst;
    std::vector<StkLight> lights;
    std::vector<StkMaterial> mats;
    std::vector<StkBvhNode> nodes;
    std::vector<StkPackedNode> packed;
    std::vector<StkTriangle> tris, wideTris, emit;
    std::vector<StkWideNode> wide;
    std::vector<stk::SceneDesc> descs(n);
    try {
        wide.reserve(nw);
        wideTris.reserve(nt);
        emit.reserve(ne);
        vb.reserve(nv);
        ib.reserve(ni);
        inst.reserve(nin);
        lights.reserve(nl);
        mats.reserve(nm);
        nodes.reserve(nn);
        packed.reserve(np);
        tris.reserve(nt);
        for (size_t i = 0; i < n; ++i) {
            const BatchScene& b = c->batch[i];
            const uint32_t vBase = (uint32_t)vb.size(), iBase = (uint32_t)ib.size(), instBase = (uint32_t)inst.size();
            const uint32_t lBase = (uint32_t)lights.size(), mBase = (uint32_t)mats.size();
            const uint32_t nBase = (uint32_t)nodes.size(), pBase = (uint32_t)packed.size();
            const uint32_t tBase = (uint32_t)tris.size();
            vb.insert(vb.end(), b.vb.begin(), b.vb.end());
            for (uint32_t x : b.ib) ib.push_back(x + vBase);
            for (StkInstance ic : b.inst) {
                ic.indexOffset += (int32_t)iBase;
                ic.materialId = (ic.materialId >= 0 && (size_t)ic.materialId < b.mats.size())
                                    ? ic.materialId + (int32_t)mBase
                                    : -1;
                if (ic.lightId >= 0) ic.lightId += (int32_t)lBase;
                inst.push_back(ic);
            }
            lights.insert(lights.end(), b.lights.begin(), b.lights.end());
            mats.insert(mats.end(), b.mats.begin(), b.mats.end());
            // skip-link nodes: device copy carries the triangle index in maxBounds[0]
            for (size_t k = 0; k < b.bvh.nodes.size(); ++k) {
                StkBvhNode nd = b.bvh.nodes[k];
                if ((uint32_t)nd.instId != STK_INVALID_INDEX) {
                    nd.instId += (int32_t)instBase;
                    nd.maxBounds[0] = stk::bits_to_float(b.bvh.leafToTri[k] + tBase);
                }
                if ((uint32_t)nd.nodeOffset != STK_INVALID_INDEX) nd.nodeOffset += (int32_t)nBase;
                nodes.push_back(nd);
            }
            for (StkPackedNode pn : b.bvh.packed) {
                for (int ch = 0; ch < 2; ++ch)
                    pn.child[ch] = pn.child[ch] >= 0 ? pn.child[ch] + (int32_t)pBase : ~(~pn.child[ch] + (int32_t)tBase);
                packed.push_back(pn);
            }
            auto rebased = [&](StkTriangle t) {
                t.leaf += nBase;
                t.instId += instBase;
                return t;
            };
            for (const StkTriangle& t : b.bvh.tris) tris.push_back(rebased(t));
            const uint32_t wBase = (uint32_t)wide.size(), wtBase = (uint32_t)wideTris.size();
            const uint32_t eBase = (uint32_t)emit.size();
            for (StkWideNode w : b.bvh.wide) {
                w.childBase += wBase;
                w.triBase += wtBase;
                wide.push_back(w);
            }
            for (const StkTriangle& t : b.bvh.wideTris) wideTris.push_back(rebased(t));
            for (const StkTriangle& t : b.bvh.emitters) emit.push_back(rebased(t));
            stk::SceneDesc& d = descs[i];
            memset(&d, 0, sizeof(d));
            d.hdr = b.bvh.header;
            if (d.hdr.root >= 0)
                d.hdr.root += (int32_t)pBase;
            else if (d.hdr.root != INT32_MIN)
                d.hdr.root = ~(~d.hdr.root + (int32_t)tBase);
            d.refRoot = b.bvh.nodes.empty() ? STK_INVALID_INDEX : nBase;
            d.wideRoot = wideUsable(b.bvh) ? wBase : STK_INVALID_INDEX;
            d.emitBase = eBase;
            d.emitCount = (uint32_t)b.bvh.emitters.size();
            d.lightBase = lBase;
            d.lightCount = (uint32_t)b.lights.size();
            d.width = (uint32_t)b.cam.dimension[0];
            d.height = (uint32_t)b.cam.dimension[1];
            memcpy(d.sky, b.sky, sizeof(d.sky));
            d.cam = b.cam;
        }
    } catch (const std::bad_alloc&) {
        return fail(c, STK_E_OOM, "stk_batch_commit: host allocation failed");
    }
    (void)hipSetDevice(c->device);
    stk_status st;
    if ((st = devUpload(c, c->fVb, vb.data(), vb.size() * sizeof(StkVertex))) != STK_OK) return st;
    if ((st = devUpload(c, c->fIb, ib.data(), ib.size() * sizeof(uint32_t))) != STK_OK) return st;
    if ((st = devUpload(c, c->fInst, inst.data(), inst.size() * sizeof(StkInstance))) != STK_OK) return st;
    if ((st = devUpload(c, c->fLights, lights.data(), lights.size() * sizeof(StkLight))) != STK_OK) return st;
    if ((st = devUpload(c, c->fMats, mats.data(), mats.size() * sizeof(StkMaterial))) != STK_OK) return st;
    if ((st = devUpload(c, c->fRefNodes, nodes.data(), nodes.size() * sizeof(StkBvhNode))) != STK_OK) return st;
    if ((st = devUpload(c, c->fPacked, packed.data(), packed.size() * sizeof(StkPackedNode))) != STK_OK) return st;
    if ((st = devUpload(c, c->fTris, tris.data(), tris.size() * sizeof(StkTriangle))) != STK_OK) return st;
    if ((st = devUpload(c, c->fWide, wide.data(), wide.size() * sizeof(StkWideNode))) != STK_OK) return st;
    if ((st = devUpload(c, c->fWideTris, wideTris.data(), wideTris.size() * sizeof(StkTriangle))) != STK_OK)
        return st;
    if ((st = devUpload(c, c->fEmit, emit.data(), emit.size() * sizeof(StkTriangle))) != STK_OK) return st;
    c->fNumMats = (uint32_t)mats.size();
    c->batchDescs = std::move(descs);
    c->batchReady = true;
    // the host node arrays are rebuilt by the next commit; keep only the header
    for (BatchScene& b : c->batch) {
        std::vector<StkBvhNode>().swap(b.bvh.nodes);
        std::vector<StkPackedNode>().swap(b.bvh.packed);
        std::vector<StkTriangle>().swap(b.bvh.tris);
        std::vector<uint32_t>().swap(b.bvh.leafToTri);
        std::vector<StkWideNode>().swap(b.bvh.wide);
        std::vector<StkTriangle>().swap(b.bvh.wideTris);
        std::vector<StkTriangle>().swap(b.bvh.emitters);
    }
    return STK_OK;
}

STK_API stk_status stk_batch_info(stk_ctx* c, uint32_t scene, StkBatchViewInfo* out)
{
    if (!validCtx(c) || !out) return STK_E_INVALID;
    if (!c->batchReady) return fail(c, STK_E_STATE, "stk_batch_info: batch not committed");
    if (scene >= c->batchDescs.size()) return fail(c, STK_E_INVALID, "stk_batch_info: scene out of range");
    const stk::SceneDesc& d = c->batchDescs[scene];
    uint32_t pix = 0;
    for (uint32_t i = 0; i < scene; ++i) pix += c->batchDescs[i].width * c->batchDescs[i].height;
    out->pixelBase = pix;
    out->width = d.width;
    out->height = d.height;
    out->numTriangles = d.hdr.numTriangles;
    out->bvhDepth = c->batch[scene].bvh.depth;
    out->numScenes = (uint32_t)c->batchDescs.size();
    memset(out->pad, 0, sizeof(out->pad));
    return STK_OK;
}

STK_API stk_status stk_batch_render(stk_ctx* c, const StkRenderParams* p, float* outRgba, int outIsDevice)
{
    if (!validCtx(c) || !p) return STK_E_INVALID;
    if (!c->batchReady) return fail(c, STK_E_STATE, "stk_batch_render: batch not committed");
    std::vector<stk::SceneDesc> descs = c->batchDescs;
    for (stk::SceneDesc& d : descs) d.cam.frameNumber = p->pt.frameNumber;
    stk::SceneView sv{};
    sv.bvh.packed = static_cast<const float4*>(c->fPacked.ptr);
    sv.bvh.tris = static_cast<const float4*>(c->fTris.ptr);
    sv.bvh.refNodes = static_cast<const StkBvhNode*>(c->fRefNodes.ptr);
    sv.bvh.instances = static_cast<const StkInstance*>(c->fInst.ptr);
    sv.bvh.wide = static_cast<const float4*>(c->fWide.ptr);
    sv.bvh.wideTris = static_cast<const float4*>(c->fWideTris.ptr);
    sv.bvh.emitters = static_cast<const float4*>(c->fEmit.ptr);
    sv.bvh.wideRoot = STK_INVALID_INDEX; // per scene: SceneDesc.wideRoot
    // the deepest view's wide tree sizes the walks' stacks
    sv.bvh.wideDepth = 0;
    for (const BatchScene& b : c->batch) sv.bvh.wideDepth = std::max(sv.bvh.wideDepth, b.bvh.wideDepth);
    sv.bvh.packedOk = 1;                  // host-built trees fit the packed walk's stack
    sv.vb = static_cast<const StkVertex*>(c->fVb.ptr);
    sv.ib = static_cast<const uint32_t*>(c->fIb.ptr);
    sv.inst = static_cast<const StkInstance*>(c->fInst.ptr);
    sv.lights = static_cast<const StkLight*>(c->fLights.ptr);
    sv.mats = static_cast<const StkMaterial*>(c->fMats.ptr);
    sv.numMats = c->fNumMats;
    sv.textures = nullptr;
    sv.numTextures = 0;
    sv.texels = nullptr;
    const uint64_t key = (uint64_t)(uintptr_t)c->fPacked.ptr ^ ((uint64_t)descs.size() << 48);
    return renderForest(c, sv, descs, p, outRgba, outIsDevice, false, key);
}

STK_API stk_status stk_set_launch_config(stk_ctx* c, const StkLaunchConfig* lc)
{
    if (!validCtx(c) || !lc) return STK_E_INVALID;
    if (lc->blockSize != 64 && lc->blockSize != 128 && lc->blockSize != 256)
        return fail(c, STK_E_INVALID, "stk_set_launch_config: blockSize must be 64, 128 or 256");
    if (lc->tile == 0 || lc->tile > 64) return fail(c, STK_E_INVALID, "stk_set_launch_config: tile must be 1..64");
    if (lc->persistentBlocks > (1u << 20))
        return fail(c, STK_E_INVALID, "stk_set_launch_config: persistentBlocks too large");
    if (lc->passStreams > kMaxPassStreams)
        return fail(c, STK_E_INVALID, "stk_set_launch_config: passStreams must be 0..4");
    if (lc->pathsInFlight != 0 && lc->pathsInFlight < 4096)
        return fail(c, STK_E_INVALID, "stk_set_launch_config: pathsInFlight must be 0 or at least 4096");
    if (lc->frameBlocks > (1u << 16))
        return fail(c, STK_E_INVALID, "stk_set_launch_config: frameBlocks must be at most 65536");
    if (lc->frameChunk % 256 != 0 || lc->frameChunk > 8192)
        return fail(c, STK_E_INVALID, "stk_set_launch_config: frameChunk must be a multiple of 256, at most 8192");
    if ((uint64_t)lc->frameBlocks * std::max<uint32_t>(256u, lc->frameChunk) > (1ull << 26))
        return fail(c, STK_E_INVALID, "stk_set_launch_config: frameBlocks x frameChunk too large");
    c->launch.blockSize = lc->blockSize;
    c->launch.tile = lc->tile;
    c->launch.persistentBlocks = lc->persistentBlocks;
    c->launch.pathsInFlight = lc->pathsInFlight;
    c->launch.binaryWalk = lc->binaryWalk ? 1u : 0u;
    c->launch.stripeThreads = lc->stripeThreads;
    c->launch.passStreams = std::max<uint32_t>(1u, lc->passStreams);
    c->launch.frameBlocks = lc->frameBlocks;
    c->launch.frameChunk = lc->frameChunk;
    c->useGraph = lc->useGraph != 0;
    invalidateGraph(c);
    return STK_OK;
}

STK_API stk_status stk_device_info(stk_ctx* c, StkDeviceInfo* out)
{
    if (!validCtx(c) || !out) return STK_E_INVALID;
    hipDeviceProp_t pr;
    STK_HIP(c, hipGetDeviceProperties(&pr, c->device));
    memset(out, 0, sizeof(*out));
    snprintf(out->name, sizeof(out->name), "%s", pr.name);
    snprintf(out->arch, sizeof(out->arch), "%s", pr.gcnArchName);
    out->computeUnits = (uint32_t)pr.multiProcessorCount;
    out->wavefrontSize = (uint32_t)pr.warpSize;
    out->globalMemBytes = (uint64_t)pr.totalGlobalMem;
    out->l2Bytes = (uint32_t)pr.l2CacheSize;
    out->clockKhz = (uint32_t)pr.clockRate;
    out->memClockKhz = (uint32_t)pr.memoryClockRate;
    out->memBusWidth = (uint32_t)pr.memoryBusWidth;
    return STK_OK;
}

STK_API stk_status stk_get_launch_config(stk_ctx* c, StkLaunchConfig* lc)
{
    if (!validCtx(c) || !lc) return STK_E_INVALID;
    lc->blockSize = c->launch.blockSize;
    lc->tile = c->launch.tile;
    lc->persistentBlocks = c->launch.persistentBlocks;
    lc->useGraph = c->useGraph ? 1u : 0u;
    lc->pathsInFlight = c->launch.pathsInFlight;
    lc->binaryWalk = c->launch.binaryWalk;
    lc->stripeThreads = c->launch.stripeThreads;
    lc->passStreams = c->launch.passStreams;
    lc->frameBlocks = c->launch.frameBlocks;
    lc->frameChunk = c->launch.frameChunk;
    return STK_OK;
}

STK_API stk_status stk_render_image(stk_ctx* c, const StkRenderParams* p, const StkImageDesc* img)
{
    if (!validCtx(c) || !p || !img || !img->data) return STK_E_INVALID;
    if (img->format != STK_IMAGE_RGBA32F && img->format != STK_IMAGE_RGBA8)
        return fail(c, STK_E_INVALID, "stk_render_image: unknown format");
    if ((int64_t)img->width != p->pt.dimension[0] || (int64_t)img->height != p->pt.dimension[1])
        return fail(c, STK_E_INVALID, "stk_render_image: image size differs from the camera's dimension");
    const size_t px = img->format == STK_IMAGE_RGBA8 ? 4 : 16;
    const size_t row = (size_t)img->width * px;
    const size_t pitch = img->pitchBytes ? img->pitchBytes : row;
    if (pitch < row) return fail(c, STK_E_INVALID, "stk_render_image: pitch smaller than a row");
    stk_status st = stk_render(c, p, nullptr, 0);
    if (st != STK_OK) return st;
    const size_t npix = (size_t)img->width * img->height;
    if ((st = devReserve(c, c->staging, npix * px)) != STK_OK) return st;
    STK_HIP(c, stk::launchPackImage(static_cast<const float4*>(c->out.ptr), (uint32_t)npix, img->format,
                                    c->staging.ptr, c->stream));
    STK_HIP(c, hipMemcpy2DAsync(img->data, pitch, c->staging.ptr, row, row, img->height,
                                img->onDevice ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
    if (!img->onDevice) STK_HIP(c, hipStreamSynchronize(c->stream));
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
    if (count > c->imagePixels * 4 || !c->history.ptr)
        return fail(c, STK_E_STATE, "stk_get_accumulation: nothing rendered or count too large");
    STK_HIP(c, hipMemcpyAsync(out, c->history.ptr, count * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    return STK_OK;
}

// ---- checkpoint / resume ------------------------------------------------------
//
// The progressive state the reference keeps in GPU images (the accumulation
// image and mPtIteration, src/render/ptrender.cpp:714-726,839) and the
// radiance cache's training state, written to one file and read back into a
// context so a long convergence run or a training run continues where it
// stopped.  Layout (little-endian, no padding between records):
//   CkptHeader; [CkptAccum, float4 history[pixels]] if section bit 0;
//   [CkptCache, grid P, M, V, mlp P, M, V, and when emaDecay > 0 grid E, EW,
//   mlp E] if bit 1; uint64 FNV-1a hash of every byte before it.

namespace {

constexpr char kCkptMagic[8] = {'S', 'T', 'K', 'C', 'K', 'P', 'T', '1'};
constexpr uint32_t kCkptVersion = 1;

struct CkptHeader {
    char magic[8];
    uint32_t version;
    uint32_t sections; // bit 0 accumulation, bit 1 cache
    uint64_t userFrame;
};
struct CkptAccum {
    uint64_t imageKey;
    uint64_t pixels;
    uint32_t iteration;
    uint32_t pad;
};
struct CkptCache {
    StkCacheConfig cfg;
    stk::GridDesc grid;
    uint32_t step, emaSteps;
    uint64_t gridParams, mlpParams;
};

uint64_t fnv1a(uint64_t h, const void* p, size_t n)
{
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 0x100000001B3ull;
    }
    return h;
}

struct CkptWriter {
    FILE* f = nullptr;
    uint64_t hash = 0xCBF29CE484222325ull;
    bool ok = true;
    void put(const void* p, size_t n)
    {
        if (!ok || n == 0) return;
        hash = fnv1a(hash, p, n);
        ok = std::fwrite(p, 1, n, f) == n;
    }
};

struct CkptReader {
    const std::vector<unsigned char>& buf;
    size_t pos = 0;
    bool take(void* dst, size_t n)
    {
        if (buf.size() - pos < n) return false;
        std::memcpy(dst, buf.data() + pos, n);
        pos += n;
        return true;
    }
    const unsigned char* view(size_t n)
    {
        if (buf.size() - pos < n) return nullptr;
        const unsigned char* p = buf.data() + pos;
        pos += n;
        return p;
    }
};

// Waits for the context's stream and every pass stream.
stk_status syncAll(stk_ctx* c)
{
    STK_HIP(c, hipStreamSynchronize(c->stream));
    for (hipStream_t s : c->xStream)
        if (s) STK_HIP(c, hipStreamSynchronize(s));
    return STK_OK;
}

stk_status download(stk_ctx* c, std::vector<float>& dst, const void* src, size_t floats)
{
    dst.resize(floats);
    if (floats) STK_HIP(c, hipMemcpyAsync(dst.data(), src, floats * 4, hipMemcpyDeviceToHost, c->stream));
    return STK_OK;
}

} // namespace

STK_API stk_status stk_checkpoint_save(stk_ctx* c, const char* path, uint64_t userFrame)
{
    if (!validCtx(c) || !path) return STK_E_INVALID;
    (void)hipSetDevice(c->device);
    stk_status st;
    if ((st = syncAll(c)) != STK_OK) return st;
    const bool accum = c->history.ptr && c->imagePixels > 0;
    const bool cache = c->haveCache;
    std::vector<float> history, gP, gM, gV, mP, mM, mV, gE, gEW, mE;
    try {
        if (accum && (st = download(c, history, c->history.ptr, c->imagePixels * 4)) != STK_OK) return st;
        if (cache) {
            const stk::CacheBuffers& cb = c->cache;
            const size_t gp = cb.gridParams, mp = stk::kCacheMlpParams;
            if ((st = download(c, gP, cb.gridP, gp)) != STK_OK || (st = download(c, gM, cb.gridM, gp)) != STK_OK ||
                (st = download(c, gV, cb.gridV, gp)) != STK_OK || (st = download(c, mP, cb.mlpP, mp)) != STK_OK ||
                (st = download(c, mM, cb.mlpM, mp)) != STK_OK || (st = download(c, mV, cb.mlpV, mp)) != STK_OK)
                return st;
            if (cb.emaDecay > 0.0f &&
                ((st = download(c, gE, cb.gridE, gp)) != STK_OK || (st = download(c, gEW, cb.gridEW, gp)) != STK_OK ||
                 (st = download(c, mE, cb.mlpE, mp)) != STK_OK))
                return st;
        }
    } catch (const std::bad_alloc&) {
        return fail(c, STK_E_OOM, "stk_checkpoint_save: host allocation failed");
    }
    STK_HIP(c, hipStreamSynchronize(c->stream));

    // written next to the target and renamed over it, so an interrupted save
    // leaves the previous checkpoint intact
    const std::string tmp = std::string(path) + ".part";
    CkptWriter w;
    w.f = std::fopen(tmp.c_str(), "wb");
    if (!w.f) return fail(c, STK_E_INVALID, std::string("stk_checkpoint_save: cannot open ") + tmp);
    CkptHeader h;
    std::memcpy(h.magic, kCkptMagic, 8);
    h.version = kCkptVersion;
    h.sections = (accum ? 1u : 0u) | (cache ? 2u : 0u);
    h.userFrame = userFrame;
    w.put(&h, sizeof(h));
    if (accum) {
        CkptAccum a{c->imageKey, (uint64_t)c->imagePixels, c->iteration, 0};
        w.put(&a, sizeof(a));
        w.put(history.data(), history.size() * 4);
    }
    if (cache) {
        CkptCache k;
        std::memset(&k, 0, sizeof(k));
        k.cfg = c->cacheCfg;
        k.grid = c->cache.grid;
        k.step = c->cache.step;
        k.emaSteps = c->cache.emaSteps;
        k.gridParams = c->cache.gridParams;
        k.mlpParams = stk::kCacheMlpParams;
        w.put(&k, sizeof(k));
        for (const std::vector<float>* v : {&gP, &gM, &gV, &mP, &mM, &mV, &gE, &gEW, &mE})
            w.put(v->data(), v->size() * 4);
    }
    const uint64_t hash = w.hash;
    w.put(&hash, sizeof(hash));
    const bool closed = std::fclose(w.f) == 0;
    if (!w.ok || !closed || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return fail(c, STK_E_INVALID, std::string("stk_checkpoint_save: write failed: ") + path);
    }
    return STK_OK;
}

STK_API stk_status stk_checkpoint_load(stk_ctx* c, const char* path, StkCheckpointInfo* info)
{
    if (!validCtx(c) || !path) return STK_E_INVALID;
    std::vector<unsigned char> buf;
    {
        FILE* f = std::fopen(path, "rb");
        if (!f) return fail(c, STK_E_INVALID, std::string("stk_checkpoint_load: cannot open ") + path);
        std::fseek(f, 0, SEEK_END);
        const long size = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        try {
            buf.resize(size > 0 ? (size_t)size : 0);
        } catch (const std::bad_alloc&) {
            std::fclose(f);
            return fail(c, STK_E_OOM, "stk_checkpoint_load: host allocation failed");
        }
        const bool ok = size > 0 && std::fread(buf.data(), 1, buf.size(), f) == buf.size();
        std::fclose(f);
        if (!ok) return fail(c, STK_E_INVALID, "stk_checkpoint_load: read failed");
    }
    if (buf.size() < sizeof(CkptHeader) + 8) return fail(c, STK_E_INVALID, "stk_checkpoint_load: truncated file");
    uint64_t stored;
    std::memcpy(&stored, buf.data() + buf.size() - 8, 8);
    if (fnv1a(0xCBF29CE484222325ull, buf.data(), buf.size() - 8) != stored)
        return fail(c, STK_E_INVALID, "stk_checkpoint_load: checksum mismatch (truncated or damaged file)");
    CkptReader r{buf};
    CkptHeader h;
    r.take(&h, sizeof(h));
    if (std::memcmp(h.magic, kCkptMagic, 8) != 0 || h.version != kCkptVersion || (h.sections & ~3u))
        return fail(c, STK_E_INVALID, "stk_checkpoint_load: not a checkpoint of this version");

    // everything is validated before any device state changes
    CkptAccum a{};
    const unsigned char* history = nullptr;
    if (h.sections & 1u) {
        if (!r.take(&a, sizeof(a)) || a.pixels == 0 || a.pixels > (1ull << 28) ||
            !(history = r.view((size_t)a.pixels * 16)))
            return fail(c, STK_E_INVALID, "stk_checkpoint_load: bad accumulation section");
    }
    CkptCache k{};
    const unsigned char* cacheData = nullptr;
    size_t gp = 0, mp = 0;
    bool ema = false;
    if (h.sections & 2u) {
        if (!r.take(&k, sizeof(k))) return fail(c, STK_E_INVALID, "stk_checkpoint_load: bad cache section");
        if (!c->haveCache)
            return fail(c, STK_E_STATE, "stk_checkpoint_load: the checkpoint holds a cache; create one with its "
                                        "configuration first (stk_cache_create)");
        if (std::memcmp(&k.cfg, &c->cacheCfg, sizeof(k.cfg)) != 0 ||
            std::memcmp(&k.grid, &c->cache.grid, sizeof(k.grid)) != 0 || k.gridParams != c->cache.gridParams ||
            k.mlpParams != stk::kCacheMlpParams)
            return fail(c, STK_E_STATE, "stk_checkpoint_load: the context's cache has another configuration or "
                                        "scene bounds than the checkpoint's");
        gp = (size_t)k.gridParams;
        mp = (size_t)k.mlpParams;
        ema = k.cfg.emaDecay > 0.0f;
        if (!(cacheData = r.view((3 * gp + 3 * mp + (ema ? 2 * gp + mp : 0)) * 4)))
            return fail(c, STK_E_INVALID, "stk_checkpoint_load: bad cache section");
    }
    if (r.pos != buf.size() - 8) return fail(c, STK_E_INVALID, "stk_checkpoint_load: trailing bytes");

    (void)hipSetDevice(c->device);
    stk_status st;
    if ((st = syncAll(c)) != STK_OK) return st;
    if (h.sections & 1u) {
        const void* before = c->history.ptr;
        if ((st = devReserve(c, c->history, (size_t)a.pixels * 16)) != STK_OK) return st;
        if (c->history.ptr != before) invalidateGraph(c);
        STK_HIP(c, hipMemcpyAsync(c->history.ptr, history, (size_t)a.pixels * 16, hipMemcpyHostToDevice, c->stream));
        c->imageKey = a.imageKey;
        c->imagePixels = (size_t)a.pixels;
        c->iteration = a.iteration;
    }
    if (h.sections & 2u) {
        stk::CacheBuffers& cb = c->cache;
        const unsigned char* p = cacheData;
        auto up = [&](void* dst, size_t floats) -> hipError_t {
            hipError_t e = hipMemcpyAsync(dst, p, floats * 4, hipMemcpyHostToDevice, c->stream);
            p += floats * 4;
            return e;
        };
        STK_HIP(c, up(cb.gridP, gp));
        STK_HIP(c, up(cb.gridM, gp));
        STK_HIP(c, up(cb.gridV, gp));
        STK_HIP(c, up(cb.mlpP, mp));
        STK_HIP(c, up(cb.mlpM, mp));
        STK_HIP(c, up(cb.mlpV, mp));
        if (ema) {
            STK_HIP(c, up(cb.gridE, gp));
            STK_HIP(c, up(cb.gridEW, gp));
            STK_HIP(c, up(cb.mlpE, mp));
        }
        cb.step = k.step;
        cb.emaSteps = k.emaSteps;
        STK_HIP(c, hipMemcpyAsync(cb.stepsDone, &k.step, 4, hipMemcpyHostToDevice, c->stream));
        STK_HIP(c, stk::cacheRestore(cb, c->stream));
    }
    STK_HIP(c, hipStreamSynchronize(c->stream));
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->sections = h.sections;
        info->iteration = a.iteration;
        info->pixels = a.pixels;
        info->userFrame = h.userFrame;
        info->cacheStep = k.step;
    }
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
                      &c->cGridE, &c->cGridEW, &c->cGridHE, &c->cMlpE, &c->cFwdE, &c->cBwdE, &c->cGridGD,
                      &c->cSortKeys, &c->cSortVals, &c->cSortTemp};
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
        cfg.maxBatch > (1u << 24) || !(cfg.learningRate > 0.0f) || !(cfg.emaDecay >= 0.0f && cfg.emaDecay < 1.0f) ||
        !(cfg.lrHalfLife >= 0.0f && cfg.lrHalfLife < 1e9f) ||
        !(cfg.minLearningRate >= 0.0f && cfg.minLearningRate <= cfg.learningRate))
        return fail(c, STK_E_INVALID, "stk_cache_create: configuration out of range");
    stk_cache_destroy(c);
    stk::CacheBuffers& cb = c->cache;
    stk::GridDesc& gd = cb.grid;
    // the scene box: the flat tree's root box, or the box a two-level
    // structure keeps in the header (its root field is then unset)
    const StkBvhHeader& h = c->bvh.header;
    const bool box = h.root >= 0 || c->twoLevel;
    float ext = 0.0f;
    for (int a = 0; a < 3; ++a) {
        gd.lo[a] = box ? h.lo[a] : 0.0f;
        if (box) ext = std::max(ext, h.hi[a] - h.lo[a]);
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
    gd.denseEntries = 0;
    for (int l = 0; l < stk::kCacheLevels; ++l) {
        const uint64_t s = (uint64_t)gd.res[l] + 1;
        gd.denseBase[l] = gd.dense[l] ? gd.denseEntries : 0xFFFFFFFFu;
        if (gd.dense[l]) gd.denseEntries += (uint32_t)(s * s * s);
    }
    cb.gridParams = (size_t)stk::kCacheLevels * T * stk::kCacheFeatures;
    cb.nCap = (cfg.maxBatch + 1023u) & ~1023u;
    cb.lr = cfg.learningRate;
    cb.lrHalfLife = cfg.lrHalfLife;
    cb.lrMin = cfg.minLearningRate;
    cb.beta1 = cfg.beta1;
    cb.beta2 = cfg.beta2;
    cb.eps = cfg.epsilon;
    cb.step = 0;
    const size_t gp = cb.gridParams, mp = stk::kCacheMlpParams;
    size_t sortBytes = 0;
    if (stk::cacheSortBytes(cb.nCap, sortBytes) != hipSuccess)
        return fail(c, STK_E_DEVICE, "stk_cache_create: sort storage query failed");
    struct {
        DevBuf* b;
        size_t bytes;
    } allocs[] = {{&c->cGridH, gp * 2},  {&c->cGridP, gp * 4}, {&c->cGridM, gp * 4},   {&c->cGridV, gp * 4},
                  {&c->cGridG, gp * 4},  {&c->cMlpP, mp * 4},  {&c->cMlpM, mp * 4},    {&c->cMlpV, mp * 4},
                  {&c->cMlpG, mp * 4},   {&c->cFwd, 68 * 256 * 2}, {&c->cBwd, 68 * 256 * 2},
                  {&c->cXT, (size_t)5 * 64 * cb.nCap * 2}, {&c->cDyT, (size_t)272 * cb.nCap * 2},
                  {&c->cLoss, 64},
                  {&c->cGridGD, (size_t)stk::kDenseCopies * gd.denseEntries * 2 * 4 + 16},
                  {&c->cSortKeys, (size_t)2 * cb.nCap * 4}, {&c->cSortVals, (size_t)2 * cb.nCap * 4},
                  {&c->cSortTemp, sortBytes + 16}};
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
        } ema[] = {{&c->cGridE, gp * 4}, {&c->cGridEW, gp * 4}, {&c->cGridHE, gp * 2}, {&c->cMlpE, mp * 4},
                   {&c->cFwdE, 68 * 256 * 2}, {&c->cBwdE, 68 * 256 * 2}};
        for (auto& a : ema) {
            stk_status st = devReserve(c, *a.b, a.bytes);
            if (st != STK_OK) {
                stk_cache_destroy(c);
                return st;
            }
        }
        cb.gridE = static_cast<float*>(c->cGridE.ptr);
        cb.gridEW = static_cast<float*>(c->cGridEW.ptr);
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
    cb.gridGDense = static_cast<float*>(c->cGridGD.ptr);
    cb.sortKeys = static_cast<uint32_t*>(c->cSortKeys.ptr);
    cb.sortVals = static_cast<uint32_t*>(c->cSortVals.ptr);
    cb.sortTemp = c->cSortTemp.ptr;
    cb.sortBytes = sortBytes;
    cb.mlpP = static_cast<float*>(c->cMlpP.ptr);
    cb.mlpM = static_cast<float*>(c->cMlpM.ptr);
    cb.mlpV = static_cast<float*>(c->cMlpV.ptr);
    cb.mlpG = static_cast<float*>(c->cMlpG.ptr);
    cb.wFwd = static_cast<__half*>(c->cFwd.ptr);
    cb.wBwd = static_cast<__half*>(c->cBwd.ptr);
    cb.xT = static_cast<__half*>(c->cXT.ptr);
    cb.dyT = static_cast<__half*>(c->cDyT.ptr);
    cb.loss = static_cast<float*>(c->cLoss.ptr);
    cb.stepsDone = static_cast<uint32_t*>(c->cLoss.ptr) + 8;
    STK_HIP(c, hipMemsetAsync(c->cLoss.ptr, 0, 64, c->stream));
    // initialisation: grid U(-1e-4, 1e-4); hidden weights He-uniform
    // U(-sqrt(6/64), sqrt(6/64)); the output layer zero, so the cache starts
    // by predicting no radiance instead of an arbitrary value it must unlearn
    // (measured: tonemapped RMSE after 4 cached frames 0.054 against 0.083)
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
        for (size_t o = 0; o < 16; ++o) mlp[16384 + in * 16 + o] = 0.0f; // output layer (3 of 16 columns used)
    STK_HIP(c, hipMemcpyAsync(cb.gridP, grid.data(), gp * 4, hipMemcpyHostToDevice, c->stream));
    STK_HIP(c, hipMemcpyAsync(cb.mlpP, mlp.data(), mp * 4, hipMemcpyHostToDevice, c->stream));
    STK_HIP(c, hipMemsetAsync(cb.gridM, 0, gp * 4, c->stream));
    STK_HIP(c, hipMemsetAsync(cb.gridV, 0, gp * 4, c->stream));
    STK_HIP(c, hipMemsetAsync(cb.gridG, 0, gp * 4, c->stream));
    STK_HIP(c, hipMemsetAsync(cb.gridGDense, 0, (size_t)stk::kDenseCopies * gd.denseEntries * 2 * 4, c->stream));
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
    STK_HIP(c, stk::cacheAdam(c->cache, 1.0f / (float)n, nullptr, n, nullptr, c->stream));
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
    if (!c->passState(c->lastPassStream).counters) return fail(c, STK_E_STATE, "stk_get_counters: nothing rendered");
    STK_HIP(c, hipMemcpyAsync(out, c->passState(c->lastPassStream).counters, 8 * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    return STK_OK;
}

STK_API stk_status stk_cache_last_records(stk_ctx* c, uint32_t which, StkCacheRecord* out, uint32_t capacity,
                                          uint32_t* n)
{
    if (!validCtx(c) || !n || which > 1) return STK_E_INVALID;
    if (!c->haveCache) return fail(c, STK_E_STATE, "stk_cache_last_records: no cache");
    const stk::PathState& ps = c->passState(c->lastPassStream);
    const DevBuf& buf = which == 0 ? (c->lastStreams > 1 ? c->xQueryRecs[c->lastPassStream] : c->queryRecs)
                                   : c->trainBatch;
    if (!ps.counters || !buf.ptr) return fail(c, STK_E_STATE, "stk_cache_last_records: no cached render yet");
    (void)hipSetDevice(c->device);
    for (hipStream_t s : c->xStream)
        if (s) STK_HIP(c, hipStreamSynchronize(s));
    uint32_t cnt[8];
    STK_HIP(c, hipMemcpyAsync(cnt, ps.counters, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    const size_t held = buf.bytes / sizeof(stk::CacheRecord);
    uint32_t count = which == 0 ? cnt[3] : std::min(cnt[4], c->lastTrainCap);
    count = (uint32_t)std::min<size_t>(count, held);
    *n = count;
    if (!out || count == 0) return STK_OK;
    if (capacity < count) return fail(c, STK_E_INVALID, "stk_cache_last_records: capacity too small");
    STK_HIP(c, hipMemcpyAsync(out, buf.ptr, (size_t)count * sizeof(stk::CacheRecord), hipMemcpyDeviceToHost,
                              c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    return STK_OK;
}

STK_API stk_status stk_get_ray_totals(stk_ctx* c, uint64_t out[2])
{
    if (!validCtx(c) || !out) return STK_E_INVALID;
    if (!c->passState(0).counters) return fail(c, STK_E_STATE, "stk_get_ray_totals: nothing rendered");
    STK_HIP(c, hipMemcpyAsync(out, c->passState(0).counters + 8, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                              c->stream));
    uint64_t more[kMaxPassStreams - 1][2] = {};
    for (uint32_t i = 1; i < c->lastStreams; ++i)
        STK_HIP(c, hipMemcpyAsync(more[i - 1], c->passState(i).counters + 8, sizeof(more[0]), hipMemcpyDeviceToHost,
                                  c->stream));
    STK_HIP(c, hipStreamSynchronize(c->stream));
    for (uint32_t i = 1; i < c->lastStreams; ++i) {
        out[0] += more[i - 1][0];
        out[1] += more[i - 1][1];
    }
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

STK_API stk_status stk_host_alloc(stk_ctx* c, size_t bytes, void** out)
{
    if (!validCtx(c) || !out) return STK_E_INVALID;
    *out = nullptr;
    if (bytes == 0) return STK_OK;
    (void)hipSetDevice(c->device);
    STK_HIP(c, hipHostMalloc(out, bytes, hipHostMallocMapped | hipHostMallocPortable));
    return STK_OK;
}

STK_API stk_status stk_host_free(stk_ctx* c, void* p)
{
    if (!validCtx(c)) return STK_E_INVALID;
    if (p) STK_HIP(c, hipHostFree(p));
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
