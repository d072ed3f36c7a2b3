// pathtrace.hip - wavefront path tracer for gfx950 over one or many views.
//
// Replaces the per-(pixel, sample) megakernel computeMain/pathTraceCameraRays
// (shaders/pathtracerMdl.hlsl:204-423) with wavefront stages:
//   reset -> raygen -> per bounce { setup, extend (closest hit), shade (NEE
//   sample, BSDF sample, RR; compacts continuing paths and shadow rays into
//   queues), connect (any hit, adds NEE) } -> [cache answers] -> finalize ->
//   [cache training].
// Queues are striped: a queue is kStripes segments of segCap entries, block b
// of a producer appends to segment b % kStripes (one atomic per wave on that
// segment's counter), and consumers run stripe-major grids.  Appends to one
// counter serialise in L2 (about 11 ns each on MI355X, 0.37 ms for the 32,400
// waves of a 1080p pass); 64 counters cut that to a few microseconds.
// A path's sequence of random numbers and of additions to its radiance is the
// reference's, so its value does not depend on the wavefront organisation or
// on how many views share the launch.
//
// Many views: the scenes of all views live in one forest (concatenated node,
// triangle, instance, light and material arrays); every path slot knows its
// scene.  Slots are ordered by view, then by square tile, then row-major in
// the tile, so a wave's rays are spatially coherent; views are padded to whole
// tiles and the padding slots never enter a queue.
#include "kernels.hpp"
#include "pt_dev.hpp"

namespace stk {
namespace {

using namespace dev;

constexpr int kBlock = 256;
constexpr int kWaveBlock = 64; // one-wave blocks of the shadow-ray walks
constexpr float kMinAlbedo = 1e-3f; // floor of the albedo the cache's targets are divided by
constexpr int kTrainWords = 6;      // float4s recorded per training vertex

__device__ __forceinline__ float4 f4(F3 v) { return make_float4(v.x, v.y, v.z, 0.0f); }
__device__ __forceinline__ F3 f3(float4 v) { return mk3(v.x, v.y, v.z); }

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

__device__ __forceinline__ uint32_t scene_of_slot(const SceneView& sc, uint32_t slot)
{
    uint32_t lo = 0, hi = sc.numScenes; // last scene whose slotBase <= slot
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (sc.scenes[mid].slotBase <= slot)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ bool is_train(const PtCache& pc, const PassState& pass, uint32_t slot)
{
    return (pc.mode == 1 || pc.mode == 3 || pc.mode == 4) && slot % pc.trainStride == pass.trainOffset;
}

__global__ void k_pass_set(PassState* pass, PassState v)
{
    if (threadIdx.x == 0) *pass = v;
}

// Clears the pass counters and every stripe counter (one 64-thread block).
// The watchdog's host word into the pass's abort word (and the context's
// abortAny when set); one lane, from a one-wave bookkeeping kernel.
__device__ __forceinline__ void poll_abort(const PathState& ps)
{
    if (!ps.abortHost) return;
    if (__hip_atomic_load(ps.abortHost, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u) {
        ps.counters[kAbortCounter] = 1u;
        *ps.abortAny = 1u;
    }
}

// Whether this pass was abandoned.  Only this pass stream's one-wave kernels
// and its persistent extend write the word, never while another kernel of the
// stream runs, so every thread of a launch reads the same value.
__device__ __forceinline__ bool pass_aborted(const PathState& ps) { return ps.counters[kAbortCounter] != 0u; }

__global__ void k_pass_reset(PathState ps)
{
    const uint32_t t = threadIdx.x;
    if (t < 8) ps.counters[t] = 0;
    if (t < kStripes) {
        for (int k = 0; k < kQueueKinds; ++k) ps.stripes[k * kStripes + t] = 0;
        ps.work[t] = 0;
    }
    if (t == 0) poll_abort(ps);
}

// Stripe-major consumer grids of B-thread blocks: block b reads segment
// b % kStripes, entries (b / kStripes) * B + lane, then every
// (gridDim / kStripes) * B further while the segment has entries (a grid
// capped below one thread per possible entry loops).  The loop bound is
// uniform over the block, so wave-wide appends inside see whole waves.
__device__ __forceinline__ uint32_t stripe_first(uint32_t B) { return (blockIdx.x / kStripes) * B; }
__device__ __forceinline__ uint32_t stripe_step(uint32_t B) { return (gridDim.x / kStripes) * B; }

// Sum over a 64-thread block (one wave).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o);
    return v;
}

// One slot's camera ray: the pass's sample of the slot's pixel, or an empty
// path for a padding slot.  Per-slot arrays are indexed by `local`; `slot` is
// the global slot (which view and pixel).  Returns whether the path starts.
__device__ __forceinline__ bool raygen_one(const SceneView& sc, const PtLaunch& lp, PathState& ps,
                                           const PassState& pass, uint32_t local, uint32_t slot)
{
    const uint32_t sid = scene_of_slot(sc, slot);
    const SceneDesc& d = sc.scenes[sid];
    const uint32_t T = sc.tile, inView = slot - d.slotBase;
    const uint32_t tileIdx = inView / (T * T), within = inView % (T * T);
    const uint32_t px = (tileIdx % d.tilesX) * T + within % T;
    const uint32_t py = (tileIdx / d.tilesX) * T + within / T;
    bool valid = px < d.width && py < d.height;
    // mode 4 traces the training paths only (cache warm-up)
    if (lp.cache.mode == 4 && !is_train(lp.cache, pass, local)) valid = false;
    ps.scene[local] = sid;
    ps.pixel[local] = valid ? d.pixelBase + py * d.width + px : STK_INVALID_INDEX;
    if (valid) {
        const uint32_t frame = pass.frameNumber * lp.spp + pass.sampleNum;
        uint32_t rng = init_rng(px, py, d.width, frame);
        float4 o, dir;
        camera_ray(d.cam, px, py, rng, frame, o, dir);
        ps.rayO[local] = o;
        ps.rayD[local] = dir;
        ps.thr[local] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        ps.col[local] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        ps.rng[local] = rng;
    } else {
        ps.col[local] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    // padding slots too: the target kernel reads every training slot's count
    // (the cache runs with one pass over all slots, so local == slot there)
    if (is_train(lp.cache, pass, local)) lp.cache.trainCount[local / lp.cache.trainStride] = 0;
    return valid;
}

// Path state arrays are indexed by the slot's position in the pass (`local`);
// a pass covers global slots [slotBegin, slotBegin + slotCount).
__global__ __launch_bounds__(kBlock) void k_pt_raygen(SceneView sc, PtLaunch lp, PathState ps)
{
    if (pass_aborted(ps)) return;
    const uint32_t local = blockIdx.x * kBlock + threadIdx.x;
    const PassState pass = *ps.pass;
    bool valid = false;
    if (local < pass.slotCount) valid = raygen_one(sc, lp, ps, pass, local, pass.slotBegin + local);
    const uint32_t stripe = blockIdx.x % kStripes;
    const uint32_t qi = wave_push(&ps.stripes[stripe], valid);
    if (valid) ps.queue[0][stripe * ps.segCap + qi] = local;
}

// Ray totals of a render call: two 64-bit words after the eight pass counters
// (closest-hit rays, shadow rays), cleared by the host before the first pass.
__device__ __forceinline__ unsigned long long* ray_totals(const PathState& ps)
{
    return reinterpret_cast<unsigned long long*>(ps.counters + 8);
}

// Start of a bounce (one 64-thread block): tallies the rays of this depth's
// queue and the previous depth's shadow rays, clears the next queue, the
// shadow queue and the extend work counters.
__global__ void k_pt_setup(PathState ps, int depth)
{
    const uint32_t t = threadIdx.x;
    const int cur = depth & 1;
    const uint32_t nq = wave_sum(ps.stripes[cur * kStripes + t]);
    const uint32_t ns = wave_sum(ps.stripes[2 * kStripes + t]);
    ps.stripes[(1 - cur) * kStripes + t] = 0;
    ps.stripes[2 * kStripes + t] = 0;
    ps.work[t] = 0;
    if (t == 0) {
        unsigned long long* tot = ray_totals(ps);
        tot[0] += nq; // rays extended at this depth
        tot[1] += ns; // shadow rays of the previous depth (0 at depth 0)
        ps.counters[cur] = nq;
        if (depth > 0) ps.counters[2] = ns;
        poll_abort(ps);
    }
}

__global__ void k_pt_tally(PathState ps)
{
    const uint32_t ns = wave_sum(ps.stripes[2 * kStripes + threadIdx.x]);
    if (threadIdx.x == 0) {
        ray_totals(ps)[1] += ns;
        ps.counters[2] = ns;
    }
}

__global__ void k_totals_clear(PathState ps)
{
    if (threadIdx.x < 2) ray_totals(ps)[threadIdx.x] = 0;
}

static_assert(2 * kWideStack <= kStackDepth, "the wide walk's two stacks share the packed walk's LDS");

// `stack` is this lane's column of 2 S words (BLOCK apart): the wide walks'
// two node-group stacks of S entries, or, with S = kWideStack, the packed
// walk's stack of kStackDepth words.  With S = kWideStackShallow (trees of at
// most S + 1 wide levels) the packed walk's stack does not fit, so a view
// without a wide tree takes the stackless skip-link walk (same hits).
// TL: the view's scene is a two-level structure (stacks of kWideStack).
template <int BLOCK, bool TL, int S = kWideStack>
__device__ __forceinline__ void extend_one(const SceneView& sc, PathState& ps, int cur, uint32_t e, uint32_t* stack,
                                           int binary)
{
    static_assert(S == kWideStack || S == kWideStackShallow, "wide stack sizes");
    constexpr bool kPackedFits = 2 * S >= kStackDepth;
    const uint32_t slot = ps.queue[cur][e];
    const RayF r = make_ray(ps.rayO[slot], ps.rayD[slot]);
    const SceneDesc& d = sc.scenes[ps.scene[slot]];
    HitF h;
    uint32_t ik;
    if (TL)
        closest_two_level(sc.bvh, r, stack, stack + kWideStack * BLOCK, BLOCK, h, ik);
    else if ((!kPackedFits || !binary) && d.wideRoot != STK_INVALID_INDEX)
        closest_wide(sc.bvh, d.wideRoot, r, stack, stack + S * BLOCK, BLOCK, h);
    else if (kPackedFits)
        closest_packed(sc.bvh, d.hdr, r, stack, BLOCK, h);
    else
        closest_ref(sc.bvh, d.hdr, d.refRoot, r, h);
    ps.hit[e] = make_float4(h.u, h.v, h.t, bits_to_float(h.inst));
    ps.hitPrim[e] = h.prim;
}

// Closest hit of every queued path.  persistent = 0: one lane per queue
// entry, stripe-major grid.  persistent = 1: a fixed grid of waves; wave w
// takes 64 entries at a time from segment (w + k) % kStripes, k = 0, 1, ...,
// moving on when a segment is exhausted (a plain load checks before the
// atomic), so every wave ends after at most kStripes segments.
template <int BLOCK, bool TL, int S>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(S == kWideStackShallow ? 7 : 5, 8))) void k_pt_extend(SceneView sc, PathState ps, int depth, int persistent,
                                                     int binary)
{
    __shared__ uint32_t stack[2 * (TL ? kWideStack : S) * BLOCK];
    if (pass_aborted(ps)) return;
    const int cur = depth & 1;
    const uint32_t* counts = ps.stripes + cur * kStripes;
    // the shallow-stack form is launched with stripe grids only
    if (S == kWideStackShallow || !persistent) {
        const uint32_t stripe = blockIdx.x % kStripes, cnt = counts[stripe];
        for (uint32_t base = stripe_first(BLOCK); base < cnt; base += stripe_step(BLOCK)) {
            const uint32_t local = base + threadIdx.x;
            if (local < cnt)
                extend_one<BLOCK, TL, S>(sc, ps, cur, stripe * ps.segCap + local, stack + threadIdx.x, binary);
        }
        return;
    }
    const uint32_t lane = threadIdx.x % 64;
    const uint32_t wave = (blockIdx.x * BLOCK + threadIdx.x) / 64;
    uint32_t fetched = 0;
    for (uint32_t k = 0; k < kStripes; ++k) {
        const uint32_t sg = (wave + k) % kStripes;
        const uint32_t cnt = counts[sg];
        for (;;) {
            // drain: every 16th fetch lane 0 polls the watchdog's host word;
            // once it is set the wave stops taking entries (the pass's later
            // kernels return at entry)
            uint32_t stop = 0;
            if (lane == 0 && (++fetched & 15u) == 0u && ps.abortHost &&
                __hip_atomic_load(ps.abortHost, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u) {
                ps.counters[kAbortCounter] = 1u;
                *ps.abortAny = 1u;
                stop = 1;
            }
            if (__shfl((int)stop, 0)) return;
            uint32_t base = cnt;
            if (lane == 0 && __hip_atomic_load(&ps.work[sg], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < cnt)
                base = atomicAdd(&ps.work[sg], 64u);
            base = (uint32_t)__shfl((int)base, 0);
            if (base >= cnt) break;
            const uint32_t q = base + lane;
            if (q < cnt) extend_one<BLOCK, TL, S>(sc, ps, cur, sg * ps.segCap + q, stack + threadIdx.x, binary);
        }
    }
}

// One queue entry of the shade step (active: q is a valid entry; every lane of
// the block calls it, for the wave-wide appends at the end).
// The continuing path goes to the next path queue at nextBase + (count
// appended through nextCnt), a shadow ray to the shadow queue likewise.
__device__ __forceinline__ void shade_one(const SceneView& sc, const PtLaunch& lp, PathState& ps, int depth, bool active,
                                          uint32_t q, uint32_t* nextCnt, uint32_t nextBase, uint32_t* shadowCnt,
                                          uint32_t shadowBase)
{
    const int cur = depth & 1;
    bool cont = false, shadow = false, cacheQuery = false;
    uint32_t slot = 0;
    F3 nO = mk3(0, 0, 0), nD = mk3(0, 0, 0);
    float nPdf = 0.0f; // solid-angle pdf of the sampled direction (corrected estimator's MIS)
    F3 sO = mk3(0, 0, 0), sD = mk3(0, 0, 0), cd = mk3(0, 0, 0), cg = mk3(0, 0, 0);
    float sT = 0.0f;
    CacheRecord qrec;
    const PtCache& pc = lp.cache;
    if (active) {
        slot = ps.queue[cur][q];
        const SceneDesc& desc = sc.scenes[ps.scene[slot]];
        const float4 h = ps.hit[q];
        const uint32_t inst = float_to_bits(h.w);
        F3 thr = f3(ps.thr[slot]);
        F3 col = f3(ps.col[slot]);
        uint32_t rng = ps.rng[slot];
        const bool isTrain = is_train(pc, *ps.pass, slot);
        if (isTrain && depth > 0) {
            // the radiance this path had gathered when it left vertex depth - 1
            // (that vertex's direct light included): its training target is
            // what the path gathers from here on
            float4* prev = pc.trainVerts + ((size_t)(slot / pc.trainStride) * lp.maxDepth + depth - 1) * kTrainWords;
            prev[5] = make_float4(col.x, col.y, col.z, 1.0f);
        }
        if (inst == STK_INVALID_INDEX) {
            col = col + thr * mk3(desc.sky[0], desc.sky[1], desc.sky[2]);
            ps.col[slot] = f4(col);
        } else {
            const StkInstance& ic = sc.inst[inst];
            if (ic.lightId != -1) {
                const bool own = (uint32_t)ic.lightId < desc.lightBase + desc.lightCount;
                // rayD.w < 0: the ray left a specular vertex (specularBounce)
                const bool afterSpecular = depth > 0 && ps.rayD[slot].w < 0.0f;
                if (lp.refCompat) {
                    // emitter: counted on camera rays and after specular bounces
                    // (shaders/pathtracerMdl.hlsl:229-234); a BSDF-sampled emitter
                    // hit after a diffuse or glossy bounce adds nothing and ends the path
                    if ((depth == 0 || afterSpecular) && own) {
                        col = col + thr * ld3(sc.lights[ic.lightId].color);
                        ps.col[slot] = f4(col);
                    }
                } else if (own) {
                    // corrected: the front face emits; after a diffuse or glossy
                    // bounce the hit is weighted against the light sample of the
                    // previous vertex (balance heuristic, both pdfs in solid angle)
                    const StkLight& l = sc.lights[ic.lightId];
                    if (dot(f3(ps.rayD[slot]), rect_light_normal(l)) < 0.0f) {
                        float w = 1.0f;
                        if (depth > 0 && !afterSpecular) {
                            const float bpdf = ps.rayD[slot].w;
                            const float S = sphquad_init(l, f3(ps.rayO[slot])).S;
                            const float lpdf = S > 0.0f ? (1.0f / (float)desc.lightCount) / S : 0.0f;
                            w = bpdf / (bpdf + lpdf);
                        }
                        col = col + (thr * ld3(l.color)) * w;
                        ps.col[slot] = f4(col);
                    }
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
                StkMaterial mat;
                if (ic.materialId >= 0 && (uint32_t)ic.materialId < sc.numMats) {
                    mat = sc.mats[ic.materialId];
                } else {
                    mat.type = STK_MAT_LAMBERT;
                    mat.roughness = 1.0f;
                    mat.albedoTexture = mat.roughnessTexture = mat.normalTexture = mat.emissionTexture = 0;
                    mat.albedo[0] = mat.albedo[1] = mat.albedo[2] = 0.5f;
                    mat.albedo[3] = 1.0f;
                    mat.emission[0] = mat.emission[1] = mat.emission[2] = 0.0f;
                    mat.ior = 0.0f;
                }
                F3 N = nrm(interp(n0, n1, n2, h.x, h.y));
                F3 emit = ld3(mat.emission);
                const bool albTex = mat.albedoTexture != 0 && mat.albedoTexture <= sc.numTextures;
                const bool roughTex = mat.roughnessTexture != 0 && mat.roughnessTexture <= sc.numTextures;
                const bool normTex = mat.normalTexture != 0 && mat.normalTexture <= sc.numTextures;
                const bool emitTex = mat.emissionTexture != 0 && mat.emissionTexture <= sc.numTextures;
                if (albTex || roughTex || normTex || emitTex) {
                    // texture coordinates as shaders/pathtracerMdl.hlsl:261-265
                    float s0, t0, s1, t1, s2, t2;
                    unpack_uv(v0.uv, s0, t0);
                    unpack_uv(v1.uv, s1, t1);
                    unpack_uv(v2.uv, s2, t2);
                    const float w = 1.0f - h.x - h.y;
                    const float s = (s0 * w + s1 * h.x) + s2 * h.y, t = (t0 * w + t1 * h.x) + t2 * h.y;
                    if (albTex) {
                        const F3 tc = tex_lookup(sc.textures[mat.albedoTexture - 1], sc.texels, s, t);
                        mat.albedo[0] = mat.albedo[0] * tc.x;
                        mat.albedo[1] = mat.albedo[1] * tc.y;
                        mat.albedo[2] = mat.albedo[2] * tc.z;
                    }
                    if (roughTex)
                        mat.roughness = mat.roughness * tex_lookup(sc.textures[mat.roughnessTexture - 1], sc.texels, s, t).x;
                    if (emitTex) emit = emit * tex_lookup(sc.textures[mat.emissionTexture - 1], sc.texels, s, t);
                    if (normTex) {
                        // the world tangent as the reference forms it (:247-257)
                        const F3 ta = xform3(ic.normalMatrix, unpack_normal(v0.tangent));
                        const F3 tb = xform3(ic.normalMatrix, unpack_normal(v1.tangent));
                        const F3 tcn = xform3(ic.normalMatrix, unpack_normal(v2.tangent));
                        const F3 T = nrm(interp(ta, tb, tcn, h.x, h.y));
                        const F3 bn = bumped_normal(N, T, tex_lookup(sc.textures[mat.normalTexture - 1], sc.texels, s, t));
                        // a tangent along the normal leaves no frame: keep the normal
                        if (!(isnan(bn.x) || isnan(bn.y) || isnan(bn.z))) N = bn;
                    }
                }
                const F3 k1 = -f3(ps.rayD[slot]);
                // the EDF: the front face emits (shaders/pathtracerMdl.hlsl:299-304)
                if ((emit.x != 0.0f || emit.y != 0.0f || emit.z != 0.0f) && dot(k1, Ng) > 0.0f) {
                    col = col + thr * emit;
                    ps.col[slot] = f4(col);
                }
                const bool specular = is_specular(mat);
                const float rough = mat.type == STK_MAT_GGX ? mat.roughness : (specular ? 0.0f : 1.0f);
                // the cache learns radiance divided by the vertex's albedo (a
                // Lambert vertex's outgoing radiance is albedo times a smooth
                // irradiance term), so albedo detail need not be learned
                const F3 alb = mk3(fmaxf(mat.albedo[0], kMinAlbedo), fmaxf(mat.albedo[1], kMinAlbedo),
                                   fmaxf(mat.albedo[2], kMinAlbedo));
                // a training path past its suffix queries like the others (self-training)
                const bool selfQuery =
                    isTrain && pc.trainSuffix != 0u && depth >= (int)(pc.queryDepth + pc.trainSuffix);
                if ((pc.mode == 1 || pc.mode == 2) && (!isTrain || selfQuery) && !specular &&
                    depth >= (int)pc.queryDepth && rough >= pc.queryMinRough) {
                    // the path ends here after this vertex's direct light: the
                    // radiance scattered from the BSDF-sampled direction on
                    // comes from the cache (a training path's earlier vertices
                    // then learn from that answer; this vertex is no record)
                    cacheQuery = true;
                    qrec.posRough = make_float4(P.x, P.y, P.z, rough);
                    qrec.normal = make_float4(N.x, N.y, N.z, bits_to_float(slot));
                    qrec.dir = make_float4(k1.x, k1.y, k1.z, 0.0f);
                    qrec.target = make_float4(alb.x, alb.y, alb.z, 0.0f); // read back by k_cache_apply
                } else if (isTrain) {
                    float4* v = pc.trainVerts + ((size_t)(slot / pc.trainStride) * lp.maxDepth + depth) * kTrainWords;
                    v[0] = make_float4(P.x, P.y, P.z, rough);
                    v[1] = make_float4(N.x, N.y, N.z, alb.x);
                    v[2] = make_float4(k1.x, k1.y, k1.z, alb.y);
                    v[3] = make_float4(col.x, col.y, col.z, alb.z);
                    v[4] = f4(thr);
                    v[5] = make_float4(0.0f, 0.0f, 0.0f, 0.0f); // set when the path goes on
                    pc.trainCount[slot / pc.trainStride] = (uint32_t)depth + 1;
                }
                // sampleLights / estimateDirectLighting (shaders/pathtracerMdl.hlsl:31-92)
                F3 toLight = mk3(0, 0, 0), radiance = mk3(0, 0, 0);
                float lightPdf = 0.0f, dist = 0.0f;
                if (desc.lightCount > 0) {
                    const uint32_t lid = (uint32_t)((float)desc.lightCount * rnd(rng));
                    const float selPdf = 1.0f / (float)desc.lightCount;
                    const StkLight& l =
                        sc.lights[desc.lightBase + (lid < desc.lightCount ? lid : desc.lightCount - 1)];
                    const float u0 = rnd(rng);
                    const float u1 = rnd(rng);
                    const LightSample ls = sample_rect_light(l, u0, u1, P);
                    toLight = ls.L;
                    const F3 Li = ld3(l.color);
                    if (dot(N, toLight) > 0.0f && -dot(toLight, ls.normal) > 0.0f && Li.x != 0.0f && Li.y != 0.0f &&
                        Li.z != 0.0f) {
                        dist = ls.dist;
                        if (lp.refCompat) {
                            // the area-measure pdf against solid-angle samples (Q6), and
                            // the cosine the BSDF evaluation already carries applied again
                            lightPdf = ls.pdf;
                            radiance = (Li * 1.0f) * saturate(dot(N, toLight)); // visibility applied by connect
                        } else {
                            lightPdf = ls.solidAngle > 0.0f ? 1.0f / ls.solidAngle : 0.0f;
                            radiance = Li;
                        }
                    }
                    lightPdf *= selPdf;
                }
                bool alive = !(isnan(radiance.x) || isnan(radiance.y) || isnan(radiance.z) || isnan(lightPdf));
                // a query that covers the vertex's direct light sends no shadow ray
                if (alive && (dot(toLight, Ng) > 0.0f) && lightPdf != 0.0f && !(cacheQuery && pc.queryDirect)) {
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
                    const float x2 = rnd(rng);
                    (void)rnd(rng);
                    const BsdfSample bs = bsdf_sample(mat, N, Ng, k1, x0, x1, x2);
                    if (bs.event != 0) {
                        thr = thr * bs.weight;
                        bool survive = true;
                        if (depth > (int)lp.rrStartDepth) {
                            if (lp.refCompat) {
                                // p = max(rgb), unclamped, and 1 / (p + 1e-5) (Q9)
                                const float pr = fmaxf(thr.x, fmaxf(thr.y, thr.z));
                                if (rnd(rng) > pr)
                                    survive = false;
                                else
                                    thr = thr * (1.0f / (pr + 1e-5f));
                            } else {
                                const float pr = fminf(fmaxf(thr.x, fmaxf(thr.y, thr.z)), 1.0f);
                                if (rnd(rng) > pr || !(pr > 0.0f))
                                    survive = false;
                                else
                                    thr = thr * (1.0f / pr);
                            }
                        }
                        if (survive) {
                            cont = true;
                            // offset to the side the new ray leaves by (:368-370)
                            nO = offset_ray(P, dot(bs.k2, Ng) > 0.0f ? Ng : -Ng);
                            nD = bs.k2;
                            nPdf = bs.event >= kEventSpecularReflection ? -1.0f : (lp.refCompat ? 0.0f : bs.pdf);
                        }
                    }
                }
            }
        }
        ps.rng[slot] = rng;
        if (cont) {
            ps.thr[slot] = f4(thr);
            ps.rayO[slot] = make_float4(nO.x, nO.y, nO.z, 1e9f);
            ps.rayD[slot] = make_float4(nD.x, nD.y, nD.z, nPdf);
        }
    }
    const uint32_t qi = wave_push(nextCnt, cont);
    if (cont) ps.queue[1 - cur][nextBase + qi] = slot;
    const uint32_t si = wave_push(shadowCnt, shadow);
    if (shadow) {
        const uint32_t se = shadowBase + si;
        ps.shadowQ[se] = slot;
        ps.shO[se] = make_float4(sO.x, sO.y, sO.z, sT);
        ps.shD[se] = make_float4(sD.x, sD.y, sD.z, 0.0f);
        ps.shCd[se] = f4(cd);
        ps.shCg[se] = f4(cg);
    }
    if (pc.mode == 1 || pc.mode == 2) {
        const uint32_t ci = wave_push(&ps.counters[3], cacheQuery);
        if (cacheQuery) pc.queryRecs[ci] = qrec;
    }
}

__global__ __launch_bounds__(kBlock) void k_pt_shade(SceneView sc, PtLaunch lp, PathState ps, int depth)
{
    if (pass_aborted(ps)) return;
    const int cur = depth & 1;
    const uint32_t stripe = blockIdx.x % kStripes, cnt = ps.stripes[cur * kStripes + stripe];
    uint32_t* nextCnt = &ps.stripes[(1 - cur) * kStripes + stripe];
    uint32_t* shadowCnt = &ps.stripes[2 * kStripes + stripe];
    const uint32_t segBase = stripe * ps.segCap;
    for (uint32_t base = stripe_first(kBlock); base < cnt; base += stripe_step(kBlock)) {
        const uint32_t local = base + threadIdx.x;
        shade_one(sc, lp, ps, depth, local < cnt, segBase + local, nextCnt, segBase, shadowCnt, segBase);
    }
}

// Emitter triangles a shadow ray tests before the wide walk (beyond this the
// scene's shadow rays take the skip-link walk).
constexpr uint32_t kMaxEmitters = 64;

// One shadow-ray entry s of the connect step.
// refCompat: the reference's rule (the first hit leaf in preorder decides,
// and a light hit counts as visible, Q5); otherwise any non-emitter triangle
// before tmax occludes (the emitter walks with no emitter triangles given).
template <bool TL, int BLOCK, int S = kWideStack>
__device__ __forceinline__ void connect_one(const SceneView& sc, PathState& ps, uint32_t s, uint32_t* stack, int binary,
                                            int refCompat)
{
    const uint32_t slot = ps.shadowQ[s];
    const SceneDesc& desc = sc.scenes[ps.scene[slot]];
    const RayF r = make_ray(ps.shO[s], ps.shD[s]);
    const uint32_t emitCount = refCompat ? desc.emitCount : 0u;
    bool vis;
    if (TL) {
        vis = visible_two_level(sc.bvh, sc.bvh.emitters + 3 * (size_t)desc.emitBase, emitCount, r,
                                stack + threadIdx.x, stack + kWideStack * BLOCK + threadIdx.x, BLOCK);
    } else if (!binary && desc.wideRoot != STK_INVALID_INDEX && desc.emitCount <= kMaxEmitters) {
        vis = visible_wide(sc.bvh, desc.wideRoot, sc.bvh.emitters + 3 * (size_t)desc.emitBase, emitCount, r,
                           stack + threadIdx.x, stack + S * BLOCK + threadIdx.x, BLOCK);
    } else if (!refCompat) {
        vis = !occluded_ref(sc.bvh, desc.hdr, desc.refRoot, r);
    } else {
        HitF h;
        vis = !any_ref(sc.bvh, desc.hdr, desc.refRoot, r, h);
        if (!vis && sc.inst[h.inst].lightId != -1) vis = true; // a light hit counts as visible
    }
    if (vis) {
        F3 col = f3(ps.col[slot]);
        col = col + f3(ps.shCd[s]);
        col = col + f3(ps.shCg[s]);
        ps.col[slot] = f4(col);
    }
}

// Shadow rays: the wide visibility walk (the reference's first-in-preorder
// any-hit rule, decided in any order), or the stackless skip-link walk when
// the scene has no usable wide tree or many emitter triangles; TL: the
// two-level visibility walk.  One-wave blocks (a finished wave frees its
// stacks at once); S wide stack entries per lane.
template <bool TL, int S>
__global__ __launch_bounds__(kWaveBlock) __attribute__((amdgpu_waves_per_eu(S == kWideStackShallow ? 7 : 5, 8))) void k_pt_connect(SceneView sc, PathState ps, int binary,
                                                                                         int refCompat)
{
    __shared__ uint32_t stack[2 * (TL ? kWideStack : S) * kWaveBlock];
    if (pass_aborted(ps)) return;
    const uint32_t stripe = blockIdx.x % kStripes, cnt = ps.stripes[2 * kStripes + stripe];
    for (uint32_t base = stripe_first(kWaveBlock); base < cnt; base += stripe_step(kWaveBlock)) {
        const uint32_t local = base + threadIdx.x;
        if (local < cnt)
            connect_one<TL, kWaveBlock, S>(sc, ps, stripe * ps.segCap + local, stack, binary, refCompat);
    }
}


// Adds the cache's answer to the paths that ended at a query:
// radiance += throughput * albedo * max(prediction, 0).
__global__ __launch_bounds__(kBlock) void k_cache_apply(PtLaunch lp, PathState ps)
{
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (pass_aborted(ps) || q >= ps.counters[3]) return;
    const CacheRecord& rec = lp.cache.queryRecs[q];
    const uint32_t slot = float_to_bits(rec.normal.w);
    const float4 o = lp.cache.queryOut[q];
    const F3 w = f3(ps.thr[slot]) * f3(rec.target);
    const F3 col = f3(ps.col[slot]) + w * mk3(fmaxf(o.x, 0.0f), fmaxf(o.y, 0.0f), fmaxf(o.z, 0.0f));
    ps.col[slot] = f4(col);
}

// One slot's sample into its pixel (reduction.hlsl:11-32: color += invSpp *
// sample, in sample order) and, when kept, into the samples buffer.
__device__ __forceinline__ void finalize_one(const PtLaunch& lp, const PathState& ps, float4* __restrict__ accum,
                                             float* __restrict__ samples, uint32_t local, uint32_t sampleNum)
{
    const uint32_t pixel = ps.pixel[local];
    if (pixel == STK_INVALID_INDEX) return;
    const float4 c = ps.col[local];
    float4 a = accum[pixel];
    a.x = a.x + lp.invSpp * c.x;
    a.y = a.y + lp.invSpp * c.y;
    a.z = a.z + lp.invSpp * c.z;
    a.w = 1.0f;
    accum[pixel] = a;
    if (samples) {
        float* o = samples + ((size_t)pixel * lp.spp + sampleNum) * 3;
        o[0] = c.x;
        o[1] = c.y;
        o[2] = c.z;
    }
}

__global__ __launch_bounds__(kBlock) void k_pt_finalize(PtLaunch lp, PathState ps, float4* __restrict__ accum,
                                                        float* __restrict__ samples)
{
    const uint32_t local = blockIdx.x * kBlock + threadIdx.x;
    if (pass_aborted(ps) || local >= ps.pass->slotCount) return;
    finalize_one(lp, ps, accum, samples, local, ps.pass->sampleNum);
}

// Training targets: the radiance a training path gathered after leaving vertex
// d (everything but d's own direct-light sample, which counts too when
// queries cover their vertex's direct light), divided by the throughput at
// d and by d's albedo - a one-sample estimate, under the plain path tracer's
// estimator, of what a query at d stands in for.
__global__ __launch_bounds__(kBlock) void k_cache_targets(PtLaunch lp, PathState ps)
{
    if (pass_aborted(ps)) return;
    const PtCache& pc = lp.cache;
    const PassState pass = *ps.pass;
    const uint32_t maxDepth = lp.maxDepth;
    const uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t tidx = id / maxDepth, d = id % maxDepth;
    bool emit = false;
    CacheRecord rec;
    if (tidx < pass.numTrainPaths && d < pc.trainCount[tidx]) {
        const uint32_t slot = tidx * pc.trainStride + pass.trainOffset;
        const float4* v = pc.trainVerts + ((size_t)tidx * maxDepth + d) * kTrainWords;
        const float4 thr = v[4], fin = ps.col[slot];
        // queries that cover their vertex's direct light learn it too: the
        // target starts before d's light sample (v[3], after d's emission);
        // otherwise after it (v[5]; a path that ended at d gathered nothing
        // after it)
        const float4 before = pc.queryDirect ? v[3] : (v[5].w != 0.0f ? v[5] : fin);
        const float ax = v[1].w, ay = v[2].w, az = v[3].w; // the vertex's albedo
        if (thr.x > 1e-8f && thr.y > 1e-8f && thr.z > 1e-8f) {
            // one-sample targets are clamped to [0, kTargetMax]: Russian-roulette
            // survivors and near-light NEE samples can be arbitrarily large, and
            // a few such fireflies would dominate the loss
            const float kTargetMax = 100.0f;
            rec.posRough = v[0];
            rec.normal = make_float4(v[1].x, v[1].y, v[1].z, 0.0f);
            rec.dir = make_float4(v[2].x, v[2].y, v[2].z, 0.0f);
            rec.target = make_float4(fminf(fmaxf((fin.x - before.x) / (thr.x * ax), 0.0f), kTargetMax),
                                     fminf(fmaxf((fin.y - before.y) / (thr.y * ay), 0.0f), kTargetMax),
                                     fminf(fmaxf((fin.z - before.z) / (thr.z * az), 0.0f), kTargetMax), 0.0f);
            emit = isfinite(fin.x) && isfinite(fin.y) && isfinite(fin.z);
        }
    }
    const uint32_t i = wave_push(&ps.counters[4], emit);
    if (emit && i < pc.trainCap) pc.trainBatch[i] = rec;
}

// One launch per frame over every view and every sample.  Resident workgroups
// of BLOCK threads (one wave or four) take chunks of `chunk` consecutive slots
// (a multiple of 256; 256 slots are 4 raygen tiles of 8x8 at the default tile) from a
// counter and run all of a chunk's samples, in sample order, through raygen,
// the bounces (closest hit, shading, shadow rays) and the reduction, with the
// path state, queues and hits in a region private to the workgroup (entries
// base .. base + chunk - 1) and the queue counters in LDS; a stage loops over
// its queue BLOCK entries at a time, so a chunk larger than the workgroup
// keeps its lanes busy in the late bounces.  Stages are separated by
// workgroup barriers only: no wave ever waits for another workgroup, and
// every workgroup leaves when the counter passes the last chunk (or at the
// next sample once the watchdog's word is set).  A chunk's slots are its
// pixels alone, so every pixel adds its samples in sample order and each path
// runs the same per-path functions as the per-stage launches: the image is
// theirs bit for bit.  Plain rendering only (the radiance cache trains on
// whole frames).
template <int BLOCK, bool TL>
__global__ __launch_bounds__(BLOCK) void k_pt_frame(SceneView sc, PtLaunch lp, PathState ps, uint32_t frameNumber,
                                                     uint32_t totalSlots, uint32_t chunk, int extendBinary,
                                                     int connectBinary, float4* __restrict__ accum,
                                                     float* __restrict__ samples)
{
    __shared__ uint32_t stack[kStackDepth * BLOCK];
    __shared__ uint32_t counts[3]; // path queues 0 and 1, shadow queue
    __shared__ uint32_t taken, stopped;
    const uint32_t t = threadIdx.x;
    const uint32_t base = blockIdx.x * chunk; // this workgroup's private slots and queue entries
    const uint32_t numChunks = (totalSlots + chunk - 1) / chunk;
    unsigned long long closest = 0, shadows = 0; // rays this workgroup traced (thread 0)
    bool stop = false;
    while (!stop) {
        if (t == 0) taken = pass_aborted(ps) ? numChunks : atomicAdd(&ps.work[0], 1u);
        __syncthreads();
        const uint32_t c = taken;
        if (c >= numChunks) break;
        const uint32_t first = c * chunk; // the chunk's first global slot
        PassState pass{};
        pass.frameNumber = frameNumber;
        pass.slotBegin = first;
        pass.slotCount = min(chunk, totalSlots - first);
        for (uint32_t smp = 0; smp < lp.spp; ++smp) {
            pass.sampleNum = smp;
            if (t == 0) {
                counts[0] = counts[1] = counts[2] = 0u;
                uint32_t a = 0u;
                if (ps.abortHost &&
                    __hip_atomic_load(ps.abortHost, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u) {
                    ps.counters[kAbortCounter] = 1u;
                    *ps.abortAny = 1u;
                    a = 1u;
                }
                stopped = a;
            }
            __syncthreads();
            if (stopped) {
                stop = true;
                break;
            }
            for (uint32_t i = t; i < chunk; i += BLOCK) { // the loop bound is uniform: whole waves append
                const uint32_t ws = base + i, slot = first + i;
                bool valid = false;
                if (slot < totalSlots)
                    valid = raygen_one(sc, lp, ps, pass, ws, slot);
                else
                    ps.pixel[ws] = STK_INVALID_INDEX;
                const uint32_t qi = wave_push(&counts[0], valid);
                if (valid) ps.queue[0][base + qi] = ws;
            }
            __syncthreads();
            for (int depth = 0; depth < (int)lp.maxDepth; ++depth) {
                const int cur = depth & 1;
                const uint32_t nq = counts[cur];
                if (nq == 0u) break; // every thread read the same count after the barrier
                for (uint32_t e = t; e < nq; e += BLOCK)
                    extend_one<BLOCK, TL>(sc, ps, cur, base + e, stack + t, extendBinary);
                if (t == 0) {
                    counts[1 - cur] = 0u;
                    counts[2] = 0u;
                    closest += nq;
                }
                __syncthreads();
                const uint32_t rounds = (nq + BLOCK - 1) / BLOCK;
                for (uint32_t r = 0; r < rounds; ++r) {
                    const uint32_t e = r * BLOCK + t;
                    shade_one(sc, lp, ps, depth, e < nq, base + e, &counts[1 - cur], base, &counts[2], base);
                }
                __syncthreads();
                const uint32_t ns = counts[2];
                if (t == 0) shadows += ns;
                for (uint32_t e = t; e < ns; e += BLOCK)
                    connect_one<TL, BLOCK>(sc, ps, base + e, stack, connectBinary, (int)lp.refCompat);
                __syncthreads();
            }
            for (uint32_t i = t; i < chunk; i += BLOCK)
                if (first + i < totalSlots) finalize_one(lp, ps, accum, samples, base + i, smp);
            __syncthreads();
        }
    }
    if (t == 0) {
        unsigned long long* tot = ray_totals(ps);
        atomicAdd(&tot[0], closest);
        atomicAdd(&tot[1], shadows);
    }
}

inline uint32_t blocks(size_t n, uint32_t b = kBlock) { return (uint32_t)((n + b - 1) / b); }

// Stripe-major consumer grid of B-thread blocks (segCap is a multiple of 256).
inline dim3 stripeGrid(const PathState& ps, uint32_t B, uint32_t stripeThreads = 0)
{
    uint32_t perStripe = ps.segCap / B;
    if (stripeThreads) perStripe = std::min(perStripe, std::max(1u, stripeThreads / B));
    return dim3(kStripes * perStripe);
}

template <int BLOCK, bool TL, int S>
void launchExtendWalk(const SceneView& sc, PathState& ps, int depth, const LaunchConfig& lc, int binary, hipStream_t s)
{
    if (lc.persistentBlocks)
        hipLaunchKernelGGL((k_pt_extend<BLOCK, TL, S>), dim3(lc.persistentBlocks), dim3(BLOCK), 0, s, sc, ps, depth,
                           1, binary);
    else
        hipLaunchKernelGGL((k_pt_extend<BLOCK, TL, S>), stripeGrid(ps, BLOCK, lc.stripeThreads), dim3(BLOCK), 0, s,
                           sc, ps, depth, 0, binary);
}

// Whether every view's wide walks fit stacks of kWideStackShallow entries:
// wideDepth is the deepest view's wide tree (0: none).
inline bool shallowWalks(const SceneView& sc)
{
    return !sc.bvh.tlas && sc.bvh.wideDepth > 0 && sc.bvh.wideDepth <= (uint32_t)kWideStackShallow + 1;
}

template <int BLOCK>
void launchExtend(const SceneView& sc, PathState& ps, int depth, const LaunchConfig& lc, hipStream_t s)
{
    // the binary walk only where its stack holds the tree
    const int binary = (lc.binaryWalk && sc.bvh.packedOk) ? 1 : 0;
    if (sc.bvh.tlas)
        launchExtendWalk<BLOCK, true, kWideStack>(sc, ps, depth, lc, binary, s);
    else if (!binary && !lc.persistentBlocks && shallowWalks(sc))
        launchExtendWalk<BLOCK, false, kWideStackShallow>(sc, ps, depth, lc, 0, s);
    else
        launchExtendWalk<BLOCK, false, kWideStack>(sc, ps, depth, lc, binary, s);
}

} // namespace

hipError_t launchSetPass(PathState& ps, const PassState& v, hipStream_t s)
{
    hipLaunchKernelGGL(k_pass_set, dim3(1), dim3(64), 0, s, ps.pass, v);
    return hipGetLastError();
}

hipError_t launchPathTraceFrame(const SceneView& sc, const PtLaunch& lp, const LaunchConfig& lc, PathState& ps,
                                float4* accum, float* samples, uint32_t frameNumber, uint32_t totalSlots,
                                hipStream_t s)
{
    const uint32_t chunk = lc.frameChunk ? lc.frameChunk : (uint32_t)kBlock;
    const size_t region = (size_t)lc.frameBlocks * chunk;
    if (lc.frameBlocks == 0 || chunk % 256 != 0 || region > ps.capacity || region > ps.queueCapacity ||
        lp.cache.mode != 0 || totalSlots == 0)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_pass_reset, dim3(1), dim3(64), 0, s, ps); // counters, chunk counter, abort word
    const int eb = (lc.binaryWalk && sc.bvh.packedOk) ? 1 : 0, cb = lc.binaryWalk ? 1 : 0;
    // one-wave workgroups for blockSize 64 (a finished wave waits for no
    // other wave at the stage barriers), else four waves
    if (lc.blockSize == 64) {
        if (sc.bvh.tlas)
            hipLaunchKernelGGL((k_pt_frame<64, true>), dim3(lc.frameBlocks), dim3(64), 0, s, sc, lp, ps, frameNumber,
                               totalSlots, chunk, eb, 0, accum, samples);
        else
            hipLaunchKernelGGL((k_pt_frame<64, false>), dim3(lc.frameBlocks), dim3(64), 0, s, sc, lp, ps,
                               frameNumber, totalSlots, chunk, eb, cb, accum, samples);
    } else if (sc.bvh.tlas) {
        hipLaunchKernelGGL((k_pt_frame<kBlock, true>), dim3(lc.frameBlocks), dim3(kBlock), 0, s, sc, lp, ps,
                           frameNumber, totalSlots, chunk, eb, 0, accum, samples);
    } else {
        hipLaunchKernelGGL((k_pt_frame<kBlock, false>), dim3(lc.frameBlocks), dim3(kBlock), 0, s, sc, lp, ps,
                           frameNumber, totalSlots, chunk, eb, cb, accum, samples);
    }
    return hipGetLastError();
}

hipError_t launchClearRayTotals(PathState& ps, hipStream_t s)
{
    hipLaunchKernelGGL(k_totals_clear, dim3(1), dim3(64), 0, s, ps);
    return hipGetLastError();
}

hipError_t launchPathTracePass(const SceneView& sc, const PtLaunch& lp, const LaunchConfig& lc, PathState& ps,
                               float4* accum, float* samples, CacheBuffers* cb, hipStream_t s,
                               hipEvent_t finalizeAfter, hipEvent_t finalizeDone, hipEvent_t cacheAfter,
                               hipEvent_t cacheDone)
{
    const uint32_t n = lp.passSlots;
    if (n == 0 || n > ps.capacity || ps.segCap % 256 != 0 || (size_t)ps.segCap * kStripes > ps.queueCapacity ||
        (size_t)ps.segCap * kStripes < n)
        return hipErrorInvalidValue;
    const dim3 g(blocks(n)), b(kBlock);
    const PtCache& pc = lp.cache;
    const bool query = pc.mode == 1 || pc.mode == 2, train = pc.mode == 1 || pc.mode == 3 || pc.mode == 4;
    hipLaunchKernelGGL(k_pass_reset, dim3(1), dim3(64), 0, s, ps);
    hipLaunchKernelGGL(k_pt_raygen, g, b, 0, s, sc, lp, ps);
    for (int depth = 0; depth < (int)lp.maxDepth; ++depth) {
        hipLaunchKernelGGL(k_pt_setup, dim3(1), dim3(64), 0, s, ps, depth);
        switch (lc.blockSize) {
        case 64: launchExtend<64>(sc, ps, depth, lc, s); break;
        case 128: launchExtend<128>(sc, ps, depth, lc, s); break;
        default: launchExtend<256>(sc, ps, depth, lc, s); break;
        }
        hipLaunchKernelGGL(k_pt_shade, stripeGrid(ps, kBlock, lc.stripeThreads), b, 0, s, sc, lp, ps, depth);
        const int rc = lp.refCompat ? 1 : 0;
        const dim3 cg = stripeGrid(ps, kWaveBlock, lc.stripeThreads), cb(kWaveBlock);
        if (sc.bvh.tlas)
            hipLaunchKernelGGL((k_pt_connect<true, kWideStack>), cg, cb, 0, s, sc, ps, 0, rc);
        else if (shallowWalks(sc))
            hipLaunchKernelGGL((k_pt_connect<false, kWideStackShallow>), cg, cb, 0, s, sc, ps, lc.binaryWalk ? 1 : 0,
                               rc);
        else
            hipLaunchKernelGGL((k_pt_connect<false, kWideStack>), cg, cb, 0, s, sc, ps, lc.binaryWalk ? 1 : 0, rc);
    }
    hipLaunchKernelGGL(k_pt_tally, dim3(1), dim3(64), 0, s, ps);
    if (cacheAfter && cb && (query || train)) {
        // the network and its optimiser state as the previous pass's training left them
        hipError_t e = hipStreamWaitEvent(s, cacheAfter, 0);
        if (e != hipSuccess) return e;
    }
    if (query && cb) {
        hipError_t e = cacheInfer(*cb, pc.queryRecs, n, ps.counters + 3, pc.queryOut, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_cache_apply, g, b, 0, s, lp, ps);
    }
    if (finalizeAfter) {
        hipError_t e = hipStreamWaitEvent(s, finalizeAfter, 0);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_pt_finalize, g, b, 0, s, lp, ps, accum, samples);
    if (finalizeDone) {
        hipError_t e = hipEventRecord(finalizeDone, s);
        if (e != hipSuccess) return e;
    }
    if (train && cb && pc.maxTrainPaths > 0) {
        hipLaunchKernelGGL(k_cache_targets, dim3(blocks((size_t)pc.maxTrainPaths * lp.maxDepth)), b, 0, s, lp, ps);
        // trainSteps optimiser steps over the frame's records
        for (uint32_t k = 0; k < pc.trainSteps; ++k) {
            hipError_t e = cacheBackward(*cb, pc.trainBatch, pc.trainCap, ps.counters + 4, s);
            if (e != hipSuccess) return e;
            e = cacheAdam(*cb, 1.0f, ps.counters + 4, pc.trainCap, ps.counters + kAbortCounter, s);
            if (e != hipSuccess) return e;
        }
    }
    if (cacheDone) {
        hipError_t e = hipEventRecord(cacheDone, s);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

} // namespace stk
