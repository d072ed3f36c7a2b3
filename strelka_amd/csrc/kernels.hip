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
// One wave per block for the walks with LDS stacks: the stacks bound the
// occupancy whatever the block size, and single-wave blocks retire one by one.
constexpr int kWalkBlock = 64;

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

static_assert(2 * kWideStack <= kStackDepth, "the wide walk's two stacks share the packed walk's LDS");

// The ray a lane of a one-wave block takes.  width == 0: the block's 64
// consecutive rays.  width > 0: the rays are an image's, width per row in
// row-major order, and block b takes the 8x8 tile b of it (camera rays of a
// tile stay close in the scene; 64 rays of one row span 64 pixels); lanes of
// the padding tiles past the right and bottom edges get no ray.  Hits go to
// the rays' own indices either way.
__device__ __forceinline__ bool wave_ray(uint32_t width, uint32_t n, uint32_t& i)
{
    static_assert(kWalkBlock == 64, "one 8x8 tile per wave");
    i = xcd_block(blockIdx.x, gridDim.x) * kWalkBlock + threadIdx.x;
    if (width) {
        const uint32_t tilesX = (width + 7) / 8, tile = i / 64, within = i % 64;
        const uint32_t px = (tile % tilesX) * 8 + within % 8, py = (tile / tilesX) * 8 + within / 8;
        const uint64_t j = (uint64_t)py * width + px;  // the padded rows can pass 2^32 - 1
        if (px >= width || j >= n) return false;
        i = (uint32_t)j;
    }
    return i < n;
}

// Closest hit over the compressed 8-wide tree (two LDS node-group stacks per
// lane).  The binary walk has a kernel of its own (56 registers instead of
// the wide walk's 72).  The 11-entry form is held to 72 registers, so 7
// waves per SIMD fit beside its 5.5 KiB of stacks per wave.
// S: stack entries per lane.  At a node of level L (the root is level 1) the
// stack holds at most one group of remaining siblings per level 2..L, and only
// a node with inner children pushes, so a walk over a tree of D levels holds
// at most D - 1 entries: a tree of at most S + 1 levels fits.  Fewer entries
// mean less LDS per wave and more waves per CU.
template <int S>
__global__ __launch_bounds__(kWalkBlock) __attribute__((amdgpu_waves_per_eu(S <= 11 ? 7 : 5, 8))) void k_trace_closest(
    BvhView bvh, const StkRay* __restrict__ rays, uint32_t n, StkHit* __restrict__ hits, uint32_t width)
{
    __shared__ uint32_t stack[2 * S * kWalkBlock];
    uint32_t i;
    if (!wave_ray(width, n, i)) return;
    const RayF r = load_ray(rays, i);
    HitF h;
    closest_wide(bvh, bvh.wideRoot, r, stack + threadIdx.x, stack + S * kWalkBlock + threadIdx.x, kWalkBlock, h);
    store_hit(hits, i, h);
}

// Closest hit over the packed binary tree.
__global__ __launch_bounds__(kWalkBlock) void k_trace_closest_packed(BvhView bvh, const StkRay* __restrict__ rays,
                                                                 uint32_t n, StkHit* __restrict__ hits, uint32_t width)
{
    __shared__ uint32_t stack[kStackDepth * kWalkBlock];
    uint32_t i;
    if (!wave_ray(width, n, i)) return;
    const RayF r = load_ray(rays, i);
    HitF h;
    closest_packed(bvh, bvh.hdr, r, stack + threadIdx.x, kWalkBlock, h);
    store_hit(hits, i, h);
}

// Closest hit and any hit over a two-level structure.
// five waves per SIMD: the two-level walk fits 96 registers without spills
__global__ __launch_bounds__(kWalkBlock) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_trace_closest_tl(BvhView bvh, const StkRay* __restrict__ rays, uint32_t n,
                                                             StkHit* __restrict__ hits, uint32_t width)
{
    __shared__ uint32_t stack[2 * kWideStack * kWalkBlock];
    uint32_t i;
    if (!wave_ray(width, n, i)) return;
    const RayF r = load_ray(rays, i);
    HitF h;
    uint32_t ik;
    closest_two_level(bvh, r, stack + threadIdx.x, stack + kWideStack * kWalkBlock + threadIdx.x, kWalkBlock, h, ik);
    store_hit(hits, i, h);
}

__global__ __launch_bounds__(kWalkBlock) void k_trace_any_tl(BvhView bvh, const StkRay* __restrict__ rays, uint32_t n,
                                                         StkHit* __restrict__ hits)
{
    __shared__ uint32_t stack[2 * kWideStack * kWalkBlock];
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kWalkBlock + threadIdx.x;
    if (i >= n) return;
    const RayF r = load_ray(rays, i);
    HitF h;
    any_two_level(bvh, r, stack + threadIdx.x, stack + kWideStack * kWalkBlock + threadIdx.x, kWalkBlock, h);
    store_hit(hits, i, h);
}

__global__ __launch_bounds__(kBlock) void k_trace_closest_ref(BvhView bvh, const StkRay* __restrict__ rays,
                                                              uint32_t n, StkHit* __restrict__ hits)
{
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const RayF r = load_ray(rays, i);
    HitF h;
    closest_ref(bvh, bvh.hdr, 0, r, h);
    store_hit(hits, i, h);
}

__global__ __launch_bounds__(kWalkBlock) void k_trace_any(BvhView bvh, const StkRay* __restrict__ rays, uint32_t n,
                                                      StkHit* __restrict__ hits)
{
    __shared__ uint32_t stack[kStackDepth * kWalkBlock];
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kWalkBlock + threadIdx.x;
    if (i >= n) return;
    const RayF r = load_ray(rays, i);
    HitF h;
    any_packed(bvh, bvh.hdr, r, stack + threadIdx.x, kWalkBlock, h);
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
    any_ref(bvh, bvh.hdr, 0, r, h);
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

__device__ __forceinline__ F3 f3(float4 v) { return mk3(v.x, v.y, v.z); }

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
                                                               int doTonemap, uint32_t type,
                                                               const uint32_t* __restrict__ abortAny)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= npix || (abortAny && *abortAny)) return;
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

// RGBA32F pixels to a packed image of `format` (0 RGBA32F, 1 RGBA8).
__global__ __launch_bounds__(kBlock) void k_pack_image(const float4* __restrict__ src, uint32_t npix, uint32_t format,
                                                       void* __restrict__ dst)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= npix) return;
    const float4 v = src[i];
    if (format == 0) {
        reinterpret_cast<float4*>(dst)[i] = v;
        return;
    }
    auto q = [](float x) -> uint32_t {
        const float c = fminf(fmaxf(x, 0.0f), 1.0f); // minNum: NaN -> 0
        return (uint32_t)(c * 255.0f + 0.5f);
    };
    reinterpret_cast<uint32_t*>(dst)[i] = q(v.x) | (q(v.y) << 8) | (q(v.z) << 16) | (q(v.w) << 24);
}

inline uint32_t blocks(size_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }
inline uint32_t walkBlocks(size_t n) { return (uint32_t)((n + kWalkBlock - 1) / kWalkBlock); }

// Every instance's world triangles (world_triangle, the walk's own form):
// block i writes instance i's count[i] triangles, mesh leaves from
// meshBase[i] on, to worldBase[i] on.
__global__ __launch_bounds__(256) void k_world_triangles(const StkInstance* __restrict__ inst,
                                                         const float4* __restrict__ blasTris,
                                                         const uint32_t* __restrict__ meshBase,
                                                         const uint32_t* __restrict__ count,
                                                         const uint64_t* __restrict__ worldBase,
                                                         float4* __restrict__ worldTris)
{
    const uint32_t i = blockIdx.x, n = count[i];
    const float* M = inst[i].objectToWorld;
    for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) {
        float4 a, b, c;
        world_triangle(M, blasTris + 3 * (size_t)(meshBase[i] + t), a, b, c);
        float4* o = worldTris + 3 * (size_t)(worldBase[i] + t);
        o[0] = a;
        o[1] = b;
        o[2] = c;
    }
}

} // namespace

hipError_t launchPackImage(const float4* src, uint32_t npix, uint32_t format, void* dst, hipStream_t s)
{
    if (npix == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_image, dim3(blocks(npix)), dim3(kBlock), 0, s, src, npix, format, dst);
    return hipGetLastError();
}

hipError_t launchTraceClosest(const BvhView& bvh, const StkRay* rays, uint32_t n, StkHit* hits, uint32_t flags,
                              hipStream_t s, uint32_t width)
{
    if (n == 0) return hipSuccess;
    // with an image width, one wave per 8x8 tile (the last tile row and column padded)
    const uint32_t grid = width ? ((width + 7) / 8) * (uint32_t)(((n + width - 1) / width + 7) / 8) : walkBlocks(n);
    if (bvh.tlas) {
        if (flags & (STK_TRACE_REFERENCE_ORDER | STK_TRACE_BVH2)) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_trace_closest_tl, dim3(grid), dim3(kWalkBlock), 0, s, bvh, rays, n, hits, width);
    } else if (!bvh.packedOk && (flags & STK_TRACE_BVH2)) {
        return hipErrorInvalidValue; // deeper than the packed walk's stack
    } else if ((flags & STK_TRACE_REFERENCE_ORDER) || (!bvh.packedOk && bvh.wideRoot == STK_INVALID_INDEX)) {
        hipLaunchKernelGGL(k_trace_closest_ref, dim3(blocks(n)), dim3(kBlock), 0, s, bvh, rays, n, hits);
    } else if ((flags & STK_TRACE_BVH2) || bvh.wideRoot == STK_INVALID_INDEX) {
        hipLaunchKernelGGL(k_trace_closest_packed, dim3(grid), dim3(kWalkBlock), 0, s, bvh, rays, n, hits, width);
    } else {
        if (closestWalkStack(bvh) == (uint32_t)kWideStackShallow)
            hipLaunchKernelGGL(k_trace_closest<kWideStackShallow>, dim3(grid), dim3(kWalkBlock), 0, s, bvh, rays, n,
                               hits, width);
        else
            hipLaunchKernelGGL(k_trace_closest<kWideStack>, dim3(grid), dim3(kWalkBlock), 0, s, bvh, rays, n, hits,
                               width);
    }
    return hipGetLastError();
}

hipError_t launchWorldTriangles(const StkInstance* inst, const float4* blasTris, const uint32_t* meshBase,
                                const uint32_t* count, const uint64_t* worldBase, uint32_t numInstances,
                                float4* worldTris, hipStream_t s)
{
    if (numInstances == 0) return hipSuccess;
    hipLaunchKernelGGL(k_world_triangles, dim3(numInstances), dim3(256), 0, s, inst, blasTris, meshBase, count,
                       worldBase, worldTris);
    return hipGetLastError();
}

uint32_t closestWalkStack(const BvhView& bvh)
{
    if (bvh.tlas || bvh.wideRoot == STK_INVALID_INDEX) return 0;
    return bvh.wideDepth <= (uint32_t)kWideStackShallow + 1 ? (uint32_t)kWideStackShallow : (uint32_t)kWideStack;
}

hipError_t launchTraceAny(const BvhView& bvh, const StkRay* rays, uint32_t n, StkHit* hits, uint32_t flags,
                          hipStream_t s)
{
    if (n == 0) return hipSuccess;
    if (bvh.tlas) {
        if (flags & STK_TRACE_REFERENCE_ORDER) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_trace_any_tl, dim3(walkBlocks(n)), dim3(kWalkBlock), 0, s, bvh, rays, n, hits);
    } else if ((flags & STK_TRACE_REFERENCE_ORDER) || !bvh.packedOk)
        hipLaunchKernelGGL(k_trace_any_ref, dim3(blocks(n)), dim3(kBlock), 0, s, bvh, rays, n, hits);
    else
        hipLaunchKernelGGL(k_trace_any, dim3(walkBlocks(n)), dim3(kWalkBlock), 0, s, bvh, rays, n, hits);
    return hipGetLastError();
}

hipError_t launchCameraRays(const StkPathTracerParam& p, uint32_t sampleNum, StkRay* rays, hipStream_t s)
{
    const size_t n = (size_t)p.dimension[0] * (size_t)p.dimension[1];
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_camera_rays, dim3(blocks(n)), dim3(kBlock), 0, s, p, sampleNum, rays);
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
                                   const uint32_t* abortAny, hipStream_t s)
{
    if (npix == 0) return hipSuccess;
    hipLaunchKernelGGL(k_accumulate_tonemap, dim3(blocks(npix)), dim3(kBlock), 0, s, cur, history, out, npix,
                       iteration, accumulate, doTonemap, tonemapType, abortAny);
    return hipGetLastError();
}

} // namespace stk
