// gpubuild.hip - BVH construction on the GPU: a linear BVH (Karras 2012,
// "Maximizing Parallelism in the Construction of BVHs, Octrees, and k-d
// Trees") emitted in the layouts the walks read: the reference's skip-link
// preorder (include/render/bvh.h:17-24), the packed binary nodes and the
// compressed 8-wide nodes with the host builder's quantisation.
//
// Steps, each a kernel over triangles, nodes or one tree level:
//   1. world triangles as getTriangle forms them (shaders/raytracing.h:134-153)
//      and their boxes (max padded by 1e-6 as include/render/bvh.h:148-150);
//   2. centroid bounds (ordered-integer atomics), 63-bit Morton codes;
//   3. radix sort of (code, triangle) pairs (rocPRIM);
//   4. the binary radix tree from longest common prefixes (ties broken by
//      index), parents of every node;
//   5. boxes and leaf counts bottom-up (the second child to arrive at a node
//      finishes it; agent-scope acquire/release atomics order the writes);
//   4'-5'. or, for the clustered tree (kGpuBuildPloc), parallel locally-ordered
//      clustering (Meister and Bittner 2018, "Parallel Locally-Ordered
//      Clustering for Bounding Volume Hierarchy Construction"): the clusters
//      in Morton order each pick the neighbour within kPlocRadius positions
//      whose merged box has the smallest surface area, mutual pairs merge
//      into a new inner node, and the survivors are compacted in order, until
//      one cluster is left;
//   6. preorder indices top-down, one kernel per level: left child p + 1,
//      right child p + 2 x leaves(left); skip link = p + nodes of the subtree;
//   7. the skip-link, packed and triangle arrays;
//   8. the 8-wide collapse: a bottom-up dynamic program picks which binary
//      nodes become wide nodes (least summed wide-node area), then one wide
//      level at a time each job opens its descendants as planned, assigns
//      octant slots,
//      quantises outwards (checked against the float decode), then an
//      exclusive scan of child / leaf counts places the next level.
#include "gpubuild.hpp"

#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "geom.hpp"

namespace stk {
namespace {

constexpr uint32_t kB = 256;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kNone = 0xFFFFFFFFu;

inline uint32_t blocksFor(size_t n) { return (uint32_t)((n + kB - 1) / kB); }

__device__ __forceinline__ uint32_t ordered(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float unordered(uint32_t u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// 1. world triangles and boxes
__global__ __launch_bounds__(kB) void k_tri_prep(const StkVertex* __restrict__ vb, const uint32_t* __restrict__ ib,
                                                 const StkInstance* __restrict__ inst,
                                                 const uint32_t* __restrict__ instStart, uint32_t ninst, uint32_t n,
                                                 float4* __restrict__ boxLo, float4* __restrict__ boxHi,
                                                 StkTriangle* __restrict__ tris)
{
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k >= n) return;
    // the instance: instStart[i] <= k < instStart[i + 1] (instStart[ninst] = n)
    uint32_t lo = 0, hi = ninst;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (instStart[mid] <= k)
            lo = mid;
        else
            hi = mid;
    }
    const StkInstance& ic = inst[lo];
    const uint32_t prim = k - instStart[lo];
    const uint32_t base = (uint32_t)ic.indexOffset + 3 * prim;
    const F3 v0 = xform4(ic.objectToWorld, ld3(vb[ib[base]].position), 1.0f);
    const F3 v1 = xform4(ic.objectToWorld, ld3(vb[ib[base + 1]].position), 1.0f);
    const F3 v2 = xform4(ic.objectToWorld, ld3(vb[ib[base + 2]].position), 1.0f);
    boxLo[k] = make_float4(fminf(v0.x, fminf(v1.x, v2.x)), fminf(v0.y, fminf(v1.y, v2.y)),
                           fminf(v0.z, fminf(v1.z, v2.z)), 0.0f);
    boxHi[k] = make_float4(fmaxf(v0.x, fmaxf(v1.x, v2.x)) + 1e-6f, fmaxf(v0.y, fmaxf(v1.y, v2.y)) + 1e-6f,
                           fmaxf(v0.z, fmaxf(v1.z, v2.z)) + 1e-6f, 0.0f);
    const F3 e0 = v1 - v0, e1 = v2 - v0;
    StkTriangle t;
    t.v0[0] = v0.x;
    t.v0[1] = v0.y;
    t.v0[2] = v0.z;
    t.leaf = 0;
    t.e0[0] = e0.x;
    t.e0[1] = e0.y;
    t.e0[2] = e0.z;
    t.instId = lo;
    t.e1[0] = e1.x;
    t.e1[1] = e1.y;
    t.e1[2] = e1.z;
    t.primId = prim;
    tris[k] = t;
}

// 2. centroid bounds and Morton codes
__global__ void k_bounds_init(uint32_t* b)
{
    if (threadIdx.x < 3) b[threadIdx.x] = ordered(INFINITY);
    else if (threadIdx.x < 6) b[threadIdx.x] = ordered(-INFINITY);
}

__global__ __launch_bounds__(kB) void k_centroid_bounds(const float4* __restrict__ boxLo,
                                                        const float4* __restrict__ boxHi, uint32_t n,
                                                        uint32_t* __restrict__ bounds)
{
    __shared__ float red[6][kB];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB) {
        const float4 l = boxLo[k], h = boxHi[k];
        const float c[3] = {(l.x + h.x) * 0.5f, (l.y + h.y) * 0.5f, (l.z + h.z) * 0.5f};
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], c[a]);
            mx[a] = fmaxf(mx[a], c[a]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = mn[a];
        red[3 + a][threadIdx.x] = mx[a];
    }
    __syncthreads();
    for (uint32_t w = kB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int a = 0; a < 3; ++a) {
                red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + w]);
                red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; ++a) {
            atomicMin(&bounds[a], ordered(red[a][0]));
            atomicMax(&bounds[3 + a], ordered(red[3 + a][0]));
        }
}

__device__ __forceinline__ uint64_t spread21(uint32_t v)
{
    uint64_t x = v & 0x1FFFFFu;
    x = (x | x << 32) & 0x1F00000000FFFFull;
    x = (x | x << 16) & 0x1F0000FF0000FFull;
    x = (x | x << 8) & 0x100F00F00F00F00Full;
    x = (x | x << 4) & 0x10C30C30C30C30C3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ __launch_bounds__(kB) void k_morton(const float4* __restrict__ boxLo, const float4* __restrict__ boxHi,
                                               uint32_t n, const uint32_t* __restrict__ bounds,
                                               uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k >= n) return;
    const float4 l = boxLo[k], h = boxHi[k];
    const float c[3] = {(l.x + h.x) * 0.5f, (l.y + h.y) * 0.5f, (l.z + h.z) * 0.5f};
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
        const float lo = unordered(bounds[a]), hi = unordered(bounds[3 + a]);
        const float ext = hi - lo;
        const float u = ext > 0.0f ? (c[a] - lo) / ext * 2097152.0f : 0.0f;
        q[a] = (uint32_t)fminf(fmaxf(u, 0.0f), 2097151.0f);
    }
    keys[k] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    vals[k] = k;
}

// 4. binary radix tree
__device__ __forceinline__ int prefix(const uint64_t* K, int n, int a, int b)
{
    if (b < 0 || b > n - 1) return -1;
    const uint64_t x = K[a] ^ K[b];
    return x ? __clzll((long long)x) : 64 + __clz((int)((uint32_t)a ^ (uint32_t)b));
}

__global__ __launch_bounds__(kB) void k_radix_tree(const uint64_t* __restrict__ K, int n, uint32_t* __restrict__ childL,
                                                   uint32_t* __restrict__ childR, uint32_t* __restrict__ parentInner,
                                                   uint32_t* __restrict__ parentLeaf, uint32_t* __restrict__ count)
{
    const int i = (int)(blockIdx.x * kB + threadIdx.x);
    if (i >= n - 1) return;
    const int d = (prefix(K, n, i, i + 1) - prefix(K, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = prefix(K, n, i, i - d);
    int lmax = 2;
    while (prefix(K, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (prefix(K, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = prefix(K, n, i, j);
    int s = 0;
    for (int div = 2;; div <<= 1) {
        const int t = (l + div - 1) / div;
        if (prefix(K, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const uint32_t left = (min(i, j) == gamma) ? ((uint32_t)gamma | kLeafBit) : (uint32_t)gamma;
    const uint32_t right = (max(i, j) == gamma + 1) ? ((uint32_t)(gamma + 1) | kLeafBit) : (uint32_t)(gamma + 1);
    childL[i] = left;
    childR[i] = right;
    count[i] = (uint32_t)(max(i, j) - min(i, j) + 1); // node i covers the sorted leaves [min(i, j), max(i, j)]
    if (left & kLeafBit)
        parentLeaf[left & ~kLeafBit] = (uint32_t)i;
    else
        parentInner[left] = (uint32_t)i;
    if (right & kLeafBit)
        parentLeaf[right & ~kLeafBit] = (uint32_t)i;
    else
        parentInner[right] = (uint32_t)i;
    if (i == 0) parentInner[0] = kNone;
}

// 4'-5'. clustered tree.  Surface area without the factor 2 (only compared).
constexpr int kPlocRadius = 32;

__device__ __forceinline__ float half_area(float4 lo, float4 hi)
{
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return (dx * dy + dy * dz) + dz * dx;
}

__device__ __forceinline__ float4 box_min(float4 a, float4 b)
{
    return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.0f);
}

__device__ __forceinline__ float4 box_max(float4 a, float4 b)
{
    return make_float4(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z), 0.0f);
}

// clusters = the sorted leaves
__global__ __launch_bounds__(kB) void k_ploc_init(uint32_t n, const uint32_t* __restrict__ vals,
                                                  const float4* __restrict__ triLo, const float4* __restrict__ triHi,
                                                  uint32_t* __restrict__ C, float4* __restrict__ lo,
                                                  float4* __restrict__ hi, uint32_t* __restrict__ cnt)
{
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k >= n) return;
    const uint32_t t = vals[k];
    C[k] = k | kLeafBit;
    lo[k] = make_float4(triLo[t].x, triLo[t].y, triLo[t].z, 0.0f);
    hi[k] = make_float4(triHi[t].x, triHi[t].y, triHi[t].z, 0.0f);
    cnt[k] = 1;
}

// Order of candidate pairs {i, j} after the area: nearer positions first,
// then pairs whose lower position is even, then the lower position.  A
// strict total order on pairs, so the least pair of a round is mutual (every
// round merges), and equal boxes pair up as (0, 1), (2, 3), ...
__device__ __forceinline__ uint32_t pair_rank(uint32_t i, uint32_t j)
{
    const uint32_t lo = min(i, j);
    return (max(i, j) - lo) * 2u + (lo & 1u);
}

// nearest neighbour of each cluster within the radius
__global__ __launch_bounds__(kB) void k_ploc_nn(uint32_t m, const float4* __restrict__ lo,
                                                const float4* __restrict__ hi, uint32_t* __restrict__ nn)
{
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    const float4 li = lo[i], hi_ = hi[i];
    const uint32_t j0 = i > (uint32_t)kPlocRadius ? i - kPlocRadius : 0u;
    const uint32_t j1 = min(m - 1, i + (uint32_t)kPlocRadius);
    float best = INFINITY;
    uint32_t bj = i == 0 ? 1u : i - 1, bk = 0xFFFFFFFFu;
    for (uint32_t j = j0; j <= j1; ++j) {
        if (j == i) continue;
        const float a = half_area(box_min(li, lo[j]), box_max(hi_, hi[j]));
        const uint32_t k = pair_rank(i, j);
        // (for equal area and rank the lower position is the smaller j: j
        // rises through the window)
        if (a < best || (a == best && k < bk)) {
            best = a;
            bk = k;
            bj = j;
        }
    }
    nn[i] = bj;
}

// merge[i]: i merges with its neighbour into a new node (mutual, i the
// smaller); keep[i]: position i survives (all but the larger of a pair)
__global__ __launch_bounds__(kB) void k_ploc_flags(uint32_t m, const uint32_t* __restrict__ nn,
                                                   uint32_t* __restrict__ merge, uint32_t* __restrict__ keep)
{
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    const uint32_t j = nn[i];
    const bool mutual = nn[j] == i;
    merge[i] = (mutual && i < j) ? 1u : 0u;
    keep[i] = (mutual && i > j) ? 0u : 1u;
}

// writes the next round's clusters and the new inner nodes; the k-th merge of
// the build gets inner index ninner - 1 - k, so the last one is the root, 0
__global__ __launch_bounds__(kB) void k_ploc_write(uint32_t m, uint32_t base, uint32_t ninner,
                                                   const uint32_t* __restrict__ C, const float4* __restrict__ lo,
                                                   const float4* __restrict__ hi, const uint32_t* __restrict__ cnt,
                                                   const uint32_t* __restrict__ nn, const uint32_t* __restrict__ merge,
                                                   const uint32_t* __restrict__ keep,
                                                   const uint32_t* __restrict__ mrank, const uint32_t* __restrict__ pos,
                                                   uint32_t* __restrict__ C2, float4* __restrict__ lo2,
                                                   float4* __restrict__ hi2, uint32_t* __restrict__ cnt2,
                                                   uint32_t* __restrict__ childL, uint32_t* __restrict__ childR,
                                                   uint32_t* __restrict__ count, float4* __restrict__ nodeLo,
                                                   float4* __restrict__ nodeHi, uint32_t* __restrict__ mOut)
{
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    if (i == m - 1) *mOut = pos[i] + keep[i];
    if (!keep[i]) return;
    const uint32_t p = pos[i];
    if (merge[i]) {
        const uint32_t j = nn[i];
        const uint32_t id = ninner - 1 - (base + mrank[i]);
        const float4 l = box_min(lo[i], lo[j]), h = box_max(hi[i], hi[j]);
        const uint32_t c = cnt[i] + cnt[j];
        childL[id] = C[i];
        childR[id] = C[j];
        count[id] = c;
        nodeLo[id] = l;
        nodeHi[id] = h;
        C2[p] = id;
        lo2[p] = l;
        hi2[p] = h;
        cnt2[p] = c;
    } else {
        C2[p] = C[i];
        lo2[p] = lo[i];
        hi2[p] = hi[i];
        cnt2[p] = cnt[i];
    }
}

// 5. boxes and leaf counts bottom-up
__global__ __launch_bounds__(kB) void k_refit(uint32_t n, const uint32_t* __restrict__ vals,
                                              const float4* __restrict__ triLo, const float4* __restrict__ triHi,
                                              const uint32_t* __restrict__ childL, const uint32_t* __restrict__ childR,
                                              const uint32_t* __restrict__ parentInner,
                                              const uint32_t* __restrict__ parentLeaf, float4* nodeLo, float4* nodeHi,
                                              uint32_t* flags)
{
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k >= n) return;
    uint32_t p = parentLeaf[k];
    while (p != kNone) {
        // The second child to arrive finishes p.  Full agent-scope fences on
        // both sides of the arrival: this lane's box stores have completed
        // and are written back before it, and the sibling's are read after
        // it, across the XCDs' caches.
        __threadfence();
        const uint32_t old = __hip_atomic_fetch_add(&flags[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) return;
        __threadfence();
        float4 lo[2], hi[2];
        const uint32_t kids[2] = {childL[p], childR[p]};
        for (int c = 0; c < 2; ++c) {
            if (kids[c] & kLeafBit) {
                const uint32_t t = vals[kids[c] & ~kLeafBit];
                lo[c] = triLo[t];
                hi[c] = triHi[t];
            } else {
                lo[c] = nodeLo[kids[c]];
                hi[c] = nodeHi[kids[c]];
            }
        }
        nodeLo[p] = make_float4(fminf(lo[0].x, lo[1].x), fminf(lo[0].y, lo[1].y), fminf(lo[0].z, lo[1].z), 0.0f);
        nodeHi[p] = make_float4(fmaxf(hi[0].x, hi[1].x), fmaxf(hi[0].y, hi[1].y), fmaxf(hi[0].z, hi[1].z), 0.0f);
        p = parentInner[p];
    }
}

// Checks the finished inner nodes: finite boxes holding their children's, and
// leaf counts adding up; any failure sets *error (the build then fails
// instead of emitting a tree the walks would mistrust).
__global__ __launch_bounds__(kB) void k_check_inner(uint32_t ninner, const uint32_t* __restrict__ vals,
                                                    const float4* __restrict__ triLo, const float4* __restrict__ triHi,
                                                    const uint32_t* __restrict__ childL,
                                                    const uint32_t* __restrict__ childR,
                                                    const float4* __restrict__ nodeLo, const float4* __restrict__ nodeHi,
                                                    const uint32_t* __restrict__ count, uint32_t* error)
{
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= ninner) return;
    const float4 lo = nodeLo[i], hi = nodeHi[i];
    bool ok = isfinite(lo.x) && isfinite(lo.y) && isfinite(lo.z) && isfinite(hi.x) && isfinite(hi.y) &&
              isfinite(hi.z) && lo.x <= hi.x && lo.y <= hi.y && lo.z <= hi.z;
    uint32_t cnt = 0;
    const uint32_t kids[2] = {childL[i], childR[i]};
    for (int c = 0; c < 2; ++c) {
        float4 l, h;
        if (kids[c] & kLeafBit) {
            const uint32_t t = vals[kids[c] & ~kLeafBit];
            l = triLo[t];
            h = triHi[t];
            cnt += 1;
        } else {
            l = nodeLo[kids[c]];
            h = nodeHi[kids[c]];
            cnt += count[kids[c]];
        }
        ok = ok && l.x >= lo.x && l.y >= lo.y && l.z >= lo.z && h.x <= hi.x && h.y <= hi.y && h.z <= hi.z;
    }
    if (!ok || cnt != count[i]) atomicOr(error, 2u);
}

// 6. preorder, one level of inner nodes per launch
__global__ __launch_bounds__(kB) void k_preorder_level(const uint32_t* __restrict__ frontier, uint32_t nf,
                                                       const uint32_t* __restrict__ childL,
                                                       const uint32_t* __restrict__ childR,
                                                       const uint32_t* __restrict__ count, uint32_t* pre,
                                                       uint32_t* preLeaf, uint32_t* next, uint32_t* nextCount)
{
    const uint32_t f = blockIdx.x * kB + threadIdx.x;
    if (f >= nf) return;
    const uint32_t i = frontier[f];
    const uint32_t p = pre[i];
    const uint32_t L = childL[i], R = childR[i];
    const uint32_t leavesL = (L & kLeafBit) ? 1u : count[L];
    const uint32_t pl = p + 1, pr = p + 2 * leavesL;
    if (L & kLeafBit) {
        preLeaf[L & ~kLeafBit] = pl;
    } else {
        pre[L] = pl;
        next[atomicAdd(nextCount, 1u)] = L;
    }
    if (R & kLeafBit) {
        preLeaf[R & ~kLeafBit] = pr;
    } else {
        pre[R] = pr;
        next[atomicAdd(nextCount, 1u)] = R;
    }
}

// 7. skip-link (device form), packed and triangle arrays
__global__ __launch_bounds__(kB) void k_emit_inner(uint32_t n, const uint32_t* __restrict__ pre,
                                                   const uint32_t* __restrict__ count,
                                                   const float4* __restrict__ nodeLo, const float4* __restrict__ nodeHi,
                                                   const uint32_t* __restrict__ childL,
                                                   const uint32_t* __restrict__ childR, StkBvhNode* __restrict__ ref,
                                                   StkPackedNode* __restrict__ packed)
{
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n - 1) return;
    const uint32_t total = 2 * n - 1;
    const uint32_t p = pre[i];
    const uint32_t skip = p + 2 * count[i] - 1;
    const float4 lo = nodeLo[i], hi = nodeHi[i];
    StkBvhNode nd;
    nd.minBounds[0] = lo.x;
    nd.minBounds[1] = lo.y;
    nd.minBounds[2] = lo.z;
    nd.instId = -1;
    nd.maxBounds[0] = hi.x;
    nd.maxBounds[1] = hi.y;
    nd.maxBounds[2] = hi.z;
    nd.nodeOffset = skip == total ? -1 : (int32_t)skip;
    ref[p] = nd;
    StkPackedNode pk;
    for (int a = 0; a < 4; ++a) pk.x[a] = pk.y[a] = pk.z[a] = 0.0f;
    pk.pad[0] = pk.pad[1] = 0;
    const uint32_t kids[2] = {childL[i], childR[i]};
    for (int c = 0; c < 2; ++c) {
        if (kids[c] & kLeafBit) {
            pk.child[c] = ~(int32_t)(kids[c] & ~kLeafBit);
        } else {
            const float4 l = nodeLo[kids[c]], h = nodeHi[kids[c]];
            pk.child[c] = (int32_t)kids[c];
            pk.x[2 * c] = l.x;
            pk.x[2 * c + 1] = h.x;
            pk.y[2 * c] = l.y;
            pk.y[2 * c + 1] = h.y;
            pk.z[2 * c] = l.z;
            pk.z[2 * c + 1] = h.z;
        }
    }
    packed[i] = pk;
}

__global__ __launch_bounds__(kB) void k_emit_leaves(uint32_t n, const uint32_t* __restrict__ vals,
                                                    const uint32_t* __restrict__ preLeaf,
                                                    const StkTriangle* __restrict__ trisIn,
                                                    const StkInstance* __restrict__ inst, StkTriangle* __restrict__ trisOut,
                                                    StkBvhNode* __restrict__ ref, StkTriangle* __restrict__ emit,
                                                    uint32_t* emitCount)
{
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k >= n) return;
    const uint32_t total = 2 * n - 1;
    const uint32_t p = preLeaf[k];
    StkTriangle t = trisIn[vals[k]];
    t.leaf = p;
    trisOut[k] = t;
    StkBvhNode nd;
    nd.minBounds[0] = __uint_as_float(t.primId);
    nd.minBounds[1] = 0.0f;
    nd.minBounds[2] = 0.0f;
    nd.instId = (int32_t)t.instId;
    nd.maxBounds[0] = __uint_as_float(k); // device form: the triangle's index
    nd.maxBounds[1] = 0.0f;
    nd.maxBounds[2] = 0.0f;
    nd.nodeOffset = p + 1 == total ? -1 : (int32_t)(p + 1);
    ref[p] = nd;
    if (inst[t.instId].lightId != -1) emit[atomicAdd(emitCount, 1u)] = t;
}

// 8. 8-wide collapse, one wide level per pair of launches
struct WideChildRef {
    uint32_t ref; // inner node index, or triangle index | kLeafBit
    float lo[3], hi[3];
};

__device__ void child_box(uint32_t ref, const float4* nodeLo, const float4* nodeHi, const StkTriangle* tris,
                          WideChildRef& c)
{
    c.ref = ref;
    if (!(ref & kLeafBit)) {
        const float4 l = nodeLo[ref], h = nodeHi[ref];
        c.lo[0] = l.x;
        c.lo[1] = l.y;
        c.lo[2] = l.z;
        c.hi[0] = h.x;
        c.hi[1] = h.y;
        c.hi[2] = h.z;
        return;
    }
    // a leaf's box: its triangle's box (the vertices as the hit test forms
    // them), widened by 1e-5 of the larger of its extent and magnitude
    const StkTriangle& t = tris[ref & ~kLeafBit];
    for (int a = 0; a < 3; ++a) {
        const float v0 = t.v0[a], v1 = t.v0[a] + t.e0[a], v2 = t.v0[a] + t.e1[a];
        const float l = fminf(v0, fminf(v1, v2)), h = fmaxf(v0, fmaxf(v1, v2));
        const float mag = fmaxf(fabsf(l), fabsf(h));
        const float pad = 1e-5f * fmaxf(h - l, mag) + 1e-30f;
        c.lo[a] = l - pad;
        c.hi[a] = h + pad;
    }
}

// Smallest exponent byte e with 2^(e-127) * 253 >= extent, at least 2^-100
// and no finer than a quarter ulp of the origin's magnitude.
__device__ uint32_t choose_exponent(float origin, float extent)
{
    const double need = fmax((double)extent / 253.0, ldexp(1.0, -100));
    const double ulp = fabs((double)origin) * ldexp(1.0, -25);
    const double target = fmax(need, ulp);
    if (!(target < 1e300)) return 254; // non-finite boxes are rejected by the inner-node check
    int e = (int)ceil(log2(target)) + 127;
    while (e > 1 && ldexp(1.0, e - 1 - 127) * 253.0 >= (double)extent && ldexp(1.0, e - 1 - 127) >= target) --e;
    while (ldexp(1.0, e - 127) < target) ++e;
    return (uint32_t)min(254, max(1, e));
}

// Parent links from child links (the clustered build writes only the latter).
__global__ __launch_bounds__(kB) void k_parents(uint32_t ninner, const uint32_t* __restrict__ childL,
                                                const uint32_t* __restrict__ childR,
                                                uint32_t* __restrict__ parentInner, uint32_t* __restrict__ parentLeaf)
{
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= ninner) return;
    const uint32_t kids[2] = {childL[i], childR[i]};
    for (int c = 0; c < 2; ++c) {
        if (kids[c] & kLeafBit)
            parentLeaf[kids[c] & ~kLeafBit] = i;
        else
            parentInner[kids[c]] = i;
    }
    if (i == 0) parentInner[0] = kNone;
}

// Which binary nodes become wide nodes: the collapse of least summed wide-node
// surface area, by dynamic programming bottom-up (Ylitie, Karras and Laine
// 2017, "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide
// BVHs"), the plan the host collapse follows, in single precision.  For inner
// node p and a = 1..8 slots of its parent: F(p, a) = least cost of p's subtree
// in at most a slots (a leaf costs 0), where p either stays one wide node
// (area(p) + the best split of 8 slots between its children) or is opened
// (a1 + a2 = a slots to its children).  openSplit(p, a) = a1 of the opened
// choice, 0 when kept; innerSplit(p) = the children's split inside a wide node
// at p.  The second child to arrive finishes its parent, as in k_refit.
__global__ __launch_bounds__(kB) void k_wide_cost(uint32_t n, const uint32_t* __restrict__ childL,
                                                  const uint32_t* __restrict__ childR,
                                                  const uint32_t* __restrict__ parentInner,
                                                  const uint32_t* __restrict__ parentLeaf,
                                                  const float4* __restrict__ nodeLo, const float4* __restrict__ nodeHi,
                                                  float* F, uint8_t* __restrict__ openSplit,
                                                  uint8_t* __restrict__ innerSplit, uint32_t* flags)
{
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k >= n) return;
    uint32_t p = parentLeaf[k];
    while (p != kNone) {
        __threadfence();
        const uint32_t old = __hip_atomic_fetch_add(&flags[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) return;
        __threadfence();
        const uint32_t L = childL[p], R = childR[p];
        float fl[8], fr[8];
        for (int a = 0; a < 8; ++a) {
            fl[a] = (L & kLeafBit) ? 0.0f : F[(size_t)L * 8 + a];
            fr[a] = (R & kLeafBit) ? 0.0f : F[(size_t)R * 8 + a];
        }
        float g = fl[0] + fr[6];
        int gs = 1; // a1 in 1..7 whatever the sums (non-finite ones compare false)
        for (int a1 = 2; a1 <= 7; ++a1) {
            const float c = fl[a1 - 1] + fr[7 - a1];
            if (c < g) {
                g = c;
                gs = a1;
            }
        }
        innerSplit[p] = (uint8_t)gs;
        const float keep = half_area(nodeLo[p], nodeHi[p]) + g;
        for (int a = 1; a <= 8; ++a) {
            float best = keep;
            int split = 0;
            for (int a1 = 1; a1 < a; ++a1) {
                const float c = fl[a1 - 1] + fr[a - a1 - 1];
                if (c < best) {
                    best = c;
                    split = a1;
                }
            }
            F[(size_t)p * 8 + (a - 1)] = best;
            openSplit[(size_t)p * 8 + (a - 1)] = (uint8_t)split;
        }
        p = parentInner[p];
    }
}

__global__ __launch_bounds__(64) void k_wide_plan(const uint32_t* __restrict__ jobs, uint32_t nj,
                                                  const uint32_t* __restrict__ childL,
                                                  const uint32_t* __restrict__ childR,
                                                  const uint8_t* __restrict__ openSplit,
                                                  const uint8_t* __restrict__ innerSplit,
                                                  const float4* __restrict__ nodeLo, const float4* __restrict__ nodeHi,
                                                  const StkTriangle* __restrict__ tris, StkWideNode* __restrict__ plan,
                                                  uint32_t* __restrict__ kidRefs, uint32_t* __restrict__ innerCnt,
                                                  uint32_t* __restrict__ leafCnt, uint32_t* __restrict__ error)
{
    const uint32_t j = blockIdx.x * 64 + threadIdx.x;
    if (j >= nj) return;
    const uint32_t job = jobs[j];
    WideChildRef kids[8];
    int nk = 0;
    if (job & kLeafBit) {
        child_box(job, nodeLo, nodeHi, tris, kids[nk++]); // the whole tree is one leaf
    } else {
        // the children in preorder as the plan opens them: every stacked or
        // taken entry holds at least one of the 8 slots, so neither exceeds 8
        uint32_t stRef[8];
        int stSlots[8];
        int sp = 0;
        const int is = min(7, max(1, (int)innerSplit[job]));
        stRef[sp] = childR[job];
        stSlots[sp++] = 8 - is;
        stRef[sp] = childL[job];
        stSlots[sp++] = is;
        while (sp > 0) {
            --sp;
            const uint32_t r = stRef[sp];
            const int a = stSlots[sp];
            const int split = ((r & kLeafBit) || a <= 1) ? 0 : openSplit[(size_t)r * 8 + (a - 1)];
            if (split <= 0 || split >= a) {
                child_box(r, nodeLo, nodeHi, tris, kids[nk++]);
                continue;
            }
            stRef[sp] = childR[r];
            stSlots[sp++] = a - split;
            stRef[sp] = childL[r];
            stSlots[sp++] = split;
        }
    }
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = kids[0].lo[a];
        hi[a] = kids[0].hi[a];
        for (int k = 1; k < nk; ++k) {
            lo[a] = fminf(lo[a], kids[k].lo[a]);
            hi[a] = fmaxf(hi[a], kids[k].hi[a]);
        }
    }
    // slots by the child's centre relative to the node's (greedy, as the host builder)
    int slotOf[8];
    uint32_t slotUsed = 0, placed = 0;
    for (int round = 0; round < nk; ++round) {
        double bestScore = -1e300;
        int bk = -1, bs = -1;
        for (int k = 0; k < nk; ++k) {
            if ((placed >> k) & 1u) continue;
            double d[3];
            for (int a = 0; a < 3; ++a)
                d[a] = 0.5 * ((double)kids[k].lo[a] + kids[k].hi[a]) - 0.5 * ((double)lo[a] + hi[a]);
            for (int sl = 0; sl < 8; ++sl) {
                if ((slotUsed >> sl) & 1u) continue;
                double score = 0.0;
                for (int a = 0; a < 3; ++a) score += ((sl >> a) & 1 ? d[a] : -d[a]);
                if (score > bestScore || bk < 0) { // the first free pair stands in for non-finite scores
                    bestScore = score;
                    bk = k;
                    bs = sl;
                }
            }
        }
        placed |= 1u << bk;
        slotUsed |= 1u << bs;
        slotOf[bk] = bs;
    }
    int kidAt[8];
    for (int sl = 0; sl < 8; ++sl) kidAt[sl] = -1;
    for (int k = 0; k < nk; ++k) kidAt[slotOf[k]] = k;
    StkWideNode w;
    uint8_t* wb = reinterpret_cast<uint8_t*>(&w);
    for (int b = 0; b < (int)sizeof(w); ++b) wb[b] = 0;
    float scale[3];
    for (int a = 0; a < 3; ++a) {
        w.origin[a] = lo[a];
        w.exps[a] = (uint8_t)choose_exponent(lo[a], hi[a] - lo[a]);
        scale[a] = __uint_as_float((uint32_t)w.exps[a] << 23);
    }
    uint32_t inner = 0, leaves = 0;
    for (int sl = 0; sl < 8; ++sl) {
        kidRefs[(size_t)j * 8 + sl] = kNone;
        if (kidAt[sl] < 0) continue;
        const WideChildRef& c = kids[kidAt[sl]];
        w.vmask |= 1u << sl;
        for (int a = 0; a < 3; ++a) {
            // outward rounding checked against the decode the traversal performs
            const double ql = floor(((double)c.lo[a] - (double)w.origin[a]) / (double)scale[a]);
            const double qh = ceil(((double)c.hi[a] - (double)w.origin[a]) / (double)scale[a]);
            uint32_t qlo = (uint32_t)fmin(255.0, fmax(0.0, ql));
            uint32_t qhi = (uint32_t)fmin(255.0, fmax(0.0, qh));
            while (qlo > 0 && fmaf((float)qlo, scale[a], w.origin[a]) > c.lo[a]) --qlo;
            while (qhi < 255 && fmaf((float)qhi, scale[a], w.origin[a]) < c.hi[a]) ++qhi;
            if (fmaf((float)qlo, scale[a], w.origin[a]) > c.lo[a] || fmaf((float)qhi, scale[a], w.origin[a]) < c.hi[a])
                atomicOr(error, 1u);
            w.qlo[a][sl] = (uint8_t)qlo;
            w.qhi[a][sl] = (uint8_t)qhi;
        }
        kidRefs[(size_t)j * 8 + sl] = c.ref;
        if (c.ref & kLeafBit) {
            ++leaves;
        } else {
            w.imask |= (uint8_t)(1u << sl);
            ++inner;
        }
    }
    plan[j] = w;
    innerCnt[j] = inner;
    leafCnt[j] = leaves;
}

__global__ __launch_bounds__(kB) void k_wide_emit(uint32_t nj, uint32_t levelStart, uint32_t triStart,
                                                  const StkWideNode* __restrict__ plan,
                                                  const uint32_t* __restrict__ kidRefs,
                                                  const uint32_t* __restrict__ innerOff,
                                                  const uint32_t* __restrict__ leafOff,
                                                  const StkTriangle* __restrict__ tris, StkWideNode* __restrict__ wide,
                                                  StkTriangle* __restrict__ wideTris, uint32_t* __restrict__ nextJobs)
{
    const uint32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= nj) return;
    StkWideNode w = plan[j];
    w.childBase = levelStart + nj + innerOff[j];
    w.triBase = triStart + leafOff[j];
    wide[levelStart + j] = w;
    uint32_t r = 0, r2 = 0;
    for (int sl = 0; sl < 8; ++sl) {
        const uint32_t ref = kidRefs[(size_t)j * 8 + sl];
        if (ref == kNone) continue;
        if (ref & kLeafBit)
            wideTris[w.triBase + r2++] = tris[ref & ~kLeafBit];
        else
            nextJobs[innerOff[j] + r++] = ref;
    }
}

__global__ void k_set_u32(uint32_t* p, uint32_t v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB) p[i] = v;
}

} // namespace

hipError_t gpuBuildBvh(const StkVertex* vb, const uint32_t* ib, const StkInstance* inst, const uint32_t* instStart,
                       uint32_t ninst, uint32_t ntri, uint32_t method,
                       const std::function<void*(int, size_t)>& alloc, GpuBvhResult& out, std::string& err,
                       hipStream_t s)
{
    out = GpuBvhResult{};
    const uint32_t n = ntri;
    if (n == 0) {
        out.header.root = INT32_MIN;
        return hipSuccess;
    }
    if (n >= (1u << 30)) {
        err = "too many triangles for the GPU builder";
        return hipErrorInvalidValue;
    }
    int slot = kGbFirstTemp;
    auto temp = [&](size_t bytes) { return alloc(slot++, bytes); };
#define STK_GB_ALLOC(T, name, count, where)                                                                          \
    T* name = static_cast<T*>(where);                                                                                 \
    if (!name && (count) > 0) {                                                                                       \
        err = "device allocation failed (" #name ")";                                                                 \
        return hipErrorOutOfMemory;                                                                                   \
    }
    STK_GB_ALLOC(uint32_t, dInstStart, ninst + 1, temp((size_t)(ninst + 1) * 4))
    STK_GB_ALLOC(float4, triLo, n, temp((size_t)n * 16))
    STK_GB_ALLOC(float4, triHi, n, temp((size_t)n * 16))
    STK_GB_ALLOC(StkTriangle, trisIn, n, temp((size_t)n * sizeof(StkTriangle)))
    STK_GB_ALLOC(uint32_t, bounds, 8, temp(64))
    STK_GB_ALLOC(uint64_t, keysA, n, temp((size_t)n * 8))
    STK_GB_ALLOC(uint64_t, keysB, n, temp((size_t)n * 8))
    STK_GB_ALLOC(uint32_t, valsA, n, temp((size_t)n * 4))
    STK_GB_ALLOC(uint32_t, valsB, n, temp((size_t)n * 4))
    hipError_t e;
    if ((e = hipMemcpyAsync(dInstStart, instStart, (size_t)(ninst + 1) * 4, hipMemcpyHostToDevice, s)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_tri_prep, dim3(blocksFor(n)), dim3(kB), 0, s, vb, ib, inst, dInstStart, ninst, n, triLo,
                       triHi, trisIn);
    hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, s, bounds);
    hipLaunchKernelGGL(k_centroid_bounds, dim3(std::min<uint32_t>(blocksFor(n), 2048)), dim3(kB), 0, s, triLo, triHi,
                       n, bounds);
    hipLaunchKernelGGL(k_morton, dim3(blocksFor(n)), dim3(kB), 0, s, triLo, triHi, n, bounds, keysA, valsA);
    size_t sortBytes = 0;
    if ((e = rocprim::radix_sort_pairs(nullptr, sortBytes, keysA, keysB, valsA, valsB, n, 0, 63, s)) != hipSuccess)
        return e;
    STK_GB_ALLOC(void, sortTemp, sortBytes, temp(sortBytes))
    if ((e = rocprim::radix_sort_pairs(sortTemp, sortBytes, keysA, keysB, valsA, valsB, n, 0, 63, s)) != hipSuccess)
        return e;
    const uint64_t* keys = keysB;
    const uint32_t* vals = valsB;

    const uint32_t ninner = n - 1;
    STK_GB_ALLOC(StkTriangle, tris, n, alloc(kGbTris, (size_t)n * sizeof(StkTriangle)))
    STK_GB_ALLOC(StkBvhNode, ref, 2 * (size_t)n - 1, alloc(kGbRefNodes, (2 * (size_t)n - 1) * sizeof(StkBvhNode)))
    STK_GB_ALLOC(StkTriangle, emit, n, alloc(kGbEmitters, (size_t)n * sizeof(StkTriangle)))
    STK_GB_ALLOC(uint32_t, childL, ninner, temp((size_t)std::max(ninner, 1u) * 4))
    STK_GB_ALLOC(uint32_t, childR, ninner, temp((size_t)std::max(ninner, 1u) * 4))
    STK_GB_ALLOC(uint32_t, parentInner, ninner, temp((size_t)std::max(ninner, 1u) * 4))
    STK_GB_ALLOC(uint32_t, parentLeaf, n, temp((size_t)n * 4))
    STK_GB_ALLOC(float4, nodeLo, ninner, temp((size_t)std::max(ninner, 1u) * 16))
    STK_GB_ALLOC(float4, nodeHi, ninner, temp((size_t)std::max(ninner, 1u) * 16))
    STK_GB_ALLOC(uint32_t, count, ninner, temp((size_t)std::max(ninner, 1u) * 4))
    STK_GB_ALLOC(uint32_t, flags, ninner, temp((size_t)std::max(ninner, 1u) * 4))
    STK_GB_ALLOC(uint32_t, pre, ninner, temp((size_t)std::max(ninner, 1u) * 4))
    STK_GB_ALLOC(uint32_t, preLeaf, n, temp((size_t)n * 4))
    STK_GB_ALLOC(uint32_t, counters, 4, temp(64))

    uint32_t depth = 0;
    if (n == 1) {
        const uint32_t zero = 0;
        if ((e = hipMemcpyAsync(preLeaf, &zero, 4, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    } else {
        if (method == kGpuBuildPloc) {
            // clusters, double-buffered, and the per-round arrays
            STK_GB_ALLOC(uint32_t, cA, n, temp((size_t)n * 4))
            STK_GB_ALLOC(uint32_t, cB, n, temp((size_t)n * 4))
            STK_GB_ALLOC(float4, loA, n, temp((size_t)n * 16))
            STK_GB_ALLOC(float4, loB, n, temp((size_t)n * 16))
            STK_GB_ALLOC(float4, hiA, n, temp((size_t)n * 16))
            STK_GB_ALLOC(float4, hiB, n, temp((size_t)n * 16))
            STK_GB_ALLOC(uint32_t, nA, n, temp((size_t)n * 4))
            STK_GB_ALLOC(uint32_t, nB, n, temp((size_t)n * 4))
            STK_GB_ALLOC(uint32_t, nn, n, temp((size_t)n * 4))
            STK_GB_ALLOC(uint32_t, mergeF, n, temp((size_t)n * 4))
            STK_GB_ALLOC(uint32_t, keepF, n, temp((size_t)n * 4))
            STK_GB_ALLOC(uint32_t, mrank, n, temp((size_t)n * 4))
            STK_GB_ALLOC(uint32_t, pos, n, temp((size_t)n * 4))
            size_t pBytes = 0;
            if ((e = rocprim::exclusive_scan(nullptr, pBytes, mergeF, mrank, 0u, n, rocprim::plus<uint32_t>(), s)) !=
                hipSuccess)
                return e;
            STK_GB_ALLOC(void, pTemp, pBytes, temp(pBytes))
            hipLaunchKernelGGL(k_ploc_init, dim3(blocksFor(n)), dim3(kB), 0, s, n, vals, triLo, triHi, cA, loA, hiA,
                               nA);
            uint32_t m = n, base = 0, rounds = 0;
            while (m > 1) {
                if (++rounds > 4096) {
                    (void)hipStreamSynchronize(s);
                    err = "clustering did not converge (non-finite triangle boxes?)";
                    return hipErrorInvalidValue;
                }
                hipLaunchKernelGGL(k_ploc_nn, dim3(blocksFor(m)), dim3(kB), 0, s, m, loA, hiA, nn);
                hipLaunchKernelGGL(k_ploc_flags, dim3(blocksFor(m)), dim3(kB), 0, s, m, nn, mergeF, keepF);
                if ((e = rocprim::exclusive_scan(pTemp, pBytes, mergeF, mrank, 0u, m, rocprim::plus<uint32_t>(), s)) !=
                    hipSuccess)
                    return e;
                if ((e = rocprim::exclusive_scan(pTemp, pBytes, keepF, pos, 0u, m, rocprim::plus<uint32_t>(), s)) !=
                    hipSuccess)
                    return e;
                hipLaunchKernelGGL(k_ploc_write, dim3(blocksFor(m)), dim3(kB), 0, s, m, base, ninner, cA, loA, hiA, nA,
                                   nn, mergeF, keepF, mrank, pos, cB, loB, hiB, nB, childL, childR, count, nodeLo,
                                   nodeHi, counters + 3);
                uint32_t m2 = 0;
                if ((e = hipMemcpyAsync(&m2, counters + 3, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) {
                    (void)hipStreamSynchronize(s);
                    return e;
                }
                if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
                if (m2 >= m || m2 == 0) {
                    err = "clustering made no progress";
                    return hipErrorInvalidValue;
                }
                base += m - m2;
                m = m2;
                std::swap(cA, cB);
                std::swap(loA, loB);
                std::swap(hiA, hiB);
                std::swap(nA, nB);
            }
            out.rounds = rounds;
        } else {
            hipLaunchKernelGGL(k_set_u32, dim3(std::min<uint32_t>(blocksFor(ninner), 1024)), dim3(kB), 0, s,
                               parentLeaf, kNone, (size_t)n);
            hipLaunchKernelGGL(k_radix_tree, dim3(blocksFor(ninner)), dim3(kB), 0, s, keys, (int)n, childL, childR,
                               parentInner, parentLeaf, count);
            hipLaunchKernelGGL(k_set_u32, dim3(std::min<uint32_t>(blocksFor(ninner), 1024)), dim3(kB), 0, s, flags, 0u,
                               (size_t)ninner);
            hipLaunchKernelGGL(k_refit, dim3(blocksFor(n)), dim3(kB), 0, s, n, vals, triLo, triHi, childL, childR,
                               parentInner, parentLeaf, nodeLo, nodeHi, flags);
        }
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(kB), 0, s, counters + 2, 0u, (size_t)1);
        hipLaunchKernelGGL(k_check_inner, dim3(blocksFor(ninner)), dim3(kB), 0, s, ninner, vals, triLo, triHi, childL,
                           childR, nodeLo, nodeHi, count, counters + 2);
        uint32_t checked = 0;
        if ((e = hipMemcpyAsync(&checked, counters + 2, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
        if (checked != 0) {
            err = "inner node boxes or leaf counts are inconsistent";
            return hipErrorInvalidValue;
        }
        // preorder, level by level from the root (pre[0] = 0)
        STK_GB_ALLOC(uint32_t, frontA, ninner, temp((size_t)ninner * 4))
        STK_GB_ALLOC(uint32_t, frontB, ninner, temp((size_t)ninner * 4))
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(kB), 0, s, pre, 0u, (size_t)1);
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(kB), 0, s, frontA, 0u, (size_t)1);
        uint32_t nf = 1;
        uint32_t* cur = frontA;
        uint32_t* nxt = frontB;
        while (nf > 0) {
            hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(kB), 0, s, counters, 0u, (size_t)1);
            hipLaunchKernelGGL(k_preorder_level, dim3(blocksFor(nf)), dim3(kB), 0, s, cur, nf, childL, childR, count,
                               pre, preLeaf, nxt, counters);
            ++depth;
            if ((e = hipMemcpyAsync(&nf, counters, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
            if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
            std::swap(cur, nxt);
        }
    }
    STK_GB_ALLOC(StkPackedNode, packed, ninner, alloc(kGbPacked, (size_t)std::max(ninner, 1u) * sizeof(StkPackedNode)))
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(kB), 0, s, counters + 1, 0u, (size_t)1);
    if (n > 1)
        hipLaunchKernelGGL(k_emit_inner, dim3(blocksFor(ninner)), dim3(kB), 0, s, n, pre, count, nodeLo, nodeHi,
                           childL, childR, ref, packed);
    hipLaunchKernelGGL(k_emit_leaves, dim3(blocksFor(n)), dim3(kB), 0, s, n, vals, preLeaf, trisIn, inst, tris, ref,
                       emit, counters + 1);

    // the 8-wide collapse, one level at a time, following the plan of least
    // summed wide-node area
    const size_t wideCap = std::max<size_t>(ninner, 1);
    STK_GB_ALLOC(float, planCost, wideCap * 8, temp(wideCap * 32))
    STK_GB_ALLOC(uint8_t, openSplit, wideCap * 8, temp(wideCap * 8))
    STK_GB_ALLOC(uint8_t, innerSplit, wideCap, temp(wideCap))
    if (n > 1) {
        if (method == kGpuBuildPloc) {
            hipLaunchKernelGGL(k_set_u32, dim3(std::min<uint32_t>(blocksFor(n), 1024)), dim3(kB), 0, s, parentLeaf,
                               kNone, (size_t)n);
            hipLaunchKernelGGL(k_parents, dim3(blocksFor(ninner)), dim3(kB), 0, s, ninner, childL, childR,
                               parentInner, parentLeaf);
        }
        hipLaunchKernelGGL(k_set_u32, dim3(std::min<uint32_t>(blocksFor(ninner), 1024)), dim3(kB), 0, s, flags, 0u,
                           (size_t)ninner);
        hipLaunchKernelGGL(k_wide_cost, dim3(blocksFor(n)), dim3(kB), 0, s, n, childL, childR, parentInner,
                           parentLeaf, nodeLo, nodeHi, planCost, openSplit, innerSplit, flags);
    }
    STK_GB_ALLOC(StkWideNode, wide, wideCap, alloc(kGbWide, wideCap * sizeof(StkWideNode)))
    STK_GB_ALLOC(StkTriangle, wideTris, n, alloc(kGbWideTris, (size_t)n * sizeof(StkTriangle)))
    STK_GB_ALLOC(uint32_t, jobsA, wideCap, temp(wideCap * 4))
    STK_GB_ALLOC(uint32_t, jobsB, wideCap, temp(wideCap * 4))
    STK_GB_ALLOC(StkWideNode, plan, wideCap, temp(wideCap * sizeof(StkWideNode)))
    STK_GB_ALLOC(uint32_t, kidRefs, wideCap * 8, temp(wideCap * 32))
    STK_GB_ALLOC(uint32_t, innerCnt, wideCap, temp(wideCap * 4))
    STK_GB_ALLOC(uint32_t, leafCnt, wideCap, temp(wideCap * 4))
    STK_GB_ALLOC(uint32_t, innerOff, wideCap, temp(wideCap * 4))
    STK_GB_ALLOC(uint32_t, leafOff, wideCap, temp(wideCap * 4))
    size_t scanBytes = 0;
    if ((e = rocprim::exclusive_scan(nullptr, scanBytes, innerCnt, innerOff, 0u, wideCap, rocprim::plus<uint32_t>(), s)) !=
        hipSuccess)
        return e;
    STK_GB_ALLOC(void, scanTemp, scanBytes, temp(scanBytes))
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(kB), 0, s, jobsA, n > 1 ? 0u : kLeafBit, (size_t)1);
    if (n == 1) hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(kB), 0, s, counters + 2, 0u, (size_t)1);
    uint32_t nj = 1, levelStart = 0, triStart = 0, wideDepth = 0;
    uint32_t* jc = jobsA;
    uint32_t* jn = jobsB;
    while (nj > 0) {
        hipLaunchKernelGGL(k_wide_plan, dim3((nj + 63) / 64), dim3(64), 0, s, jc, nj, childL, childR, openSplit,
                           innerSplit, nodeLo, nodeHi, tris, plan, kidRefs, innerCnt, leafCnt, counters + 2);
        if ((e = rocprim::exclusive_scan(scanTemp, scanBytes, innerCnt, innerOff, 0u, nj, rocprim::plus<uint32_t>(),
                                         s)) != hipSuccess)
            return e;
        if ((e = rocprim::exclusive_scan(scanTemp, scanBytes, leafCnt, leafOff, 0u, nj, rocprim::plus<uint32_t>(),
                                         s)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(k_wide_emit, dim3(blocksFor(nj)), dim3(kB), 0, s, nj, levelStart, triStart, plan, kidRefs,
                           innerOff, leafOff, tris, wide, wideTris, jn);
        // (on a failed copy, wait for the ones queued before returning: they
        // write this frame's stack)
        uint32_t last[4];
        e = hipMemcpyAsync(&last[0], innerOff + nj - 1, 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(&last[1], innerCnt + nj - 1, 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(&last[2], leafOff + nj - 1, 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(&last[3], leafCnt + nj - 1, 4, hipMemcpyDeviceToHost, s);
        const hipError_t es = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        if ((e = es) != hipSuccess) return e;
        ++wideDepth;
        levelStart += nj;
        triStart += last[2] + last[3];
        nj = last[0] + last[1];
        if (levelStart + nj > wideCap) {
            err = "wide collapse: more nodes than the binary tree has inner nodes";
            return hipErrorInvalidValue;
        }
        std::swap(jc, jn);
    }
    uint32_t tail[3];
    float4 rootBox[2] = {};
    e = hipMemcpyAsync(tail, counters, 12, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && n > 1) e = hipMemcpyAsync(&rootBox[0], nodeLo, 16, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && n > 1) e = hipMemcpyAsync(&rootBox[1], nodeHi, 16, hipMemcpyDeviceToHost, s);
    const hipError_t es = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    if ((e = es) != hipSuccess) return e;
    if (tail[2] & 2u) {
        err = "inner node boxes or leaf counts are inconsistent";
        return hipErrorInvalidValue;
    }
    if (tail[2] != 0) {
        err = "wide collapse: a child box does not fit its node's quantisation grid";
        return hipErrorInvalidValue;
    }
    if (triStart != n) {
        err = "wide collapse: leaf count differs from the binary tree";
        return hipErrorInvalidValue;
    }
#undef STK_GB_ALLOC
    out.numNodes = 2 * n - 1;
    out.numPacked = ninner;
    out.numTris = n;
    out.numWide = levelStart;
    out.numEmit = tail[1];
    out.depth = depth;
    out.wideDepth = wideDepth;
    out.header.numTriangles = n;
    if (n == 1) {
        out.header.root = ~(int32_t)0;
    } else {
        out.header.root = 0;
        out.header.lo[0] = rootBox[0].x;
        out.header.lo[1] = rootBox[0].y;
        out.header.lo[2] = rootBox[0].z;
        out.header.hi[0] = rootBox[1].x;
        out.header.hi[1] = rootBox[1].y;
        out.header.hi[2] = rootBox[1].z;
    }
    return hipGetLastError();
}

} // namespace stk
