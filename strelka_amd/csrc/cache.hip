// cache.hip - radiance-cache kernels for gfx950 (see cache.hpp).
//
// MLP data flow on v_mfma_f32_16x16x16f16, computed transposed (Y^T = W^T X^T)
// so that a layer's accumulator layout (lane holds rows 4*(lane/16)+i of
// column lane%16) is exactly the B-operand layout the next layer needs (lane
// holds K rows 4*(lane/16)+i of column lane%16): activations never leave the
// registers between layers, forwards or backwards.  One wave handles 16
// samples (one per column); weights are pre-swizzled into per-lane A fragments
// (8 B per lane per MFMA).
#include <hip/hip_fp16.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "cache.hpp"

namespace stk {
namespace {

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;
constexpr int kFrag = 256; // halfs per A fragment (64 lanes x 4)

__device__ __forceinline__ f4 mfma(h4 a, h4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0); }

__device__ __forceinline__ h4 load_frag(const __half* base, int frag, int lane)
{
    return *reinterpret_cast<const h4*>(reinterpret_cast<const _Float16*>(base) + (size_t)frag * kFrag + lane * 4);
}

// FP16 conversions saturate at +-60000 so that no activation or gradient
// becomes inf (an inf would turn into NaN at the next multiply by zero).
constexpr float kHalfMax = 60000.0f;

__device__ __forceinline__ h4 relu_h(f4 a)
{
    h4 r;
    for (int i = 0; i < 4; ++i) r[i] = (_Float16)fminf(fmaxf(a[i], 0.0f), kHalfMax);
    return r;
}

__device__ __forceinline__ _Float16 sat_h(float x) { return (_Float16)fminf(fmaxf(x, -kHalfMax), kHalfMax); }

template <class F>
__device__ __forceinline__ void for_corners(const GridDesc& gd, int lvl, float px, float py, float pz, F&& f)
{
    const uint32_t res = gd.res[lvl];
    const float r = (float)res;
    const float x = px * r, y = py * r, z = pz * r;
    float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    float tx = x - fx, ty = y - fy, tz = z - fz;
    uint32_t ix = (uint32_t)fx, iy = (uint32_t)fy, iz = (uint32_t)fz;
    if (ix >= res) { ix = res - 1; tx = 1.0f; }
    if (iy >= res) { iy = res - 1; ty = 1.0f; }
    if (iz >= res) { iz = res - 1; tz = 1.0f; }
    const uint32_t stride = res + 1;
    const bool dense = gd.dense[lvl] != 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const uint32_t cx = ix + (c & 1), cy = iy + ((c >> 1) & 1), cz = iz + (c >> 2);
        const uint32_t idx = dense ? cx + stride * (cy + stride * cz)
                                   : ((cx * 1u) ^ (cy * 2654435761u) ^ (cz * 805459861u)) & gd.tableMask;
        const float w = ((c & 1) ? tx : 1.0f - tx) * ((c & 2) ? ty : 1.0f - ty) * ((c & 4) ? tz : 1.0f - tz);
        f(idx, w);
    }
}

__device__ __forceinline__ void encode_level(const GridDesc& gd, const __half* grid, int lvl, float px, float py,
                                             float pz, float& f0, float& f1)
{
    const __half2* tab = reinterpret_cast<const __half2*>(grid) + ((size_t)lvl << gd.log2Table);
    float a = 0.0f, b = 0.0f;
    for_corners(gd, lvl, px, py, pz, [&](uint32_t idx, float w) {
        const float2 v = __half22float2(tab[idx]);
        a += w * v.x;
        b += w * v.y;
    });
    f0 = a;
    f1 = b;
}

__device__ __forceinline__ void normalised(const GridDesc& gd, float4 pr, float& px, float& py, float& pz)
{
    px = fminf(fmaxf((pr.x - gd.lo[0]) * gd.scale, 0.0f), 1.0f);
    py = fminf(fmaxf((pr.y - gd.lo[1]) * gd.scale, 0.0f), 1.0f);
    pz = fminf(fmaxf((pr.z - gd.lo[2]) * gd.scale, 0.0f), 1.0f);
}

// Input fragments of one sample column: lane group g holds features
// 16s + 4g + i.  Features 0..31: level 2k -> (2k, 2k+1) of the grid
// (levels 2g, 2g+1 in s = 0; 8+2g, 9+2g in s = 1); 32..38: normal, outgoing
// direction, roughness; 39: constant 1; 40..63: zero.
__device__ __forceinline__ void input_frags(const GridDesc& gd, const __half* grid, const CacheRecord& rec, int g,
                                            bool valid, h4 x[4])
{
    for (int s = 0; s < 4; ++s) x[s] = h4{0, 0, 0, 0};
    if (!valid) return;
    float px, py, pz;
    normalised(gd, rec.posRough, px, py, pz);
    for (int s = 0; s < 2; ++s) {
        float a0, a1, b0, b1;
        encode_level(gd, grid, 8 * s + 2 * g, px, py, pz, a0, a1);
        encode_level(gd, grid, 8 * s + 2 * g + 1, px, py, pz, b0, b1);
        x[s] = h4{(_Float16)a0, (_Float16)a1, (_Float16)b0, (_Float16)b1};
    }
    if (g == 0)
        x[2] = h4{(_Float16)rec.normal.x, (_Float16)rec.normal.y, (_Float16)rec.normal.z, (_Float16)rec.dir.x};
    else if (g == 1)
        x[2] = h4{(_Float16)rec.dir.y, (_Float16)rec.dir.z, (_Float16)rec.posRough.w, (_Float16)1.0f};
}

// Forward through the 4 hidden layers; xs[l] receives the input of layer l
// (l = 0..4).  Returns the output accumulator (rows 4g+i = outputs).
__device__ __forceinline__ f4 forward(const __half* wFwd, int lane, h4 xs[5][4])
{
    for (int l = 0; l < kCacheHidden; ++l) {
        f4 acc[4];
        for (int t = 0; t < 4; ++t) {
            acc[t] = f4{0, 0, 0, 0};
            for (int s = 0; s < 4; ++s) acc[t] = mfma(load_frag(wFwd, l * 16 + t * 4 + s, lane), xs[l][s], acc[t]);
        }
        for (int t = 0; t < 4; ++t) xs[l + 1][t] = relu_h(acc[t]);
    }
    f4 out = f4{0, 0, 0, 0};
    for (int s = 0; s < 4; ++s) out = mfma(load_frag(wFwd, 64 + s, lane), xs[4][s], out);
    return out;
}

__global__ __launch_bounds__(256) void k_cache_infer(GridDesc gd, const __half* __restrict__ grid,
                                                     const __half* __restrict__ wFwd,
                                                     const CacheRecord* __restrict__ recs, uint32_t n,
                                                     const uint32_t* __restrict__ nDev, float4* __restrict__ out)
{
    const uint32_t N = nDev ? (*nDev < n ? *nDev : n) : n;
    const int lane = threadIdx.x % kWave, r = lane % 16, g = lane / 16;
    const uint32_t wave = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    for (uint32_t grp = wave; grp * 16 < N; grp += waves) {
        const uint32_t idx = grp * 16 + r;
        const bool valid = idx < N;
        CacheRecord rec;
        if (valid) rec = recs[idx];
        h4 xs[5][4];
        input_frags(gd, grid, rec, g, valid, xs[0]);
        const f4 o = forward(wFwd, lane, xs);
        if (valid && g == 0) out[idx] = make_float4(o[0], o[1], o[2], 0.0f);
    }
}

// Training records are walked in the order of a 21-bit Morton code of their
// normalised position (7 bits per axis), so the 16 samples of a wave lie close
// together and share most cells of the coarse grid levels.
constexpr int kMortonBits = 7;

__device__ __forceinline__ uint32_t spread3(uint32_t v)
{
    uint32_t x = 0;
    for (int b = 0; b < kMortonBits; ++b) x |= ((v >> b) & 1u) << (3 * b);
    return x;
}

// key[i]: Morton code of record i, or 1 << 21 past the record count (sorted
// last); val[i] = i.
__global__ __launch_bounds__(256) void k_cache_keys(GridDesc gd, const CacheRecord* __restrict__ recs, uint32_t n,
                                                    const uint32_t* __restrict__ nDev, uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t N = nDev ? (*nDev < n ? *nDev : n) : n;
    uint32_t key = 1u << (3 * kMortonBits);
    if (i < N) {
        float px, py, pz;
        normalised(gd, recs[i].posRough, px, py, pz);
        const float m = (float)((1 << kMortonBits) - 1);
        key = spread3((uint32_t)(px * m)) | (spread3((uint32_t)(py * m)) << 1) | (spread3((uint32_t)(pz * m)) << 2);
    }
    keys[i] = key;
    vals[i] = i;
}

// Adds (v0, v1) to tab[2e], tab[2e + 1] for every active lane, after summing
// the values of consecutive active lanes of the same 16-lane row that hold the
// same entry e (a row holds one grid level's 16 samples, which are in Morton
// order, so lanes that share a cell are mostly adjacent): one atomic pair per
// run.  A segmented inclusive scan over the row - the index of the run's first
// lane by a max-scan, then Hillis-Steele steps that add only within the run -
// leaves the run's sum in its last lane.  Every lane of the row must call it
// (the shuffles read the row's other lanes).
__device__ __forceinline__ void row_add_runs(float* tab, uint32_t e, bool active, float v0, float v1)
{
    const int r = (int)(__lane_id() & 15u);
    const uint32_t key = active ? e : 0xFFFFFFFFu;
    const uint32_t prev = (uint32_t)__shfl_up((int)key, 1, 16);
    int head = (r == 0 || prev != key) ? r : 0;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
        const int h = __shfl_up(head, off, 16);
        if (r >= off) head = max(head, h);
    }
    float s0 = active ? v0 : 0.0f, s1 = active ? v1 : 0.0f;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
        const float a0 = __shfl_up(s0, off, 16), a1 = __shfl_up(s1, off, 16);
        if (r - off >= head) {
            s0 += a0;
            s1 += a1;
        }
    }
    const uint32_t next = (uint32_t)__shfl_down((int)key, 1, 16);
    if (active && (r == 15 || next != key)) {
        atomicAdd(tab + 2 * e, s0);
        atomicAdd(tab + 2 * e + 1, s1);
    }
}

// Fused encode + forward + relative-L2 loss + backward for 16 samples per
// wave.  Saves every layer's input (xT) and output gradient (dyT), transposed
// to [row][sample], for the weight-gradient GEMM, and scatters the gradient of
// the grid features into gridG.
__global__ __launch_bounds__(256) void k_cache_train(GridDesc gd, const __half* __restrict__ grid,
                                                     const __half* __restrict__ wFwd,
                                                     const __half* __restrict__ wBwd,
                                                     const CacheRecord* __restrict__ recs,
                                                     const uint32_t* __restrict__ order, uint32_t n,
                                                     const uint32_t* __restrict__ nDev, uint32_t nCap,
                                                     _Float16* __restrict__ xT, _Float16* __restrict__ dyT,
                                                     float* __restrict__ gridG, float* __restrict__ gridGDense,
                                                     float* __restrict__ lossSum)
{
    const uint32_t N = nDev ? (*nDev < n ? *nDev : n) : n;
    const int lane = threadIdx.x % kWave, r = lane % 16, g = lane / 16;
    const uint32_t grp = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (grp * 16 >= N) return;
    const uint32_t idx = grp * 16 + r; // column of the saved activations (sorted order)
    const bool valid = idx < N;
    CacheRecord rec;
    if (valid) rec = recs[order[idx]];
    h4 xs[5][4];
    input_frags(gd, grid, rec, g, valid, xs[0]);
    const f4 o = forward(wFwd, lane, xs);
    for (int l = 0; l < 5; ++l)
        for (int s = 0; s < 4; ++s)
            for (int i = 0; i < 4; ++i) xT[((size_t)(l * 64 + 16 * s + 4 * g + i)) * nCap + idx] = xs[l][s][i];
    // relative L2: sum_c (p - t)^2 / (p^2 + 0.01), gradient 2 (p - t) / (p^2 + 0.01)
    h4 dy = h4{0, 0, 0, 0};
    float loss = 0.0f;
    if (valid && g == 0) {
        const float t[3] = {rec.target.x, rec.target.y, rec.target.z};
        for (int c = 0; c < 3; ++c) {
            const float den = o[c] * o[c] + 0.01f;
            const float d = o[c] - t[c];
            loss += d * d / den;
            dy[c] = sat_h(2.0f * d / den);
        }
    }
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    if (lane == 0 && loss != 0.0f) atomicAdd(lossSum, loss);
    for (int i = 0; i < 4; ++i) dyT[((size_t)(256 + 4 * g + i)) * nCap + idx] = dy[i];
    // output layer
    h4 dcur[4];
    for (int t = 0; t < 4; ++t) {
        const f4 a = mfma(load_frag(wBwd, 64 + t, lane), dy, f4{0, 0, 0, 0});
        for (int i = 0; i < 4; ++i) dcur[t][i] = sat_h(xs[4][t][i] > (_Float16)0 ? a[i] : 0.0f);
    }
    // hidden layers 3..1: dcur is the gradient at the output of layer l
    for (int l = kCacheHidden - 1; l >= 0; --l) {
        for (int s = 0; s < 4; ++s)
            for (int i = 0; i < 4; ++i) dyT[((size_t)(l * 64 + 16 * s + 4 * g + i)) * nCap + idx] = dcur[s][i];
        if (l == 0) break;
        h4 dnext[4];
        for (int t = 0; t < 4; ++t) {
            f4 a = f4{0, 0, 0, 0};
            for (int s = 0; s < 4; ++s) a = mfma(load_frag(wBwd, l * 16 + t * 4 + s, lane), dcur[s], a);
            for (int i = 0; i < 4; ++i) dnext[t][i] = sat_h(xs[l][t][i] > (_Float16)0 ? a[i] : 0.0f);
        }
        for (int t = 0; t < 4; ++t) dcur[t] = dnext[t];
    }
    // gradient of the grid features: rows 0..31 of W0 * dY0 (every lane takes
    // part: contributions merge along each row)
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    if (valid) normalised(gd, rec.posRough, px, py, pz);
    for (int t = 0; t < 2; ++t) {
        f4 a = f4{0, 0, 0, 0};
        for (int s = 0; s < 4; ++s) a = mfma(load_frag(wBwd, t * 4 + s, lane), dcur[s], a);
        for (int k = 0; k < 2; ++k) {
            const int lvl = 8 * t + 2 * g + k;
            const float d0 = a[2 * k], d1 = a[2 * k + 1];
            const bool act = valid && !(d0 == 0.0f && d1 == 0.0f);
            float* tab = gd.dense[lvl]
                             ? gridGDense + ((size_t)(blockIdx.x % kDenseCopies) * gd.denseEntries + gd.denseBase[lvl]) * 2
                             : gridG + ((size_t)lvl << gd.log2Table) * 2;
            for_corners(gd, lvl, px, py, pz,
                        [&](uint32_t e, float w) { row_add_runs(tab, e, act, w * d0, w * d1); });
        }
    }
}

// Folds the dense levels' gradient copies into gridG and clears them: one
// thread per (dense entry, feature), the only writer of its gridG element.
__global__ __launch_bounds__(256) void k_dense_fold(GridDesc gd, float* __restrict__ gridGDense,
                                                    float* __restrict__ gridG)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; // dense entry * 2 + feature
    if (i >= gd.denseEntries * 2) return;
    const uint32_t e = i / 2;
    int lvl = 0;
    for (int l = 0; l < kCacheLevels; ++l)
        if (gd.dense[l] && gd.denseBase[l] <= e) lvl = l;
    float sum = 0.0f;
    for (int k = 0; k < kDenseCopies; ++k) {
        float* c = gridGDense + (size_t)k * gd.denseEntries * 2 + i;
        sum += *c;
        *c = 0.0f;
    }
    if (sum != 0.0f) {
        float* g = gridG + ((size_t)lvl << gd.log2Table) * 2 + (size_t)(e - gd.denseBase[lvl]) * 2 + (i & 1u);
        *g += sum;
    }
}

// dW[l] = sum_n X_l[n]^T dY_l[n]: one wave per (16x16 tile, 1024-sample chunk).
__global__ __launch_bounds__(64) void k_cache_dw(const _Float16* __restrict__ xT, const _Float16* __restrict__ dyT,
                                                 uint32_t n, const uint32_t* __restrict__ nDev, uint32_t nCap,
                                                 float* __restrict__ mlpG)
{
    uint32_t N = nDev ? (*nDev < n ? *nDev : n) : n;
    N = (N + 15u) & ~15u;
    const uint32_t n0 = blockIdx.y * 1024u;
    if (n0 >= N) return;
    const uint32_t n1 = n0 + 1024u < N ? n0 + 1024u : N;
    const int lane = threadIdx.x, r = lane % 16, g = lane / 16;
    const int T = blockIdx.x;
    int l, t, u, dyRow, outW;
    size_t base;
    if (T < 64) {
        l = T / 16;
        t = (T / 4) % 4;
        u = T % 4;
        dyRow = l * 64;
        outW = 64;
        base = (size_t)l * 4096;
    } else {
        l = 4;
        t = T - 64;
        u = 0;
        dyRow = 256;
        outW = 16;
        base = 16384;
    }
    const _Float16* xa = xT + ((size_t)(l * 64 + 16 * t + r)) * nCap;
    const _Float16* yb = dyT + ((size_t)(dyRow + 16 * u + r)) * nCap;
    f4 acc = f4{0, 0, 0, 0};
    for (uint32_t k = n0; k < n1; k += 16) {
        const h4 a = *reinterpret_cast<const h4*>(xa + k + 4 * g);
        const h4 b = *reinterpret_cast<const h4*>(yb + k + 4 * g);
        acc = mfma(a, b, acc);
    }
    for (int i = 0; i < 4; ++i) {
        const int in = 16 * t + 4 * g + i, out = 16 * u + r;
        if (acc[i] != 0.0f) atomicAdd(mlpG + base + (size_t)in * outW + out, acc[i]);
    }
}

// Adam (Kingma & Ba 2015) with bias correction; `sparse` skips entries whose
// gradient is exactly zero (hash-grid entries no sample touched).  The gradient
// is consumed (zeroed).  gradScale multiplies the raw gradient; when nDev is
// given it is divided by *nDev (mean over the batch).
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                              float* __restrict__ gr, __half* __restrict__ half, size_t n,
                                              float gradScale, const uint32_t* __restrict__ nDev, uint32_t nLimit, float lr,
                                              float b1, float b2, float eps, float bc1, float bc2, int sparse,
                                              float* __restrict__ ema, float* __restrict__ emaW,
                                              __half* __restrict__ emaHalf, float emaDecay,
                                              const uint32_t* __restrict__ abortWord, uint32_t* __restrict__ stepsDone)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (abortWord && *abortWord)) return;
    if (stepsDone && i == 0) atomicAdd(stepsDone, 1u);
    const float raw = gr[i];
    if (sparse && raw == 0.0f) return;
    if (!isfinite(raw)) {
        gr[i] = 0.0f;
        return;
    }
    gr[i] = 0.0f;
    float scale = gradScale;
    if (nDev) scale /= (float)(*nDev == 0 ? 1u : (*nDev < nLimit ? *nDev : nLimit));
    const float g = raw * scale;
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    p[i] = pi;
    if (half) half[i] = __float2half(pi);
    if (ema) {
        const float e = emaDecay * ema[i] + (1.0f - emaDecay) * pi;
        ema[i] = e;
        if (emaW) {
            const float w = emaDecay * emaW[i] + (1.0f - emaDecay);
            emaW[i] = w;
            if (emaHalf) emaHalf[i] = __float2half(e / w);
        }
    }
}

__global__ __launch_bounds__(256) void k_to_half(const float* __restrict__ p, __half* __restrict__ h, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) h[i] = __float2half(p[i]);
}

// MFMA fragments from the FP32 master weights.
__global__ __launch_bounds__(256) void k_pack_frags(const float* __restrict__ w, _Float16* __restrict__ fwd,
                                                    _Float16* __restrict__ bwd, float scale)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x; // frag * 256 + lane * 4 + i
    if (id >= 68 * kFrag) return;
    const int f = id / kFrag, lane = (id % kFrag) / 4, i = id % 4, r = lane % 16, g = lane / 16;
    if (f < 64) {
        const int l = f / 16, t = (f / 4) % 4, s = f % 4;
        const float* W = w + l * 4096;
        fwd[id] = (_Float16)(scale * W[(16 * s + 4 * g + i) * 64 + 16 * t + r]); // A[out][in] = W[in][out]
        bwd[id] = (_Float16)(scale * W[(16 * t + r) * 64 + 16 * s + 4 * g + i]); // A[in][out] = W[in][out]
    } else {
        const float* W = w + 16384;
        const int q = f - 64;
        fwd[id] = (_Float16)(scale * W[(16 * q + 4 * g + i) * 16 + r]);
        bwd[id] = (_Float16)(scale * W[(16 * q + r) * 16 + 4 * g + i]);
    }
}

// Restarts the moving average: empty average and weight; until an element is
// updated its query copy is the current parameter.
__global__ __launch_bounds__(256) void k_reset_ema(const float* __restrict__ p, float* __restrict__ e,
                                                   float* __restrict__ ew, __half* __restrict__ eh, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    e[i] = 0.0f;
    if (ew) ew[i] = 0.0f;
    if (eh) eh[i] = __float2half(p[i]);
}

// The query copy of a restored moving average: the bias-corrected average of
// an element that has been updated, else the parameter itself (what
// k_reset_ema and k_adam leave behind).
__global__ __launch_bounds__(256) void k_restore_ema_half(const float* __restrict__ p, const float* __restrict__ e,
                                                          const float* __restrict__ ew, __half* __restrict__ eh,
                                                          size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float w = ew[i];
    eh[i] = __float2half(w > 0.0f ? e[i] / w : p[i]);
}

__global__ __launch_bounds__(256) void k_fill_zero(float* __restrict__ p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

inline uint32_t blocks256(size_t n) { return (uint32_t)((n + 255) / 256); }

} // namespace

hipError_t cacheInfer(const CacheBuffers& cb, const CacheRecord* recs, uint32_t n, const uint32_t* nDev, float4* out,
                      hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const uint32_t groups = (n + 15) / 16;
    const uint32_t g = groups / 4 + 1 < 4096 ? groups / 4 + 1 : 4096;
    const bool ema = cb.emaDecay > 0.0f;
    hipLaunchKernelGGL(k_cache_infer, dim3(g), dim3(256), 0, s, cb.grid, ema ? cb.gridHE : cb.gridH,
                       ema ? cb.wFwdE : cb.wFwd, recs, n, nDev, out);
    return hipGetLastError();
}

hipError_t cacheBackward(const CacheBuffers& cb, const CacheRecord* recs, uint32_t n, const uint32_t* nDev,
                         hipStream_t s)
{
    const uint32_t nMax = n < cb.nCap ? n : cb.nCap;
    if (nMax == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_zero, dim3(1), dim3(64), 0, s, cb.loss, (size_t)1);
    // the records' spatial order
    hipLaunchKernelGGL(k_cache_keys, dim3(blocks256(nMax)), dim3(256), 0, s, cb.grid, recs, nMax, nDev, cb.sortKeys,
                       cb.sortVals);
    size_t bytes = cb.sortBytes;
    hipError_t e = rocprim::radix_sort_pairs(cb.sortTemp, bytes, cb.sortKeys, cb.sortKeys + cb.nCap, cb.sortVals,
                                             cb.sortVals + cb.nCap, nMax, 0, 3 * kMortonBits + 1, s);
    if (e != hipSuccess) return e;
    const uint32_t groups = (nMax + 15) / 16;
    hipLaunchKernelGGL(k_cache_train, dim3((groups + 3) / 4), dim3(256), 0, s, cb.grid, cb.gridH, cb.wFwd, cb.wBwd,
                       recs, cb.sortVals + cb.nCap, nMax, nDev, cb.nCap, reinterpret_cast<_Float16*>(cb.xT),
                       reinterpret_cast<_Float16*>(cb.dyT), cb.gridG, cb.gridGDense, cb.loss);
    if (cb.grid.denseEntries)
        hipLaunchKernelGGL(k_dense_fold, dim3(blocks256((size_t)cb.grid.denseEntries * 2)), dim3(256), 0, s, cb.grid,
                           cb.gridGDense, cb.gridG);
    const uint32_t chunks = (nMax + 1023) / 1024;
    hipLaunchKernelGGL(k_cache_dw, dim3(68, chunks), dim3(64), 0, s, reinterpret_cast<const _Float16*>(cb.xT),
                       reinterpret_cast<const _Float16*>(cb.dyT), nMax, nDev, cb.nCap, cb.mlpG);
    return hipGetLastError();
}

hipError_t cacheAdam(CacheBuffers& cb, float gradScale, const uint32_t* nDev, uint32_t nLimit,
                     const uint32_t* abortWord, hipStream_t s)
{
    cb.step += 1;
    const float lr = cb.lrHalfLife > 0.0f
                         ? fmaxf(cb.lrMin, cb.lr * exp2f(-(float)(cb.step - 1) / cb.lrHalfLife))
                         : cb.lr;
    const float bc1 = 1.0f - powf(cb.beta1, (float)cb.step);
    const float bc2 = 1.0f - powf(cb.beta2, (float)cb.step);
    const bool ema = cb.emaDecay > 0.0f;
    hipLaunchKernelGGL(k_adam, dim3(blocks256(cb.gridParams)), dim3(256), 0, s, cb.gridP, cb.gridM, cb.gridV, cb.gridG,
                       cb.gridH, cb.gridParams, gradScale, nDev, nLimit, lr, cb.beta1, cb.beta2, cb.eps, bc1, bc2,
                       1, ema ? cb.gridE : nullptr, ema ? cb.gridEW : nullptr, ema ? cb.gridHE : nullptr,
                       cb.emaDecay, abortWord, (uint32_t*)nullptr);
    hipLaunchKernelGGL(k_adam, dim3(blocks256(kCacheMlpParams)), dim3(256), 0, s, cb.mlpP, cb.mlpM, cb.mlpV, cb.mlpG,
                       (__half*)nullptr, (size_t)kCacheMlpParams, gradScale, nDev, nLimit, lr, cb.beta1, cb.beta2,
                       cb.eps, bc1, bc2, 0, ema ? cb.mlpE : nullptr, (float*)nullptr, (__half*)nullptr, cb.emaDecay,
                       abortWord, cb.stepsDone);
    hipLaunchKernelGGL(k_pack_frags, dim3(blocks256(68 * kFrag)), dim3(256), 0, s, cb.mlpP,
                       reinterpret_cast<_Float16*>(cb.wFwd), reinterpret_cast<_Float16*>(cb.wBwd), 1.0f);
    if (ema) {
        // the MLP's moving average is updated every step: weight 1 - decay^step
        cb.emaSteps += 1;
        const float w = 1.0f - powf(cb.emaDecay, (float)cb.emaSteps);
        hipLaunchKernelGGL(k_pack_frags, dim3(blocks256(68 * kFrag)), dim3(256), 0, s, cb.mlpE,
                           reinterpret_cast<_Float16*>(cb.wFwdE), reinterpret_cast<_Float16*>(cb.wBwdE), 1.0f / w);
    }
    return hipGetLastError();
}

hipError_t cachePack(CacheBuffers& cb, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_half, dim3(blocks256(cb.gridParams)), dim3(256), 0, s, cb.gridP, cb.gridH, cb.gridParams);
    hipLaunchKernelGGL(k_pack_frags, dim3(blocks256(68 * kFrag)), dim3(256), 0, s, cb.mlpP,
                       reinterpret_cast<_Float16*>(cb.wFwd), reinterpret_cast<_Float16*>(cb.wBwd), 1.0f);
    if (cb.emaDecay > 0.0f) {
        // the moving average restarts (empty); queries read the current
        // parameters until the first update
        cb.emaSteps = 0;
        hipLaunchKernelGGL(k_reset_ema, dim3(blocks256(cb.gridParams)), dim3(256), 0, s, cb.gridP, cb.gridE,
                           cb.gridEW, cb.gridHE, cb.gridParams);
        hipLaunchKernelGGL(k_reset_ema, dim3(blocks256(kCacheMlpParams)), dim3(256), 0, s, cb.mlpP, cb.mlpE,
                           (float*)nullptr, (__half*)nullptr, (size_t)kCacheMlpParams);
        hipLaunchKernelGGL(k_pack_frags, dim3(blocks256(68 * kFrag)), dim3(256), 0, s, cb.mlpP,
                           reinterpret_cast<_Float16*>(cb.wFwdE), reinterpret_cast<_Float16*>(cb.wBwdE), 1.0f);
    }
    return hipGetLastError();
}

hipError_t cacheRestore(CacheBuffers& cb, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_half, dim3(blocks256(cb.gridParams)), dim3(256), 0, s, cb.gridP, cb.gridH, cb.gridParams);
    hipLaunchKernelGGL(k_pack_frags, dim3(blocks256(68 * kFrag)), dim3(256), 0, s, cb.mlpP,
                       reinterpret_cast<_Float16*>(cb.wFwd), reinterpret_cast<_Float16*>(cb.wBwd), 1.0f);
    if (cb.emaDecay > 0.0f) {
        hipLaunchKernelGGL(k_restore_ema_half, dim3(blocks256(cb.gridParams)), dim3(256), 0, s, cb.gridP, cb.gridE,
                           cb.gridEW, cb.gridHE, cb.gridParams);
        // the MLP average as k_adam leaves it after emaSteps updates (the raw
        // weights before the first)
        const bool avg = cb.emaSteps > 0;
        const float w = avg ? 1.0f - powf(cb.emaDecay, (float)cb.emaSteps) : 1.0f;
        hipLaunchKernelGGL(k_pack_frags, dim3(blocks256(68 * kFrag)), dim3(256), 0, s, avg ? cb.mlpE : cb.mlpP,
                           reinterpret_cast<_Float16*>(cb.wFwdE), reinterpret_cast<_Float16*>(cb.wBwdE), 1.0f / w);
    }
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks256(cb.gridParams)), dim3(256), 0, s, cb.gridG, cb.gridParams);
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks256(kCacheMlpParams)), dim3(256), 0, s, cb.mlpG,
                       (size_t)kCacheMlpParams);
    return hipGetLastError();
}

hipError_t cacheSortBytes(uint32_t nCap, size_t& bytes)
{
    bytes = 0;
    return rocprim::radix_sort_pairs(nullptr, bytes, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                     (uint32_t*)nullptr, nCap, 0, 3 * kMortonBits + 1);
}

hipError_t cacheZeroGrads(const CacheBuffers& cb, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks256(cb.gridParams)), dim3(256), 0, s, cb.gridG, cb.gridParams);
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks256(kCacheMlpParams)), dim3(256), 0, s, cb.mlpG,
                       (size_t)kCacheMlpParams);
    return hipGetLastError();
}

} // namespace stk
