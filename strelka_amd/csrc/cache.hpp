// cache.hpp - online-trained radiance cache: multi-resolution hash grid (FP16)
// feeding a 64-wide MLP on MFMA (FP16 operands, FP32 accumulation), with
// hand-written encode, forward, backward, weight-gradient GEMM and Adam
// kernels.  No counterpart in the reference renderer.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace stk {

constexpr int kCacheLevels = 16;   // hash-grid levels
constexpr int kCacheFeatures = 2;  // features per level
constexpr int kCacheWidth = 64;    // MLP width (input padded to 64)
constexpr int kCacheHidden = 4;    // hidden layers (64x64 each, ReLU)
constexpr int kCacheOut = 16;      // output layer width (3 used: RGB)
// MLP parameters: 4 hidden 64x64 matrices (layer 0 maps the padded input)
// and one 64x16 output matrix, no biases (feature 39 is a constant 1).
constexpr int kCacheMlpParams = kCacheHidden * kCacheWidth * kCacheWidth + kCacheWidth * kCacheOut;

// Copies of the dense levels' gradient table: training block b adds into copy
// b % kDenseCopies.  Blocks are dispatched to the 8 XCDs in turn, so each copy
// is written by one XCD and stays in its L2 (the dense levels of a 2^19-entry
// grid take 2.6 MB); the copies are folded into gridG after the scatter.
// Every sample touches the same few thousand coarse entries, and one shared
// copy serialised those atomics.
constexpr int kDenseCopies = 8;

struct GridDesc {
    float lo[3];
    float scale;                 // normalised position = (p - lo) * scale, in [0,1]
    uint32_t res[kCacheLevels];  // cells per unit of normalised coordinate
    uint32_t dense[kCacheLevels];
    uint32_t denseBase[kCacheLevels]; // first entry of a dense level in the dense gradient copies
    uint32_t denseEntries;       // entries of all dense levels
    uint32_t tableMask;          // entries per level - 1 (power of two)
    uint32_t log2Table;
};

// One cache sample (query or training): 64 B.
struct CacheRecord {
    float4 posRough;   // position.xyz, roughness
    float4 normal;     // normal.xyz, w: payload (path slot bits for queries)
    float4 dir;        // outgoing direction.xyz
    float4 target;     // training target radiance.rgb (unused for queries)
};

struct CacheBuffers {
    GridDesc grid;
    // hash grid: [level][entry][feature]
    __half* gridH = nullptr;   // FP16 copy read by the kernels
    float* gridP = nullptr;    // FP32 master
    float* gridM = nullptr;    // Adam first moment
    float* gridV = nullptr;    // Adam second moment
    float* gridG = nullptr;    // gradient accumulator
    float* gridGDense = nullptr; // [kDenseCopies][denseEntries][feature] dense-level gradients (zero between steps)
    size_t gridParams = 0;
    // MLP: fp32 master weights W[l][in][out]
    float* mlpP = nullptr;
    float* mlpM = nullptr;
    float* mlpV = nullptr;
    float* mlpG = nullptr;
    __half* wFwd = nullptr;    // MFMA A fragments of W^T (forward)
    __half* wBwd = nullptr;    // MFMA A fragments of W (backward)
    // exponential moving average of the parameters, read by queries.  The
    // averages start at zero and are divided by their weight 1 - decay^k (k =
    // updates of the element) so early queries are not pulled towards the
    // initial parameters: the grid is updated sparsely, so it keeps k per
    // element as the weight array gridEW; the MLP is updated densely (k = step).
    float emaDecay = 0.0f;     // 0: queries read the raw parameters
    float* gridEW = nullptr;
    float* gridE = nullptr;
    __half* gridHE = nullptr;
    float* mlpE = nullptr;
    __half* wFwdE = nullptr;
    __half* wBwdE = nullptr;   // unused by the kernels; keeps k_pack_frags' interface
    // training scratch
    __half* xT = nullptr;      // saved layer inputs  [5][64][nCap]
    __half* dyT = nullptr;     // saved layer output gradients [4][64][nCap] + [16][nCap]
    uint32_t nCap = 0;         // capacity in samples (multiple of 1024)
    uint32_t* sortKeys = nullptr; // [2][nCap] spatial sort of the training records: keys
    uint32_t* sortVals = nullptr; // [2][nCap] and record indices (the second half: sorted)
    void* sortTemp = nullptr;
    size_t sortBytes = 0;
    float* loss = nullptr;     // [0] loss sum of the last step
    uint32_t* stepsDone = nullptr; // device count of the optimiser steps applied (equals step between frames;
                                   // fewer after a frame the watchdog abandoned)
    uint32_t step = 0;
    uint32_t emaSteps = 0;     // updates since the moving average restarted
    float lr = 1e-2f, beta1 = 0.9f, beta2 = 0.99f, eps = 1e-15f;
    float lrHalfLife = 0.0f, lrMin = 0.0f; // learning-rate decay (0: constant)
};

// In every call, n is the host-side maximum record count; when nDev is not
// null the kernels read the actual count from device memory (so a producer
// kernel can set it without a host round trip).
//
// Forward only: out[i] = radiance (rgb, w = 0) for the records.
hipError_t cacheInfer(const CacheBuffers& cb, const CacheRecord* recs, uint32_t n, const uint32_t* nDev,
                      float4* out, hipStream_t s);
// Gradients of sum_i relL2(pred_i, target_i) into gridG / mlpG (accumulated;
// Adam consumes and zeroes them).  Writes the loss sum to cb.loss[0].
hipError_t cacheBackward(const CacheBuffers& cb, const CacheRecord* recs, uint32_t n, const uint32_t* nDev,
                         hipStream_t s);
// One Adam step on every parameter from the accumulated gradients times
// gradScale (divided by *nDev when given), then refreshes the FP16 copies and
// MFMA fragments.  abortWord (may be null): a nonzero word leaves every
// parameter, moment and gradient as it was (a frame the watchdog abandoned;
// the caller then restores step / emaSteps and calls cacheRestore).
hipError_t cacheAdam(CacheBuffers& cb, float gradScale, const uint32_t* nDev, uint32_t nLimit,
                     const uint32_t* abortWord, hipStream_t s);
// Rebuilds the FP16 grid copy and the fragments from the FP32 masters.
hipError_t cachePack(CacheBuffers& cb, hipStream_t s);
// After the FP32 state (masters, moments, moving averages, their weights,
// step and emaSteps) was written back from a checkpoint: rebuilds every
// derived copy without restarting the moving average, and zeroes the
// gradient accumulators.
hipError_t cacheRestore(CacheBuffers& cb, hipStream_t s);
hipError_t cacheZeroGrads(const CacheBuffers& cb, hipStream_t s);
// Temporary bytes of the training records' sort for nCap records.
hipError_t cacheSortBytes(uint32_t nCap, size_t& bytes);

} // namespace stk
