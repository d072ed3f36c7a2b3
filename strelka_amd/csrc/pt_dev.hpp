// pt_dev.hpp - device-side path-tracing building blocks for gfx950.
//
// RNG (shaders/random.h:14-47), camera rays (shaders/pathtracerMdl.hlsl:95-202),
// ray offset (shaders/helper.h:12-28), rect-light sampling
// (shaders/lights.h:102-301), attribute unpacking (shaders/pack.h:4-33) and the
// non-MDL Lambert/GGX BSDFs (shaders/pathtracer.hlsl:108-234,
// shaders/helper.h:41-99) with the output conventions of Bsdf_sample_data /
// Bsdf_evaluate_data (shaders/mdl_types.hlsl:271-304).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/stk/stk.h"
#include "fmath.h"
#include "geom.hpp"

namespace stk {
namespace dev {

constexpr float kPi = 3.1415926535897f;
constexpr float kInvPi = 1.0f / kPi;

__device__ __forceinline__ F3 nrm(F3 a)
{
    const float inv = 1.0f / sqrtf(dot(a, a));
    return a * inv;
}
__device__ __forceinline__ float len3(F3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ float saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

// ---- RNG ---------------------------------------------------------------------
__device__ __forceinline__ uint32_t jenkins(uint32_t x)
{
    x += x << 10;
    x ^= x >> 6;
    x += x << 3;
    x ^= x >> 11;
    x += x << 15;
    return x;
}
__device__ __forceinline__ uint32_t init_rng(uint32_t px, uint32_t py, uint32_t width, uint32_t frame)
{
    return jenkins((px + py * width) ^ jenkins(frame));
}
__device__ __forceinline__ float rnd(uint32_t& s)
{
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    return bits_to_float(0x3f800000u | (s >> 9)) - 1.0f;
}
__device__ __forceinline__ uint32_t permute(uint32_t i, uint32_t l, uint32_t p)
{
    uint32_t w = l - 1;
    w |= w >> 1;
    w |= w >> 2;
    w |= w >> 4;
    w |= w >> 8;
    w |= w >> 16;
    do {
        i ^= p;
        i *= 0xe170893du;
        i ^= p >> 16;
        i ^= (i & w) >> 4;
        i ^= p >> 8;
        i *= 0x0929eb3fu;
        i ^= p >> 23;
        i ^= (i & w) >> 1;
        i *= 1u | p >> 27;
        i *= 0x6935fa69u;
        i ^= (i & w) >> 11;
        i *= 0x74dcb303u;
        i ^= (i & w) >> 2;
        i *= 0x9e501cc3u;
        i ^= (i & w) >> 2;
        i *= 0xc860a3dfu;
        i &= w;
        i ^= i >> 5;
    } while (i >= l);
    return (i + p) % l;
}
__device__ __forceinline__ float randfloat(uint32_t i, uint32_t p)
{
    i ^= p;
    i ^= i >> 17;
    i ^= i >> 10;
    i *= 0xb36534e5u;
    i ^= i >> 12;
    i ^= i >> 21;
    i *= 0x93fc4795u;
    i ^= 0xdf6e307fu;
    i ^= i >> 17;
    i *= 1u | p >> 18;
    return (float)i * (1.0f / 4294967808.0f);
}

// ---- camera ------------------------------------------------------------------
__device__ __forceinline__ void camera_ray(const StkPathTracerParam& p, uint32_t px, uint32_t py, uint32_t& rng,
                                           uint32_t s, float4& o, float4& d)
{
    float jx = 0.5f, jy = 0.5f;
    if (p.stratifiedSamplingType == 1) {
        jx = rnd(rng);
        jy = rnd(rng);
    } else if (p.stratifiedSamplingType == 2) {
        const uint32_t N = 16;
        const int m = (int)sqrtf((float)N);
        const int n = ((int)N + m - 1) / m;
        const int x = (int)((s / (uint32_t)n) % (uint32_t)n);
        const int y = (int)((s % (uint32_t)m) % (uint32_t)m);
        const float r0 = rnd(rng);
        jx = ((float)x + ((float)y + r0) / (float)n) / (float)m;
        const float r1 = rnd(rng);
        jy = ((float)y + ((float)x + r1) / (float)m) / (float)n;
    } else if (p.stratifiedSamplingType == 3) {
        const int N = 16, pp = 16;
        const int m = (int)sqrtf((float)N * 1.0f);
        const int n = (N + m - 1) / m;
        const int si = (int)permute(s, (uint32_t)N, (uint32_t)pp * 0x51633e2du);
        const int sx = (int)permute((uint32_t)(si % m), (uint32_t)m, (uint32_t)pp * 0x68bc21ebu);
        const int sy = (int)permute((uint32_t)(si / m), (uint32_t)n, (uint32_t)pp * 0x02e5be93u);
        const float fx = randfloat((uint32_t)si, (uint32_t)pp * 0x967a889bu);
        const float fy = randfloat((uint32_t)si, (uint32_t)pp * 0x368cc8b7u);
        jx = ((float)sx + ((float)sy + fx) / (float)n) / (float)m;
        jy = ((float)si + fy) / (float)N;
    }
    const float posx = (float)px + jx, posy = (float)py + jy;
    const float ndcx = (posx / (float)p.dimension[0]) * 2.0f - 1.0f;
    const float ndcy = (posy / (float)p.dimension[1]) * 2.0f - 1.0f;
    const float* C = p.clipToView;
    const F3 v = mk3((C[0] * ndcx + C[4] * ndcy) + (C[8] * 1.0f + C[12] * 1.0f),
                     (C[1] * ndcx + C[5] * ndcy) + (C[9] * 1.0f + C[13] * 1.0f),
                     (C[2] * ndcx + C[6] * ndcy) + (C[10] * 1.0f + C[14] * 1.0f));
    const F3 w = xform4(p.viewToWorld, v, 0.0f);
    const F3 org = xform4(p.viewToWorld, mk3(0.0f, 0.0f, 0.0f), 1.0f);
    const F3 dir = nrm(w);
    o = make_float4(org.x, org.y, org.z, 1e9f);
    d = make_float4(dir.x, dir.y, dir.z, 0.0f);
}

// ---- offset_ray ----------------------------------------------------------------
__device__ __forceinline__ F3 offset_ray(F3 p, F3 n)
{
    const float origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f, int_scale = 256.0f;
    const int ox = (int)(int_scale * n.x), oy = (int)(int_scale * n.y), oz = (int)(int_scale * n.z);
    const float pix = __int_as_float(__float_as_int(p.x) + ((p.x < 0) ? -ox : ox));
    const float piy = __int_as_float(__float_as_int(p.y) + ((p.y < 0) ? -oy : oy));
    const float piz = __int_as_float(__float_as_int(p.z) + ((p.z < 0) ? -oz : oz));
    return mk3(fabsf(p.x) < origin ? p.x + float_scale * n.x : pix,
               fabsf(p.y) < origin ? p.y + float_scale * n.y : piy,
               fabsf(p.z) < origin ? p.z + float_scale * n.z : piz);
}

// ---- rect light (SphQuad, Urena et al. 2013) -----------------------------------
struct LightSample {
    F3 L;
    float dist;
    F3 normal;
    float pdf;        // the reference's area-measure pdf d^2 / (cos_l A) (shaders/lights.h:298)
    float solidAngle; // S: the samples are uniform over this solid angle, pdf 1 / S
};

// The spherical rectangle of a rect light seen from o (SphQuad init,
// shaders/lights.h:102-166).
struct SphQuadF {
    F3 X, Y, Z;
    float z0, z0sq, x0, y0, y0sq, x1, y1, y1sq, b0, b1, b0sq, k, S;
};

__device__ __forceinline__ SphQuadF sphquad_init(const StkLight& l, F3 o)
{
    SphQuadF q;
    const F3 p0 = ld3(l.points[0]), p1 = ld3(l.points[1]), p3 = ld3(l.points[3]);
    const F3 ex = p1 - p0, ey = p3 - p0;
    const float exl = len3(ex), eyl = len3(ey);
    q.X = mk3(ex.x / exl, ex.y / exl, ex.z / exl);
    q.Y = mk3(ey.x / eyl, ey.y / eyl, ey.z / eyl);
    q.Z = cross(q.X, q.Y);
    const F3 d = p0 - o;
    q.z0 = dot(d, q.Z);
    if (q.z0 > 0) {
        q.Z = q.Z * -1.0f;
        q.z0 *= -1.0f;
    }
    q.z0sq = q.z0 * q.z0;
    q.x0 = dot(d, q.X);
    q.y0 = dot(d, q.Y);
    q.x1 = q.x0 + exl;
    q.y1 = q.y0 + eyl;
    q.y0sq = q.y0 * q.y0;
    q.y1sq = q.y1 * q.y1;
    const F3 v00 = mk3(q.x0, q.y0, q.z0), v01 = mk3(q.x0, q.y1, q.z0), v10 = mk3(q.x1, q.y0, q.z0),
             v11 = mk3(q.x1, q.y1, q.z0);
    const F3 n0 = nrm(cross(v00, v10)), n1 = nrm(cross(v10, v11));
    const F3 n2 = nrm(cross(v11, v01)), n3 = nrm(cross(v01, v00));
    const float g0 = stk_acosf(-dot(n0, n1)), g1 = stk_acosf(-dot(n1, n2));
    const float g2 = stk_acosf(-dot(n2, n3)), g3 = stk_acosf(-dot(n3, n0));
    q.b0 = n0.z;
    q.b1 = n2.z;
    q.b0sq = q.b0 * q.b0;
    q.k = 2.0f * kPi - g2 - g3;
    q.S = g0 + g1 - q.k;
    return q;
}

// SphQuadSample (shaders/lights.h:168-194): the point of the rect at (u, v).
__device__ __forceinline__ F3 sphquad_sample(const SphQuadF& q, F3 o, float u, float v)
{
    const float au = u * q.S + q.k;
    const float fu = (stk_cosf(au) * q.b0 - q.b1) / stk_sinf(au);
    float cu = 1.0f / sqrtf(fu * fu + q.b0sq) * (fu > 0 ? 1.0f : -1.0f);
    cu = fminf(fmaxf(cu, -1.0f), 1.0f);
    float xu = -(cu * q.z0) / sqrtf(1.0f - cu * cu);
    xu = fminf(fmaxf(xu, q.x0), q.x1);
    const float dd = sqrtf(xu * xu + q.z0sq);
    const float h0 = q.y0 / sqrtf(dd * dd + q.y0sq);
    const float h1 = q.y1 / sqrtf(dd * dd + q.y1sq);
    const float hv = h0 + v * (h1 - h0);
    const float hv2 = hv * hv;
    const float yv = (hv < 1.0f - 1e-5f) ? (hv * dd) / sqrtf(1.0f - hv2) : q.y1;
    return ((o + q.X * xu) + q.Y * yv) + q.Z * q.z0;
}

// calcLightNormal of a rect (shaders/lights.h:230-246): it emits along this.
__device__ __forceinline__ F3 rect_light_normal(const StkLight& l)
{
    const F3 p0 = ld3(l.points[0]), p1 = ld3(l.points[1]), p3 = ld3(l.points[3]);
    return -nrm(cross(p1 - p0, p3 - p0));
}

__device__ __forceinline__ LightSample sample_rect_light(const StkLight& l, float u, float v, F3 o)
{
    const SphQuadF q = sphquad_init(l, o);
    const F3 pt = sphquad_sample(q, o, u, v);
    // fillLightData (:274-282)
    const F3 p0 = ld3(l.points[0]), p1 = ld3(l.points[1]), p3 = ld3(l.points[3]);
    const F3 ex = p1 - p0, ey = p3 - p0;
    const float A = len3(cross(ex, ey));
    const F3 nl = -nrm(cross(ex, ey));
    const F3 toL = pt - o;
    const float lenL = len3(toL);
    LightSample ls;
    ls.L = mk3(toL.x / lenL, toL.y / lenL, toL.z / lenL);
    ls.dist = lenL;
    ls.normal = nl;
    ls.pdf = lenL * lenL / (-dot(ls.L, nl) * A); // area-measure pdf of the reference (:298)
    ls.solidAngle = q.S;
    return ls;
}

// ---- unpack / interpolate ------------------------------------------------------
__device__ __forceinline__ F3 unpack_normal(uint32_t v)
{
    F3 n;
    n.z = (float)((v & 0xfff00000u) >> 20) / 511.99999f * 2.0f - 1.0f;
    n.y = (float)((v & 0x000ffc00u) >> 10) / 511.99999f * 2.0f - 1.0f;
    n.x = (float)(v & 0x000003ffu) / 511.99999f * 2.0f - 1.0f;
    return n;
}
__device__ __forceinline__ F3 interp(F3 a, F3 b, F3 c, float u, float v)
{
    const float w = 1.0f - u - v;
    return (a * w + b * u) + c * v;
}

// ---- albedo textures ---------------------------------------------------------
// unpackUV (shaders/pack.h:15-23)
__device__ __forceinline__ void unpack_uv(uint32_t p, float& s, float& t)
{
    t = (float)((p & 0xffff0000u) >> 16) / 16383.99999f * 20.0f - 10.0f;
    s = (float)(p & 0x0000ffffu) / 16383.99999f * 20.0f - 10.0f;
}

// One axis of the lookup: the wrap of apply_wrap_and_crop
// (shaders/mdl_runtime.hlsl:93-120; repeat leaves the coordinate to the
// fetch, mirrored repeat folds it into [0, 1], and every wrap but repeat then
// clamps it to the edge texels' centres), then the quintic smoothing of the
// fractional texel position (apply_smootherstep_filter, :125-140).  Returns
// the filter's texel coordinate (x * n - 0.5).
__device__ __forceinline__ float tex_axis(float c, uint32_t n, uint32_t wrap)
{
    const float fn = (float)n;
    if (wrap != STK_TEX_REPEAT) {
        if (wrap == STK_TEX_MIRRORED_REPEAT) {
            const float fl = floorf(c);
            c = ((int)fl & 1) != 0 ? 1.0f - (c - fl) : c - fl;
        }
        const float h = 0.5f / fn;
        c = fminf(fmaxf(c, h), 1.0f - h);
    }
    float r = c * fn + 0.5f;
    const float i = floorf(r);
    float f = r - i;
    f = f * f * f * (f * (f * 6.0f - 15.0f) + 10.0f);
    r = i + f;
    return ((r - 0.5f) / fn) * fn - 0.5f;
}

__device__ __forceinline__ int tex_wrap(int i, int n, bool clampEdge)
{
    if (clampEdge) return i < 0 ? 0 : (i >= n ? n - 1 : i);
    const int m = i % n;
    return m < 0 ? m + n : m;
}

// Bilinear RGB of texture tx at (s, t); a clip texture reads zero outside
// [0, 1) on either axis (tex_lookup_float4_2d, :186-189).
__device__ __forceinline__ F3 tex_lookup(const StkTexture& tx, const float4* __restrict__ texels, float s, float t)
{
    if (tx.wrap == STK_TEX_CLIP && (s < 0.0f || s >= 1.0f || t < 0.0f || t >= 1.0f)) return mk3(0.0f, 0.0f, 0.0f);
    const bool ce = tx.wrap != STK_TEX_REPEAT;
    const float x = tex_axis(s, tx.width, tx.wrap), y = tex_axis(t, tx.height, tx.wrap);
    const float x0 = floorf(x), y0 = floorf(y);
    const float fx = x - x0, fy = y - y0;
    const int w = (int)tx.width, h = (int)tx.height;
    const int ix0 = tex_wrap((int)x0, w, ce), ix1 = tex_wrap((int)x0 + 1, w, ce);
    const int iy0 = tex_wrap((int)y0, h, ce), iy1 = tex_wrap((int)y0 + 1, h, ce);
    const float4* row0 = texels + tx.offset + (size_t)iy0 * tx.width;
    const float4* row1 = texels + tx.offset + (size_t)iy1 * tx.width;
    const float4 a = row0[ix0], b = row0[ix1], c = row1[ix0], d = row1[ix1];
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    return mk3((a.x * gx + b.x * fx) * gy + (c.x * gx + d.x * fx) * fy,
               (a.y * gx + b.y * fx) * gy + (c.y * gx + d.y * fx) * fy,
               (a.z * gx + b.z * fx) * gy + (c.z * gx + d.z * fx) * fy);
}

// CalcBumpedNormal (shaders/helper.h:101-115): the tangent negated and
// made orthogonal to the normal, the bitangent N x T, and the map's texel
// (rgb = (n + 1) / 2) taken from that tangent frame to world space.
__device__ __forceinline__ F3 bumped_normal(F3 normal, F3 tangent, F3 texel)
{
    const F3 n = nrm(normal);
    F3 t = -nrm(tangent);
    t = nrm(t - n * dot(t, n));
    const F3 b = cross(n, t);
    const F3 m = mk3(texel.x * 2.0f - 1.0f, texel.y * 2.0f - 1.0f, texel.z * 2.0f - 1.0f);
    return nrm((t * m.x + b * m.y) + n * m.z);
}

// ---- BSDFs --------------------------------------------------------------------
// Events of a BSDF sample: 0 absorb, 1 diffuse reflection, 2 glossy
// reflection, 3 specular reflection, 4 specular transmission (the reference's
// BSDF_EVENT_* classes, shaders/mdl_types.hlsl:235-248).  Specular samples
// report pdf 0.
constexpr int kEventSpecularReflection = 3;
constexpr int kEventSpecularTransmission = 4;
struct BsdfSample {
    F3 k2;
    float pdf;
    F3 weight; // bsdf * cos / pdf
    int event;
};

__device__ __forceinline__ bool is_specular(const StkMaterial& m)
{
    return m.type == STK_MAT_MIRROR || m.type == STK_MAT_GLASS;
}

__device__ __forceinline__ void tangent_space(F3 n, F3& t, F3& b)
{
    F3 helper = mk3(1.0f, 0.0f, 0.0f);
    if (fabsf(n.x) > 0.99f) helper = mk3(0.0f, 0.0f, 1.0f);
    t = nrm(cross(n, helper));
    b = nrm(cross(n, t));
}
__device__ __forceinline__ F3 to_world(F3 t, F3 b, F3 n, F3 v) { return (t * v.x + b * v.y) + n * v.z; }
__device__ __forceinline__ float ggx_alpha(float roughness)
{
    const float a = roughness * roughness;
    return a < 1e-4f ? 1e-4f : a;
}
__device__ __forceinline__ float ggx_d(float a2, F3 n, F3 m)
{
    const float c = dot(n, m);
    const float c2 = c * c;
    const float den = c2 * (a2 - 1.0f) + 1.0f;
    return c > 0.0f ? (a2 / (kPi * den * den)) : 0.0f;
}
__device__ __forceinline__ float ggx_g1(float a2, F3 w, F3 n, F3 m)
{
    const float NdotW = dot(n, w), MdotW = dot(m, w);
    if (MdotW * NdotW <= 0.0f) return 0.0f;
    const float c2 = NdotW * NdotW;
    const float tan2 = (1.0f - c2) / c2;
    if (tan2 == 0.0f) return 1.0f;
    return 2.0f / (1.0f + sqrtf(1.0f + a2 * tan2));
}
__device__ __forceinline__ F3 schlick(F3 f0, float c)
{
    const float w = stk_pow5f(1.0f - c);
    return mk3(f0.x + (1.0f - f0.x) * w, f0.y + (1.0f - f0.y) * w, f0.z + (1.0f - f0.z) * w);
}

__device__ __forceinline__ BsdfSample bsdf_sample(const StkMaterial& m, F3 N, F3 Ng, F3 k1, float x0, float x1,
                                                  float x2)
{
    BsdfSample s;
    s.k2 = mk3(0, 0, 0);
    s.pdf = 0.0f;
    s.weight = mk3(0, 0, 0);
    s.event = 0;
    const F3 albedo = ld3(m.albedo);
    if (m.type == STK_MAT_MIRROR) {
        const float c1 = dot(N, k1);
        const F3 k2 = N * (2.0f * c1) - k1;
        if (c1 <= 0.0f || dot(k2, Ng) <= 0.0f) return s;
        s.k2 = k2;
        s.weight = schlick(albedo, c1);
        s.event = kEventSpecularReflection;
        return s;
    }
    if (m.type == STK_MAT_GLASS) {
        // the side is the geometric normal's: entering through the front face
        const float ior = m.ior > 0.0f ? m.ior : 1.5f;
        const bool entering = dot(k1, Ng) > 0.0f;
        const F3 n = entering ? N : -N, ng = entering ? Ng : -Ng;
        const float eta = entering ? 1.0f / ior : ior; // n1 / n2
        const float c1 = dot(n, k1);
        if (c1 <= 0.0f) return s;
        const float s2 = eta * eta * (1.0f - c1 * c1); // sin^2 of the refracted angle
        float F = 1.0f, c2 = 0.0f;
        if (s2 < 1.0f) {
            // Fresnel's equations for unpolarised light
            c2 = sqrtf(1.0f - s2);
            const float rs = (eta * c1 - c2) / (eta * c1 + c2);
            const float rp = (c1 - eta * c2) / (c1 + eta * c2);
            F = 0.5f * (rs * rs + rp * rp);
        }
        if (x2 < F) {
            const F3 k2 = n * (2.0f * c1) - k1;
            if (dot(k2, ng) <= 0.0f) return s;
            s.k2 = k2;
            s.weight = mk3(1.0f, 1.0f, 1.0f);
            s.event = kEventSpecularReflection;
        } else {
            const F3 k2 = n * (eta * c1 - c2) - k1 * eta; // Snell's law
            if (dot(k2, ng) >= 0.0f) return s;
            s.k2 = k2;
            s.weight = albedo;
            s.event = kEventSpecularTransmission;
        }
        return s;
    }
    F3 T, B;
    tangent_space(N, T, B);
    if (m.type == STK_MAT_GGX) {
        const float alpha = ggx_alpha(m.roughness);
        const float a2 = alpha * alpha;
        // SampleGGXDistribution (shaders/helper.h:93-99)
        const float cosT = sqrtf((1.0f - x0) / (x0 * (alpha * alpha - 1.0f) + 1.0f));
        const float sinT = sqrtf(saturate(1.0f - cosT * cosT));
        const float phi = 2.0f * kPi * x1;
        const F3 hl = mk3(stk_cosf(phi) * sinT, stk_sinf(phi) * sinT, cosT + 1e-6f);
        const F3 h = nrm(to_world(T, B, N, hl));
        const float k1h = dot(k1, h);
        const F3 k2 = h * (2.0f * k1h) - k1;
        const float nk1 = dot(N, k1), nk2 = dot(N, k2), nh = dot(N, h);
        if (nk1 <= 0.0f || nk2 <= 0.0f || dot(k2, Ng) <= 0.0f || k1h <= 0.0f) return s;
        const float D = ggx_d(a2, N, h);
        const float G = ggx_g1(a2, k1, N, h) * ggx_g1(a2, k2, N, h);
        const F3 F = schlick(albedo, dot(k2, h));
        s.k2 = k2;
        s.pdf = D * nh / (4.0f * k1h);
        s.weight = F * (G * k1h / (nk1 * nh));
        s.event = 2;
        return s;
    }
    // Lambert: SampleHemisphereCosine (shaders/helper.h:73-79) in GetTangentSpace
    const float cosT = sqrtf(saturate(x0));
    const float sinT = sqrtf(saturate(1.0f - cosT * cosT));
    const float phi = 2.0f * kPi * x1;
    const F3 k2 = to_world(T, B, N, mk3(stk_cosf(phi) * sinT, stk_sinf(phi) * sinT, cosT + 1e-6f));
    if (dot(k2, Ng) <= 0.0f) return s;
    s.k2 = k2;
    s.pdf = kInvPi * (saturate(dot(N, k2)) + 1e-6f);
    s.weight = albedo;
    s.event = 1;
    return s;
}

// Returns bsdf*cos split into diffuse and glossy parts and the solid-angle pdf
// (zero for the specular models).
__device__ __forceinline__ void bsdf_eval(const StkMaterial& m, F3 N, F3 Ng, F3 k1, F3 k2, F3& diffuse, F3& glossy,
                                          float& pdf)
{
    diffuse = mk3(0, 0, 0);
    glossy = mk3(0, 0, 0);
    pdf = 0.0f;
    const float nk2 = dot(N, k2);
    if (is_specular(m) || nk2 <= 0.0f || dot(k2, Ng) <= 0.0f) return; // a specular BSDF is zero off its direction
    const F3 albedo = ld3(m.albedo);
    if (m.type == STK_MAT_GGX) {
        const float nk1 = dot(N, k1);
        if (nk1 <= 0.0f) return;
        const float alpha = ggx_alpha(m.roughness);
        const float a2 = alpha * alpha;
        const F3 h = nrm(k1 + k2);
        const float k1h = dot(k1, h);
        if (k1h <= 0.0f) return;
        const float nh = dot(N, h);
        const float D = ggx_d(a2, N, h);
        const float G = ggx_g1(a2, k1, N, h) * ggx_g1(a2, k2, N, h);
        const F3 F = schlick(albedo, dot(k2, h));
        glossy = F * (D * G / (4.0f * nk1));
        pdf = D * nh / (4.0f * k1h);
        return;
    }
    diffuse = albedo * (kInvPi * nk2);
    pdf = kInvPi * (nk2 + 1e-6f);
}

} // namespace dev
} // namespace stk
