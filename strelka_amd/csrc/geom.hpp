// geom.hpp - float3 arithmetic shared by host C++ and HIP device code.
//
// Every helper evaluates in a fixed order and the whole library is compiled
// with -ffp-contract=off, so host and device produce the same bits for the
// same inputs (the traversal parity depends on it).
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define STK_HD __host__ __device__ __forceinline__
#else
#define STK_HD inline
#endif

namespace stk {

struct F3 {
    float x, y, z;
};

STK_HD F3 mk3(float x, float y, float z)
{
    F3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
STK_HD F3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }
STK_HD F3 operator+(F3 a, F3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
STK_HD F3 operator-(F3 a, F3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
STK_HD F3 operator*(F3 a, F3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
STK_HD F3 operator*(F3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
STK_HD F3 operator-(F3 a) { return mk3(-a.x, -a.y, -a.z); }
STK_HD float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
STK_HD F3 cross(F3 a, F3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

// mul(M, float4(p, w)).xyz, M column-major, glm's pairwise sum order
// (c0*x + c1*y) + (c2*z + c3*w) - the order glm uses for the BVH input
// positions (src/render/ptrender.cpp:488-490).
STK_HD F3 xform4(const float* m, F3 p, float w)
{
    return mk3((m[0] * p.x + m[4] * p.y) + (m[8] * p.z + m[12] * w),
               (m[1] * p.x + m[5] * p.y) + (m[9] * p.z + m[13] * w),
               (m[2] * p.x + m[6] * p.y) + (m[10] * p.z + m[14] * w));
}

// mul((float3x3)M, v)
STK_HD F3 xform3(const float* m, F3 v)
{
    return mk3((m[0] * v.x + m[4] * v.y) + m[8] * v.z, (m[1] * v.x + m[5] * v.y) + m[9] * v.z,
               (m[2] * v.x + m[6] * v.y) + m[10] * v.z);
}

STK_HD float bits_to_float(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}
STK_HD uint32_t float_to_bits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

} // namespace stk
