/*
 * oracle_fp64.c - binary64 twin of the oracle's path sample (TEST
 * INFRASTRUCTURE ONLY; linked into liboracle.so beside oracle.c).
 *
 * orc_pt_sample_fp64 runs the estimator of orc_pt_sample - the reference's
 * pathTraceCameraRays (shaders/pathtracerMdl.hlsl:204-404) with refCompat = 1,
 * or the corrected estimator with refCompat = 0 - with every geometric and
 * shading quantity in binary64: camera rays, triangles (getTriangle's
 * objectToWorld in binary64), the closest-hit and shadow walks over the same
 * skip-link BVH, the spherical-rectangle light sample, the BSDFs and the path
 * throughput.  The random numbers are the reference's (the same xorshift
 * stream and pixel jitter), so only the arithmetic differs.  A path may take
 * other branches than its binary32 twin; converged pixel means are what the
 * tests compare, within the samples' statistical error.
 *
 * Shadow and bounce origins keep the reference's offset_ray
 * (shaders/helper.h:12-28, applied to the binary32-rounded point and normal):
 * its offset of up to 256 binary32 ulps per coordinate (about 2e-4 at |x| =
 * 10) is part of the estimator, not rounding - with an ideal offset (1e-9 of
 * the scene scale) the converged 1M-triangle view differs by up to 25% in
 * pixels of terrain creases, where the offset origins see past nearby
 * geometry.  Boxes are tested widened by 1e-9 relative, so no hit is lost at
 * box faces.  Texture lookups are not part of this twin (its scenes are
 * untextured).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    double x, y, z;
} d3;

static inline d3 D(double x, double y, double z)
{
    d3 r = {x, y, z};
    return r;
}
static inline d3 dl(const float* p) { return D(p[0], p[1], p[2]); }
static inline d3 dadd(d3 a, d3 b) { return D(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 dsub(d3 a, d3 b) { return D(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 dmul(d3 a, d3 b) { return D(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline d3 dscl(d3 a, double s) { return D(a.x * s, a.y * s, a.z * s); }
static inline double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 dcross(d3 a, d3 b) { return D(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline double dlen(d3 a) { return sqrt(ddot(a, a)); }
static inline d3 dnrm(d3 a) { return dscl(a, 1.0 / dlen(a)); }
static inline double dsat(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

static const double kPi = 3.14159265358979323846;

static d3 dxform4(const float m[16], d3 p, double w)
{
    return D((double)m[0] * p.x + (double)m[4] * p.y + (double)m[8] * p.z + (double)m[12] * w,
             (double)m[1] * p.x + (double)m[5] * p.y + (double)m[9] * p.z + (double)m[13] * w,
             (double)m[2] * p.x + (double)m[6] * p.y + (double)m[10] * p.z + (double)m[14] * w);
}
static d3 dxform3(const float m[16], d3 v)
{
    return D((double)m[0] * v.x + (double)m[4] * v.y + (double)m[8] * v.z,
             (double)m[1] * v.x + (double)m[5] * v.y + (double)m[9] * v.z,
             (double)m[2] * v.x + (double)m[6] * v.y + (double)m[10] * v.z);
}

/* the 10:10:10 normal packing over [-1, 1] (shaders/pack.h:4-13) */
static d3 dunpack_normal(uint32_t v)
{
    return D((double)(v & 0x3ffu) / 511.99999 * 2.0 - 1.0, (double)((v >> 10) & 0x3ffu) / 511.99999 * 2.0 - 1.0,
             (double)((v >> 20) & 0xfffu) / 511.99999 * 2.0 - 1.0);
}

typedef struct {
    d3 o, d;
    double tmax;
} Ray64;

typedef struct {
    double t, u, v;
    uint32_t inst, prim;
} Hit64;

static void tri64(const OrcScene* s, uint32_t inst, uint32_t prim, d3* p0, d3* p1, d3* p2)
{
    const StkInstance* ic = &s->inst[inst];
    const uint32_t base = (uint32_t)ic->indexOffset + prim * 3;
    *p0 = dxform4(ic->objectToWorld, dl(s->vb[s->ib[base]].position), 1.0);
    *p1 = dxform4(ic->objectToWorld, dl(s->vb[s->ib[base + 1]].position), 1.0);
    *p2 = dxform4(ic->objectToWorld, dl(s->vb[s->ib[base + 2]].position), 1.0);
}

/* Moller-Trumbore with the reference's rejection rules (shaders/raytracing.h:36-88) */
static int rt64(const Ray64* r, d3 v0, d3 e0, d3 e1, double* t, double* u, double* v)
{
    const d3 p = dcross(r->d, e1);
    const double det = ddot(e0, p);
    if (fabs(det) < 1e-6) return 0;
    const double inv = 1.0 / det;
    const d3 tv = dsub(r->o, v0);
    *u = ddot(tv, p) * inv;
    if (*u < 0.0 || *u > 1.0) return 0;
    const d3 q = dcross(tv, e0);
    *v = ddot(r->d, q) * inv;
    if (*v < 0.0 || *u + *v > 1.0) return 0;
    *t = ddot(e1, q) * inv;
    return *t >= 1e-6;
}

static int box64(const Ray64* r, const float lo[3], const float hi[3], double* tEnter)
{
    const double ro[3] = {r->o.x, r->o.y, r->o.z}, rd[3] = {r->d.x, r->d.y, r->d.z};
    double t0 = 0.0, t1 = INFINITY;
    for (int a = 0; a < 3; ++a) {
        const double pad = 1e-9 * (fabs((double)lo[a]) + fabs((double)hi[a]) + 1.0);
        const double l = (double)lo[a] - pad, h = (double)hi[a] + pad;
        if (rd[a] == 0.0) {
            if (ro[a] < l || ro[a] > h) return 0;
            continue;
        }
        const double inv = 1.0 / rd[a];
        double tn = (l - ro[a]) * inv, tf = (h - ro[a]) * inv;
        if (tn > tf) {
            const double x = tn;
            tn = tf;
            tf = x;
        }
        if (tn > t0) t0 = tn;
        if (tf < t1) t1 = tf;
    }
    *tEnter = t0;
    return t1 >= t0;
}

/* Walks every entered box of the skip-link tree.  mode 0: closest hit;
 * mode 1: the reference's any hit (first leaf in preorder, t < tmax); mode 2:
 * any triangle of a non-emitter instance with t < tmax. */
static int walk64(const OrcScene* s, const Ray64* r, int mode, Hit64* h)
{
    int found = 0;
    h->t = r->tmax;
    h->inst = h->prim = STK_INVALID_INDEX;
    uint32_t idx = 0;
    while (s->nnodes && idx != STK_INVALID_INDEX) {
        const StkBvhNode* n = &s->nodes[idx];
        if ((uint32_t)n->instId != STK_INVALID_INDEX) {
            uint32_t prim;
            memcpy(&prim, &n->minBounds[0], 4);
            d3 p0, p1, p2;
            tri64(s, (uint32_t)n->instId, prim, &p0, &p1, &p2);
            double t, u, v;
            if (rt64(r, p0, dsub(p1, p0), dsub(p2, p0), &t, &u, &v) && t < h->t) {
                const int emitter = s->inst[n->instId].lightId != -1;
                if (mode != 2 || !emitter) {
                    found = 1;
                    h->t = mode == 0 ? t : h->t;
                    h->u = u;
                    h->v = v;
                    h->inst = (uint32_t)n->instId;
                    h->prim = prim;
                    if (mode != 0) {
                        h->t = t;
                        return 1;
                    }
                }
            }
            idx = (uint32_t)n->nodeOffset;
            continue;
        }
        double te;
        const double limit = mode == 0 ? h->t : r->tmax;
        if (box64(r, n->minBounds, n->maxBounds, &te) && te <= limit) {
            ++idx;
            continue;
        }
        idx = (uint32_t)n->nodeOffset;
    }
    return found;
}

typedef struct {
    d3 X, Y, Z;
    double z0, z0sq, x0, y0, y0sq, x1, y1, y1sq, b0, b1, b0sq, k, S;
} Quad64;

/* SphQuad (Urena et al. 2013; shaders/lights.h:102-194) in binary64 */
static Quad64 quad64(const StkLight* l, d3 o)
{
    Quad64 q;
    const d3 p0 = dl(l->points[0]), ex = dsub(dl(l->points[1]), p0), ey = dsub(dl(l->points[3]), p0);
    const double exl = dlen(ex), eyl = dlen(ey);
    q.X = dscl(ex, 1.0 / exl);
    q.Y = dscl(ey, 1.0 / eyl);
    q.Z = dcross(q.X, q.Y);
    const d3 d = dsub(p0, o);
    q.z0 = ddot(d, q.Z);
    if (q.z0 > 0) {
        q.Z = dscl(q.Z, -1.0);
        q.z0 = -q.z0;
    }
    q.z0sq = q.z0 * q.z0;
    q.x0 = ddot(d, q.X);
    q.y0 = ddot(d, q.Y);
    q.x1 = q.x0 + exl;
    q.y1 = q.y0 + eyl;
    q.y0sq = q.y0 * q.y0;
    q.y1sq = q.y1 * q.y1;
    const d3 v00 = D(q.x0, q.y0, q.z0), v01 = D(q.x0, q.y1, q.z0), v10 = D(q.x1, q.y0, q.z0),
             v11 = D(q.x1, q.y1, q.z0);
    const d3 n0 = dnrm(dcross(v00, v10)), n1 = dnrm(dcross(v10, v11));
    const d3 n2 = dnrm(dcross(v11, v01)), n3 = dnrm(dcross(v01, v00));
    const double g0 = acos(-ddot(n0, n1)), g1 = acos(-ddot(n1, n2));
    const double g2 = acos(-ddot(n2, n3)), g3 = acos(-ddot(n3, n0));
    q.b0 = n0.z;
    q.b1 = n2.z;
    q.b0sq = q.b0 * q.b0;
    q.k = 2.0 * kPi - g2 - g3;
    q.S = g0 + g1 - q.k;
    return q;
}

static d3 quad64_sample(const Quad64* q, d3 o, double u, double v)
{
    const double au = u * q->S + q->k;
    const double fu = (cos(au) * q->b0 - q->b1) / sin(au);
    double cu = 1.0 / sqrt(fu * fu + q->b0sq) * (fu > 0 ? 1.0 : -1.0);
    cu = cu < -1.0 ? -1.0 : (cu > 1.0 ? 1.0 : cu);
    double xu = -(cu * q->z0) / sqrt(1.0 - cu * cu);
    xu = xu < q->x0 ? q->x0 : (xu > q->x1 ? q->x1 : xu);
    const double dd = sqrt(xu * xu + q->z0sq);
    const double h0 = q->y0 / sqrt(dd * dd + q->y0sq);
    const double h1 = q->y1 / sqrt(dd * dd + q->y1sq);
    const double hv = h0 + v * (h1 - h0);
    const double yv = (hv < 1.0 - 1e-5) ? (hv * dd) / sqrt(1.0 - hv * hv) : q->y1;
    return dadd(dadd(dadd(o, dscl(q->X, xu)), dscl(q->Y, yv)), dscl(q->Z, q->z0));
}

static d3 light_normal64(const StkLight* l)
{
    const d3 p0 = dl(l->points[0]);
    return dscl(dnrm(dcross(dsub(dl(l->points[1]), p0), dsub(dl(l->points[3]), p0))), -1.0);
}

static void tangent64(d3 n, d3* t, d3* b)
{
    d3 helper = fabs(n.x) > 0.99 ? D(0, 0, 1) : D(1, 0, 0);
    *t = dnrm(dcross(n, helper));
    *b = dnrm(dcross(n, *t));
}
static d3 world64(d3 t, d3 b, d3 n, d3 v) { return dadd(dadd(dscl(t, v.x), dscl(b, v.y)), dscl(n, v.z)); }
static double alpha64(const StkMaterial* m)
{
    const double a = (double)m->roughness * (double)m->roughness;
    return a < 1e-4 ? 1e-4 : a;
}
static double ggxD64(double a2, d3 n, d3 m)
{
    const double c = ddot(n, m), den = c * c * (a2 - 1.0) + 1.0;
    return c > 0.0 ? a2 / (kPi * den * den) : 0.0;
}
static double ggxG1_64(double a2, d3 w, d3 n, d3 m)
{
    const double nw = ddot(n, w);
    if (ddot(m, w) * nw <= 0.0) return 0.0;
    const double tan2 = (1.0 - nw * nw) / (nw * nw);
    return tan2 == 0.0 ? 1.0 : 2.0 / (1.0 + sqrt(1.0 + a2 * tan2));
}
static d3 schlick64(d3 f0, double c)
{
    const double w = pow(1.0 - c, 5.0);
    return D(f0.x + (1.0 - f0.x) * w, f0.y + (1.0 - f0.y) * w, f0.z + (1.0 - f0.z) * w);
}

/* bsdf * cos (diffuse or glossy) and the solid-angle pdf (the oracle's
 * Lambert / GGX, shaders/pathtracer.hlsl:108-234, shaders/pbr.hlsl:94-151) */
static d3 eval64(const StkMaterial* m, d3 N, d3 Ng, d3 k1, d3 k2, double* pdf)
{
    *pdf = 0.0;
    const double nk2 = ddot(N, k2);
    if (m->type == STK_MAT_MIRROR || m->type == STK_MAT_GLASS) return D(0, 0, 0); /* specular: a delta */
    if (nk2 <= 0.0 || ddot(k2, Ng) <= 0.0) return D(0, 0, 0);
    const d3 alb = dl(m->albedo);
    if (m->type == STK_MAT_GGX) {
        const double nk1 = ddot(N, k1);
        if (nk1 <= 0.0) return D(0, 0, 0);
        const double a = alpha64(m), a2 = a * a;
        const d3 h = dnrm(dadd(k1, k2));
        const double k1h = ddot(k1, h);
        if (k1h <= 0.0) return D(0, 0, 0);
        const double Dh = ggxD64(a2, N, h);
        *pdf = Dh * ddot(N, h) / (4.0 * k1h);
        return dscl(schlick64(alb, ddot(k2, h)), Dh * ggxG1_64(a2, k1, N, h) * ggxG1_64(a2, k2, N, h) / (4.0 * nk1));
    }
    *pdf = (nk2 + 1e-6) / kPi;
    return dscl(alb, nk2 / kPi);
}

/* Returns the event: 0 absorb, 1 diffuse, 2 glossy, 3 specular reflection,
 * 4 specular transmission (mirror and glass as the binary32 oracle's models). */
static int sample64(const StkMaterial* m, d3 N, d3 Ng, d3 k1, double x0, double x1, double x2, d3* k2, d3* weight,
                    double* pdf)
{
    const d3 alb = dl(m->albedo);
    *pdf = 0.0;
    if (m->type == STK_MAT_MIRROR) {
        const double c1 = ddot(N, k1);
        *k2 = dsub(dscl(N, 2.0 * c1), k1);
        if (c1 <= 0.0 || ddot(*k2, Ng) <= 0.0) return 0;
        *weight = schlick64(alb, c1);
        return 3;
    }
    if (m->type == STK_MAT_GLASS) {
        const double ior = m->ior > 0.0f ? (double)m->ior : 1.5;
        const int entering = ddot(k1, Ng) > 0.0;
        const d3 n = entering ? N : dscl(N, -1.0), ng = entering ? Ng : dscl(Ng, -1.0);
        const double eta = entering ? 1.0 / ior : ior;
        const double c1 = ddot(n, k1);
        if (c1 <= 0.0) return 0;
        const double s2 = eta * eta * (1.0 - c1 * c1);
        double F = 1.0, c2 = 0.0;
        if (s2 < 1.0) {
            c2 = sqrt(1.0 - s2);
            const double rs = (eta * c1 - c2) / (eta * c1 + c2), rp = (c1 - eta * c2) / (c1 + eta * c2);
            F = 0.5 * (rs * rs + rp * rp);
        }
        if (x2 < F) {
            *k2 = dsub(dscl(n, 2.0 * c1), k1);
            if (ddot(*k2, ng) <= 0.0) return 0;
            *weight = D(1, 1, 1);
            return 3;
        }
        *k2 = dsub(dscl(n, eta * c1 - c2), dscl(k1, eta));
        if (ddot(*k2, ng) >= 0.0) return 0;
        *weight = alb;
        return 4;
    }
    d3 T, B;
    tangent64(N, &T, &B);
    const double phi = 2.0 * kPi * x1;
    if (m->type == STK_MAT_GGX) {
        const double a = alpha64(m), a2 = a * a;
        const double ct = sqrt((1.0 - x0) / (x0 * (a * a - 1.0) + 1.0)), st = sqrt(dsat(1.0 - ct * ct));
        const d3 h = dnrm(world64(T, B, N, D(cos(phi) * st, sin(phi) * st, ct + 1e-6)));
        const double k1h = ddot(k1, h);
        *k2 = dsub(dscl(h, 2.0 * k1h), k1);
        const double nk1 = ddot(N, k1), nk2 = ddot(N, *k2), nh = ddot(N, h);
        if (nk1 <= 0.0 || nk2 <= 0.0 || ddot(*k2, Ng) <= 0.0 || k1h <= 0.0) return 0;
        *pdf = ggxD64(a2, N, h) * nh / (4.0 * k1h);
        *weight = dscl(schlick64(alb, ddot(*k2, h)),
                       ggxG1_64(a2, k1, N, h) * ggxG1_64(a2, *k2, N, h) * k1h / (nk1 * nh));
        return 2;
    }
    const double ct = sqrt(dsat(x0)), st = sqrt(dsat(1.0 - ct * ct));
    *k2 = world64(T, B, N, D(cos(phi) * st, sin(phi) * st, ct + 1e-6));
    if (ddot(*k2, Ng) <= 0.0) return 0;
    *pdf = (dsat(ddot(N, *k2)) + 1e-6) / kPi;
    *weight = alb;
    return 1;
}

/* offset_ray of the binary32 point and normal (the reference's origin offset) */
static d3 offset64(d3 p, d3 n)
{
    const float pf[3] = {(float)p.x, (float)p.y, (float)p.z}, nf[3] = {(float)n.x, (float)n.y, (float)n.z};
    float o[3];
    orc_offset_ray(pf, nf, o);
    return D(o[0], o[1], o[2]);
}

void orc_pt_sample_fp64(const OrcScene* s, const StkRenderParams* rp, uint32_t px, uint32_t py, uint32_t sampleNum,
                        double out[3])
{
    const StkPathTracerParam* p = &rp->pt;
    const int ref = rp->refCompat != 0;
    const uint32_t frame = p->frameNumber * (uint32_t)p->spp + sampleNum;
    uint32_t rng = orc_init_rng(px, py, (uint32_t)p->dimension[0], frame);
    float jx, jy;
    orc_pixel_jitter(p, frame, &rng, &jx, &jy);
    /* generateCameraRay (shaders/pathtracerMdl.hlsl:165-202) in binary64 */
    const double ndcx = (((double)px + jx) / (double)p->dimension[0]) * 2.0 - 1.0;
    const double ndcy = (((double)py + jy) / (double)p->dimension[1]) * 2.0 - 1.0;
    const float* C = p->clipToView;
    const d3 vv = D((double)C[0] * ndcx + (double)C[4] * ndcy + (double)C[8] + (double)C[12],
                    (double)C[1] * ndcx + (double)C[5] * ndcy + (double)C[9] + (double)C[13],
                    (double)C[2] * ndcx + (double)C[6] * ndcy + (double)C[10] + (double)C[14]);
    Ray64 ray;
    ray.o = dxform4(p->viewToWorld, D(0, 0, 0), 1.0);
    ray.d = dnrm(dxform4(p->viewToWorld, vv, 0.0));
    ray.tmax = 1e9;
    d3 color = D(0, 0, 0), thr = D(1, 1, 1);
    const d3 sky = dl(rp->skyColor);
    double prevPdf = 0.0;
    int specularBounce = 0;
    for (int depth = 0; depth < (int)p->maxDepth; ++depth) {
        Hit64 hit;
        if (!walk64(s, &ray, 0, &hit)) {
            color = dadd(color, dmul(thr, sky));
            break;
        }
        const StkInstance* ic = &s->inst[hit.inst];
        if (ic->lightId != -1) {
            const StkLight* l = &s->lights[ic->lightId];
            if (ref) {
                if (depth == 0 || specularBounce) color = dadd(color, dmul(thr, dl(l->color)));
            } else if (ddot(ray.d, light_normal64(l)) < 0.0) {
                double w = 1.0;
                if (depth > 0 && !specularBounce) {
                    const Quad64 q = quad64(l, ray.o);
                    const double lpdf = q.S > 0.0 ? (1.0 / (double)s->nlights) / q.S : 0.0;
                    w = prevPdf / (prevPdf + lpdf);
                }
                color = dadd(color, dscl(dmul(thr, dl(l->color)), w));
            }
            break;
        }
        const uint32_t base = (uint32_t)ic->indexOffset + hit.prim * 3;
        const StkVertex *v0 = &s->vb[s->ib[base]], *v1 = &s->vb[s->ib[base + 1]], *v2 = &s->vb[s->ib[base + 2]];
        d3 p0, p1, p2;
        tri64(s, hit.inst, hit.prim, &p0, &p1, &p2);
        const d3 Ng = dnrm(dcross(dsub(p1, p0), dsub(p2, p0)));
        const double w0 = 1.0 - hit.u - hit.v;
        const d3 P = dadd(dadd(dscl(p0, w0), dscl(p1, hit.u)), dscl(p2, hit.v));
        const d3 n0 = dxform3(ic->normalMatrix, dunpack_normal(v0->normal));
        const d3 n1 = dxform3(ic->normalMatrix, dunpack_normal(v1->normal));
        const d3 n2 = dxform3(ic->normalMatrix, dunpack_normal(v2->normal));
        const d3 N = dnrm(dadd(dadd(dscl(n0, w0), dscl(n1, hit.u)), dscl(n2, hit.v)));
        StkMaterial mat;
        if (ic->materialId >= 0 && (uint32_t)ic->materialId < s->nmats) {
            mat = s->mats[ic->materialId];
        } else {
            memset(&mat, 0, sizeof(mat));
            mat.type = STK_MAT_LAMBERT;
            mat.roughness = 1.0f;
            mat.albedo[0] = mat.albedo[1] = mat.albedo[2] = 0.5f;
        }
        const d3 k1 = dscl(ray.d, -1.0);
        /* the EDF: the front face emits (:299-304) */
        if (ddot(k1, Ng) > 0.0) color = dadd(color, dmul(thr, dl(mat.emission)));
        /* sampleLights / estimateDirectLighting (:31-92) */
        d3 toLight = D(0, 0, 0), radiance = D(0, 0, 0);
        double lightPdf = 0.0;
        if (s->nlights > 0) {
            const uint32_t lid = (uint32_t)((float)s->nlights * orc_rand(&rng));
            const StkLight* l = &s->lights[lid < s->nlights ? lid : s->nlights - 1];
            const double u0 = orc_rand(&rng), u1 = orc_rand(&rng);
            const Quad64 q = quad64(l, P);
            const d3 pt = quad64_sample(&q, P, u0, u1);
            const d3 toL = dsub(pt, P);
            const double dist = dlen(toL);
            toLight = dscl(toL, 1.0 / dist);
            const d3 nl = light_normal64(l);
            const d3 Li = dl(l->color);
            if (ddot(N, toLight) > 0.0 && -ddot(toLight, nl) > 0.0 && Li.x != 0.0 && Li.y != 0.0 && Li.z != 0.0) {
                Ray64 sr;
                sr.o = offset64(P, Ng);
                sr.d = toLight;
                sr.tmax = dist;
                Hit64 sh;
                double vis;
                if (ref) {
                    vis = walk64(s, &sr, 1, &sh) && s->inst[sh.inst].lightId == -1 ? 0.0 : 1.0;
                    const d3 e0 = dsub(dl(l->points[1]), dl(l->points[0]));
                    const d3 e1 = dsub(dl(l->points[3]), dl(l->points[0]));
                    lightPdf = dist * dist / (-ddot(toLight, nl) * dlen(dcross(e0, e1)));
                    radiance = dscl(Li, vis * dsat(ddot(N, toLight)));
                } else {
                    vis = walk64(s, &sr, 2, &sh) ? 0.0 : 1.0;
                    lightPdf = q.S > 0.0 ? 1.0 / q.S : 0.0;
                    radiance = dscl(Li, vis);
                }
            }
            lightPdf *= 1.0 / (double)s->nlights;
        }
        if (ddot(toLight, Ng) > 0.0 && lightPdf != 0.0) {
            double bpdf;
            const d3 f = eval64(&mat, N, Ng, k1, toLight, &bpdf);
            if (bpdf > 0.0) {
                const double mis = lightPdf / (lightPdf + bpdf);
                color = dadd(color, dmul(dscl(dmul(thr, radiance), mis / lightPdf), f));
            }
        }
        const double x0 = orc_rand(&rng), x1 = orc_rand(&rng), x2 = orc_rand(&rng);
        (void)orc_rand(&rng);
        d3 k2, weight;
        double bpdf;
        const int event = sample64(&mat, N, Ng, k1, x0, x1, x2, &k2, &weight, &bpdf);
        if (!event) break;
        specularBounce = event >= 3;
        thr = dmul(thr, weight);
        if (depth > (int)rp->rrStartDepth) {
            double pr = thr.x > thr.y ? thr.x : thr.y;
            pr = pr > thr.z ? pr : thr.z;
            if (ref) {
                if (orc_rand(&rng) > pr) break;
                thr = dscl(thr, 1.0 / (pr + 1e-5));
            } else {
                pr = pr < 1.0 ? pr : 1.0;
                if (orc_rand(&rng) > pr || !(pr > 0.0)) break;
                thr = dscl(thr, 1.0 / pr);
            }
        }
        prevPdf = bpdf;
        ray.o = offset64(P, ddot(k2, Ng) > 0.0 ? Ng : dscl(Ng, -1.0));
        ray.d = k2;
        ray.tmax = 1e9;
    }
    out[0] = color.x;
    out[1] = color.y;
    out[2] = color.z;
}

#include <pthread.h>

typedef struct {
    const OrcScene* s;
    const StkRenderParams* rp;
    uint32_t x0, y0, w, h;
    double *sum, *sumsq;
    volatile uint32_t* next;
    pthread_mutex_t* mu;
} Fp64Arg;

static void* fp64_worker(void* pv)
{
    Fp64Arg* a = (Fp64Arg*)pv;
    const uint32_t npix = a->w * a->h, spp = (uint32_t)a->rp->pt.spp;
    for (;;) {
        pthread_mutex_lock(a->mu);
        const uint32_t start = *a->next;
        *a->next = start + 1;
        pthread_mutex_unlock(a->mu);
        if (start >= npix) break;
        const uint32_t px = a->x0 + start % a->w, py = a->y0 + start / a->w;
        double s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
        for (uint32_t k = 0; k < spp; ++k) {
            double c[3];
            orc_pt_sample_fp64(a->s, a->rp, px, py, k, c);
            for (int i = 0; i < 3; ++i) {
                s1[i] += c[i];
                s2[i] += c[i] * c[i];
            }
        }
        for (int i = 0; i < 3; ++i) {
            a->sum[start * 3 + i] = s1[i];
            a->sumsq[start * 3 + i] = s2[i];
        }
    }
    return NULL;
}

void orc_pt_frame_fp64(const OrcScene* s, const StkRenderParams* rp, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                       double* sum, double* sumsq, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    volatile uint32_t next = 0;
    pthread_mutex_t mu;
    pthread_mutex_init(&mu, NULL);
    Fp64Arg a = {s, rp, x0, y0, w, h, sum, sumsq, &next, &mu};
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, fp64_worker, &a);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
    pthread_mutex_destroy(&mu);
}
