// scenegen.cpp - seeded synthetic workloads (see include/stk/stk_scene.h).
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "../../include/stk/stk_scene.h"
#include "scene.hpp"

namespace stk {

// PCG32 XSH-RR (O'Neill 2014).
struct Pcg32 {
    uint64_t state = 0, inc = 1;
    Pcg32(uint64_t seed, uint64_t seq)
    {
        inc = (seq << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next()
    {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    }
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); } // [0,1)
    float uniform(float a, float b) { return a + (b - a) * uniform(); }
};

static constexpr uint64_t kSeedBase = 0x5354524B00000000ULL; // "STRK"

namespace {

// Value noise on a 256x256 wrapped lattice, smoothstep interpolation.
struct ValueNoise {
    std::vector<float> lat;
    explicit ValueNoise(Pcg32& rng) : lat(256 * 256)
    {
        for (float& v : lat) v = rng.uniform(-1.0f, 1.0f);
    }
    float at(int x, int y) const { return lat[(size_t)(y & 255) * 256 + (size_t)(x & 255)]; }
    float eval(float x, float y) const
    {
        const float fx = std::floor(x), fy = std::floor(y);
        const int ix = (int)fx, iy = (int)fy;
        float tx = x - fx, ty = y - fy;
        tx = tx * tx * (3.0f - 2.0f * tx);
        ty = ty * ty * (3.0f - 2.0f * ty);
        const float a = at(ix, iy), b = at(ix + 1, iy), c = at(ix, iy + 1), d = at(ix + 1, iy + 1);
        return (a + (b - a) * tx) + ((c + (d - c) * tx) - (a + (b - a) * tx)) * ty;
    }
    // fBm: 6 octaves, gain 0.5, lacunarity 2, normalised to [-1, 1]
    float fbm(float x, float y) const
    {
        float sum = 0.0f, amp = 1.0f, freq = 1.0f, norm = 0.0f;
        for (int o = 0; o < 6; ++o) {
            sum += amp * eval(x * freq + 17.0f * (float)o, y * freq + 31.0f * (float)o);
            norm += amp;
            amp *= 0.5f;
            freq *= 2.0f;
        }
        return sum / norm;
    }
};

void addHeightfield(Scene& sc, Pcg32& rng, uint32_t res, float extent, float amplitude, int32_t material)
{
    ValueNoise noise(rng);
    const float baseFreq = 0.2f;
    auto height = [&](float x, float z) { return amplitude * noise.fbm(x * baseFreq + 64.0f, z * baseFreq + 64.0f); };
    const uint32_t nv = res + 1;
    std::vector<StkVertex> vb((size_t)nv * nv);
    const float step = 2.0f * extent / (float)res;
    const float h = step * 0.5f;
    for (uint32_t j = 0; j < nv; ++j)
        for (uint32_t i = 0; i < nv; ++i) {
            const float x = -extent + step * (float)i, z = -extent + step * (float)j;
            StkVertex& v = vb[(size_t)j * nv + i];
            v = StkVertex{};
            v.position[0] = x;
            v.position[1] = height(x, z);
            v.position[2] = z;
            const float dx = (height(x + h, z) - height(x - h, z)) / (2.0f * h);
            const float dz = (height(x, z + h) - height(x, z - h)) / (2.0f * h);
            F3 n = mk3(-dx, 1.0f, -dz);
            const float l = std::sqrt(dot(n, n));
            n = mk3(n.x / l, n.y / l, n.z / l);
            v.normal = packNormal(n);
            v.tangent = packNormal(mk3(1.0f, 0.0f, 0.0f));
            v.uv = packUV((float)i / (float)res, (float)j / (float)res);
        }
    std::vector<uint32_t> ib;
    ib.reserve((size_t)res * res * 6);
    for (uint32_t j = 0; j < res; ++j)
        for (uint32_t i = 0; i < res; ++i) {
            const uint32_t a = j * nv + i, b = a + 1, c = a + nv, d = c + 1;
            const uint32_t t[6] = {a, c, b, b, c, d}; // geometric normals point up (+y)
            ib.insert(ib.end(), t, t + 6);
        }
    const uint32_t mesh = sc.createMesh(vb, ib);
    sc.createInstance(mesh, material, Mat4::identity());
}

StkMaterial randomMaterial(Pcg32& rng)
{
    StkMaterial m{};
    m.type = rng.uniform() < 0.5f ? STK_MAT_LAMBERT : STK_MAT_GGX;
    m.albedo[0] = rng.uniform(0.1f, 0.9f);
    m.albedo[1] = rng.uniform(0.1f, 0.9f);
    m.albedo[2] = rng.uniform(0.1f, 0.9f);
    m.albedo[3] = 1.0f;
    m.roughness = rng.uniform(0.05f, 1.0f);
    return m;
}

void addSpheres(Scene& sc, Pcg32& geo, Pcg32& matRng, uint32_t count, float extent, int subdiv)
{
    if (count == 0) return;
    const uint32_t mesh = sc.createIcosphereMesh(subdiv, false);
    for (uint32_t i = 0; i < count; ++i) {
        const float cx = geo.uniform(-extent, extent), cy = geo.uniform(0.5f, 4.0f), cz = geo.uniform(-extent, extent);
        const float r = geo.uniform(0.05f, 0.5f);
        const int32_t mat = (int32_t)sc.addMaterial(randomMaterial(matRng));
        sc.createInstance(mesh, mat, Mat4::translate(cx, cy, cz) * Mat4::scale(r, r, r));
    }
}

void addQuadLight(Scene& sc, float y, float size, float intensity)
{
    LightDesc d;
    d.type = 0;
    d.xform = Mat4::translate(0.0f, y, 0.0f) * Mat4::rotateX(-(float)M_PI / 2.0f); // emits towards -y
    d.width = size;
    d.height = size;
    d.intensity = intensity;
    sc.createLight(d);
}

void addBox(Scene& sc, float half, int32_t material, bool inward)
{
    // six faces, each two triangles, per-face vertices with inward (or outward) normals
    std::vector<StkVertex> vb;
    std::vector<uint32_t> ib;
    for (int axis = 0; axis < 3; ++axis)
        for (int side = -1; side <= 1; side += 2) {
            const int u = (axis + 1) % 3, v = (axis + 2) % 3;
            float corners[4][3];
            const float su[4] = {-1, 1, 1, -1}, sv[4] = {-1, -1, 1, 1};
            for (int k = 0; k < 4; ++k) {
                corners[k][axis] = half * (float)side;
                corners[k][u] = half * su[k];
                corners[k][v] = half * sv[k];
            }
            F3 n = mk3(0, 0, 0);
            (&n.x)[axis] = inward ? -(float)side : (float)side;
            const uint32_t base = (uint32_t)vb.size();
            for (int k = 0; k < 4; ++k) {
                StkVertex vx{};
                vx.position[0] = corners[k][0];
                vx.position[1] = corners[k][1];
                vx.position[2] = corners[k][2];
                vx.normal = packNormal(n);
                vx.tangent = packNormal(mk3(1.0f, 0.0f, 0.0f));
                vb.push_back(vx);
            }
            // orient the winding so the geometric normal agrees with n
            const F3 p0 = ld3(corners[0]), p1 = ld3(corners[1]), p2 = ld3(corners[2]);
            const bool flip = dot(cross(p1 - p0, p2 - p0), n) < 0.0f;
            const uint32_t t[6] = {0, 1, 2, 0, 2, 3};
            for (int k = 0; k < 6; k += 3) {
                ib.push_back(base + t[k]);
                ib.push_back(base + (flip ? t[k + 2] : t[k + 1]));
                ib.push_back(base + (flip ? t[k + 1] : t[k + 2]));
            }
        }
    const uint32_t mesh = sc.createMesh(vb, ib);
    sc.createInstance(mesh, material, Mat4::identity());
}

void addInwardBox(Scene& sc, float half, int32_t material) { addBox(sc, half, material, true); }

// Torus around the y axis: major radius R, minor radius r, U x V quads.
void makeTorus(uint32_t U, uint32_t V, float R, float r, std::vector<StkVertex>& vb, std::vector<uint32_t>& ib)
{
    vb.clear();
    ib.clear();
    for (uint32_t i = 0; i < U; ++i)
        for (uint32_t j = 0; j < V; ++j) {
            const float a = 2.0f * (float)M_PI * (float)i / (float)U, b = 2.0f * (float)M_PI * (float)j / (float)V;
            const F3 n = mk3(std::cos(a) * std::cos(b), std::sin(b), std::sin(a) * std::cos(b));
            StkVertex v{};
            v.position[0] = (R + r * std::cos(b)) * std::cos(a);
            v.position[1] = r * std::sin(b);
            v.position[2] = (R + r * std::cos(b)) * std::sin(a);
            v.normal = packNormal(n);
            v.tangent = packNormal(mk3(1.0f, 0.0f, 0.0f));
            vb.push_back(v);
        }
    for (uint32_t i = 0; i < U; ++i)
        for (uint32_t j = 0; j < V; ++j) {
            const uint32_t a = i * V + j, b = ((i + 1) % U) * V + j, c = i * V + (j + 1) % V,
                           d = ((i + 1) % U) * V + (j + 1) % V;
            const uint32_t t[6] = {a, c, b, b, c, d};
            ib.insert(ib.end(), t, t + 6);
        }
}

// Uniform random rotation (Shoemake 1992, uniform unit quaternion) then a
// translation uniform in [-extent, extent]^3.
Mat4 randomRigid(Pcg32& rng, float extent)
{
    const float u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
    const float s1 = std::sqrt(1.0f - u1), s2 = std::sqrt(u1);
    const float x = s1 * std::sin(2.0f * (float)M_PI * u2), y = s1 * std::cos(2.0f * (float)M_PI * u2);
    const float z = s2 * std::sin(2.0f * (float)M_PI * u3), w = s2 * std::cos(2.0f * (float)M_PI * u3);
    Mat4 r = Mat4::identity(); // column-major rotation of the quaternion (x, y, z, w)
    r.m[0] = 1 - 2 * (y * y + z * z);
    r.m[1] = 2 * (x * y + z * w);
    r.m[2] = 2 * (x * z - y * w);
    r.m[4] = 2 * (x * y - z * w);
    r.m[5] = 1 - 2 * (x * x + z * z);
    r.m[6] = 2 * (y * z + x * w);
    r.m[8] = 2 * (x * z + y * w);
    r.m[9] = 2 * (y * z - x * w);
    r.m[10] = 1 - 2 * (x * x + y * y);
    const float tx = rng.uniform(-extent, extent), ty = rng.uniform(-extent, extent), tz = rng.uniform(-extent, extent);
    return Mat4::translate(tx, ty, tz) * r;
}

// Two-triangle floor y = 0 over [-extent, extent]^2, normal +y.
void addFloor(Scene& sc, float extent, int32_t material)
{
    std::vector<StkVertex> vb(4);
    const float xs[4] = {-extent, extent, extent, -extent}, zs[4] = {-extent, -extent, extent, extent};
    for (int k = 0; k < 4; ++k) {
        vb[k] = StkVertex{};
        vb[k].position[0] = xs[k];
        vb[k].position[2] = zs[k];
        vb[k].normal = packNormal(mk3(0.0f, 1.0f, 0.0f));
        vb[k].tangent = packNormal(mk3(1.0f, 0.0f, 0.0f));
    }
    const uint32_t mesh = sc.createMesh(vb, {0, 2, 1, 0, 3, 2}); // (p_b - p_a) x (p_c - p_a) points to +y
    sc.createInstance(mesh, material, Mat4::identity());
}

} // namespace
} // namespace stk

struct stk_scene {
    stk::Scene scene;
    // instance records (transforms, their inverses, index ranges), rebuilt by
    // stk_scene_get after the instances or lights change
    mutable std::vector<StkInstance> inst;
    mutable bool instDirty = false;
    std::vector<StkPathTracerParam> cameras; // Scene::mCameras; camera 0 is the generator's
    float sky[4] = {0, 0, 0, 0};
    uint32_t config = 0;                     // 0 for a scene built through the calls below

    const std::vector<StkInstance>& instances() const
    {
        if (instDirty) {
            inst = scene.getInstanceConstants();
            instDirty = false;
        }
        return inst;
    }
};

using namespace stk;

extern "C" STK_API stk_status stk_scene_generate(uint32_t config, uint32_t variant, uint32_t flags, stk_scene** out)
{
    if (!out) return STK_E_INVALID;
    *out = nullptr;
    std::unique_ptr<stk_scene> s(new (std::nothrow) stk_scene());
    if (!s) return STK_E_OOM;
    s->config = config;
    Scene& sc = s->scene;
    const bool refIco = (flags & STK_SCENE_REF_ICOSPHERE) != 0;
    StkPathTracerParam cam{};
    try {
        if (config == 1 || config == 7) {
            StkMaterial sphere{};
            sphere.type = STK_MAT_LAMBERT;
            sphere.albedo[0] = sphere.albedo[1] = sphere.albedo[2] = 0.8f;
            StkMaterial wall = sphere;
            const float wa = config == 7 ? 0.7f : 0.5f;
            wall.albedo[0] = wall.albedo[1] = wall.albedo[2] = wa;
            const int32_t ms = (int32_t)sc.addMaterial(sphere), mw = (int32_t)sc.addMaterial(wall);
            const uint32_t mesh = sc.createIcosphereMesh(3, refIco);
            sc.createInstance(mesh, ms, Mat4::identity());
            addInwardBox(sc, 2.0f, mw);
            const uint32_t res = config == 7 ? 128 : 256;
            if (config == 7) addQuadLight(sc, 1.9f, 1.0f, 20.0f); // closed room lit from the ceiling
            cameraMatrices(mk3(0, 0, 1.9f), mk3(0, 0, 0), mk3(0, 1, 0), 60.0f, res, res, 0.01f, 100.0f, cam);
        } else if (config == 2 || config == 3 || config == 6) {
            uint32_t hfRes = 512, spheres = 2000, width = 1920, height = 1080;
            float extent = 10.0f, lightY = 8.0f, lightSize = 2.0f, intensity = 100.0f;
            uint64_t stream = 0;
            if (config == 3) {
                if (variant >= 32) return STK_E_INVALID;
                Pcg32 pick(kSeedBase + 3, 100 + variant);
                const double lnN = std::log(1e4) + (std::log(2e6) - std::log(1e4)) * (double)pick.uniform();
                const double N = std::round(std::exp(lnN));
                hfRes = (uint32_t)std::max(2.0, std::round(std::sqrt(0.45 * N / 2.0)));
                spheres = (uint32_t)std::max(1.0, std::round(0.55 * N / 320.0));
                width = 960;
                height = 540;
                stream = 16 * (uint64_t)(variant + 1);
            } else if (config == 6) {
                hfRes = 16;
                spheres = 8;
                width = 64;
                height = 48;
                extent = 4.0f;
                lightY = 3.0f;
                lightSize = 1.0f;
                intensity = 10.0f;
            }
            Pcg32 geo(kSeedBase + config, 1 + stream), mat(kSeedBase + config, 2 + stream);
            StkMaterial ground{};
            ground.type = STK_MAT_LAMBERT;
            ground.albedo[0] = 0.55f;
            ground.albedo[1] = 0.5f;
            ground.albedo[2] = 0.45f;
            ground.roughness = 1.0f;
            const int32_t mg = (int32_t)sc.addMaterial(ground);
            addHeightfield(sc, geo, hfRes, extent, 1.5f, mg);
            addSpheres(sc, geo, mat, spheres, extent, 2);
            addQuadLight(sc, lightY, lightSize, intensity);
            if (config == 6)
                cameraMatrices(mk3(0, 3.0f, 6.0f), mk3(0, 0.3f, 0), mk3(0, 1, 0), 55.0f, width, height, 0.01f, 100.0f,
                               cam);
            else
                // 60 degrees below the horizon; the whole frame falls on the heightfield
                cameraMatrices(mk3(0, 17.0f, 9.5f), mk3(0, 0.0f, -0.3f), mk3(0, 1, 0), 24.0f, width, height, 0.01f,
                               1000.0f, cam);
            s->sky[0] = 0.5f;
            s->sky[1] = 0.6f;
            s->sky[2] = 0.8f;
        } else if (config == 5) {
            // 20 seeded meshes (10 displaced icospheres of 1,280 triangles, 10 tori
            // of 500-5,000), 10,000 instances with random rigid transforms in
            // [-100, 100]^3, one material per instance, a large rect light above.
            // variant n > 0: n instances in a box scaled to the same density,
            // camera and light scaled alike, a 384x216 view
            Pcg32 geo(kSeedBase + 5, 1), mat(kSeedBase + 5, 2);
            std::vector<uint32_t> meshes;
            for (int k = 0; k < 20; ++k) {
                std::vector<StkVertex> vb;
                std::vector<uint32_t> ib;
                if (k % 2 == 0) {
                    makeIcosphere(3, false, vb, ib);
                    const float radius = geo.uniform(1.0f, 3.0f);
                    for (StkVertex& v : vb) {
                        const F3 p = ld3(v.position);
                        const float d = radius * (1.0f + 0.15f * geo.uniform(-1.0f, 1.0f));
                        const float l = std::sqrt(dot(p, p));
                        v.position[0] = p.x / l * d;
                        v.position[1] = p.y / l * d;
                        v.position[2] = p.z / l * d;
                    }
                } else {
                    const float target = geo.uniform(500.0f, 5000.0f);
                    const uint32_t U = std::max(4u, (uint32_t)std::lround(std::sqrt(target)));
                    const uint32_t V = std::max(3u, (uint32_t)std::lround(target / (2.0f * (float)U)));
                    const float R = geo.uniform(1.0f, 3.0f);
                    makeTorus(U, V, R, R * geo.uniform(0.2f, 0.4f), vb, ib);
                }
                meshes.push_back(sc.createMesh(vb, ib));
            }
            const uint32_t count = variant ? variant : 10000u;
            const float k3 = std::cbrt((float)count / 10000.0f);
            for (uint32_t k = 0; k < count; ++k) {
                const uint32_t m = (uint32_t)(geo.next() % 20u);
                const int32_t mi = (int32_t)sc.addMaterial(randomMaterial(mat));
                sc.createInstance(meshes[m], mi, randomRigid(geo, 100.0f * k3));
            }
            LightDesc d;
            d.type = 0;
            d.xform = Mat4::translate(0.0f, 130.0f * k3, 0.0f) * Mat4::rotateX(-(float)M_PI / 2.0f);
            d.width = 80.0f * k3;
            d.height = 80.0f * k3;
            d.intensity = 8.0f * d.width * d.height; // radiance 8
            sc.createLight(d);
            cameraMatrices(mk3(0.0f, 30.0f * k3, 190.0f * k3), mk3(0, 0, 0), mk3(0, 1, 0), 55.0f, variant ? 384 : 3840,
                           variant ? 216 : 2160, 0.01f, 1000.0f, cam);
            s->sky[0] = 0.25f;
            s->sky[1] = 0.3f;
            s->sky[2] = 0.4f;
        } else if (config == 8) {
            // convex furnace: an outward-facing Lambert cube under a constant sky, no lights
            StkMaterial m{};
            m.type = STK_MAT_LAMBERT;
            m.roughness = 1.0f;
            m.albedo[0] = 0.8f;
            m.albedo[1] = 0.6f;
            m.albedo[2] = 0.4f;
            m.albedo[3] = 1.0f;
            addBox(sc, 1.0f, (int32_t)sc.addMaterial(m), false);
            cameraMatrices(mk3(3.0f, 2.4f, 3.6f), mk3(0, 0, 0), mk3(0, 1, 0), 40.0f, 96, 96, 0.01f, 100.0f,
                           cam);
            s->sky[0] = 0.5f;
            s->sky[1] = 0.7f;
            s->sky[2] = 1.0f;
        } else if (config == 9) {
            // lit floor: Lambert floor (albedo 0.5) under a 1x1 rect light at y = 2 facing down
            StkMaterial m{};
            m.type = STK_MAT_LAMBERT;
            m.roughness = 1.0f;
            m.albedo[0] = m.albedo[1] = m.albedo[2] = 0.5f;
            m.albedo[3] = 1.0f;
            addFloor(sc, 20.0f, (int32_t)sc.addMaterial(m));
            addQuadLight(sc, 2.0f, 1.0f, 10.0f);
            cameraMatrices(mk3(0, 1.0f, 0), mk3(0, 0, 0), mk3(0, 0, -1), 60.0f, 64, 64, 0.01f, 100.0f, cam);
        } else {
            return STK_E_INVALID;
        }
        s->cameras.push_back(cam);
        s->inst = sc.getInstanceConstants();
    } catch (const std::bad_alloc&) {
        return STK_E_OOM;
    }
    *out = s.release();
    return STK_OK;
}

extern "C" STK_API void stk_scene_free(stk_scene* s) { delete s; }

extern "C" STK_API stk_status stk_scene_get(const stk_scene* s, StkSceneArrays* out)
{
    if (!s || !out) return STK_E_INVALID;
    const Scene& sc = s->scene;
    try {
        s->instances();
    } catch (const std::bad_alloc&) {
        return STK_E_OOM;
    }
    out->vb = sc.getVertices().data();
    out->numVertices = (uint32_t)sc.getVertices().size();
    out->ib = sc.getIndices().data();
    out->numIndices = (uint32_t)sc.getIndices().size();
    out->inst = s->inst.data();
    out->numInstances = (uint32_t)s->inst.size();
    out->lights = sc.getLights().data();
    out->numLights = (uint32_t)sc.getLights().size();
    out->mats = sc.getMaterials().data();
    out->numMaterials = (uint32_t)sc.getMaterials().size();
    out->numTriangles = out->numIndices / 3;
    out->pad = 0;
    if (s->cameras.empty())
        std::memset(&out->camera, 0, sizeof(out->camera));
    else
        out->camera = s->cameras[0];
    for (int i = 0; i < 4; ++i) out->sky[i] = s->sky[i];
    return STK_OK;
}

extern "C" STK_API stk_status stk_scene_jittered_rays(const stk_scene* s, uint64_t seed, StkRay* rays, uint32_t n)
{
    if (!s || !rays || s->cameras.empty()) return STK_E_INVALID;
    const StkPathTracerParam& p = s->cameras[0];
    const uint32_t W = (uint32_t)p.dimension[0], H = (uint32_t)p.dimension[1];
    if ((uint64_t)W * H != n) return STK_E_INVALID;
    Pcg32 rng(seed, 3);
    const F3 eye = xform4(p.viewToWorld, mk3(0, 0, 0), 1.0f);
    for (uint32_t py = 0; py < H; ++py)
        for (uint32_t px = 0; px < W; ++px) {
            const float jx = rng.uniform(), jy = rng.uniform();
            const float ndcx = ((float)px + jx) / (float)W * 2.0f - 1.0f;
            const float ndcy = ((float)py + jy) / (float)H * 2.0f - 1.0f;
            const float* C = p.clipToView;
            const F3 v = mk3(C[0] * ndcx + C[4] * ndcy + C[8] + C[12], C[1] * ndcx + C[5] * ndcy + C[9] + C[13],
                             C[2] * ndcx + C[6] * ndcy + C[10] + C[14]);
            const F3 w = xform4(p.viewToWorld, v, 0.0f);
            const double l = std::sqrt((double)w.x * w.x + (double)w.y * w.y + (double)w.z * w.z);
            StkRay& r = rays[(size_t)py * W + px];
            r.o[0] = eye.x + rng.uniform(-1e-3f, 1e-3f);
            r.o[1] = eye.y + rng.uniform(-1e-3f, 1e-3f);
            r.o[2] = eye.z + rng.uniform(-1e-3f, 1e-3f);
            r.o[3] = 1e9f;
            r.d[0] = (float)(w.x / l);
            r.d[1] = (float)(w.y / l);
            r.d[2] = (float)(w.z / l);
            r.d[3] = 0.0f;
        }
    return STK_OK;
}

// ---------------------------------------------------------------------------
// Scene building through the C ABI: the calls oka::Scene offers its producers
// (include/scene/scene.h:231-389, src/scene/scene.cpp), over the same container
// the generators fill.  Each returns STK_E_INVALID for a null or out-of-range
// argument and STK_E_OOM when the host allocation fails; ids count from 0 in
// creation order.

namespace {

Mat4 loadMat4(const float* m)
{
    Mat4 r;
    std::memcpy(r.m, m, sizeof(r.m));
    return r;
}

bool finite16(const float* m)
{
    for (int i = 0; i < 16; ++i)
        if (!std::isfinite(m[i])) return false;
    return true;
}

bool toLightDesc(const StkLightDesc* d, LightDesc& out)
{
    if (!d || d->type < 0 || d->type > 2 || !finite16(d->xform)) return false;
    if (d->type == 0 && !(d->width > 0.0f && d->height > 0.0f)) return false;
    if (d->type != 0 && !(d->radius > 0.0f)) return false;
    if (!std::isfinite(d->intensity)) return false;
    out.type = d->type;
    out.xform = loadMat4(d->xform);
    for (int k = 0; k < 3; ++k) out.color[k] = d->color[k];
    out.intensity = d->intensity;
    out.width = d->width;
    out.height = d->height;
    out.radius = d->radius;
    return true;
}

} // namespace

extern "C" STK_API stk_status stk_scene_create(stk_scene** out)
{
    if (!out) return STK_E_INVALID;
    *out = new (std::nothrow) stk_scene();
    return *out ? STK_OK : STK_E_OOM;
}

extern "C" STK_API stk_status stk_scene_create_mesh(stk_scene* s, const StkVertex* vb, uint32_t numVertices,
                                                    const uint32_t* ib, uint32_t numIndices, uint32_t* meshId)
{
    if (!s || !meshId || !numVertices || !vb || !numIndices || !ib || numIndices % 3) return STK_E_INVALID;
    for (uint32_t i = 0; i < numIndices; ++i)
        if (ib[i] >= numVertices) return STK_E_INVALID;
    const uint64_t total = (uint64_t)s->scene.getVertices().size() + numVertices;
    if (total > 0xFFFFFFFFull || (uint64_t)s->scene.getIndices().size() + numIndices > 0x7FFFFFFFull)
        return STK_E_INVALID;
    try {
        *meshId = s->scene.createMesh(std::vector<StkVertex>(vb, vb + numVertices),
                                      std::vector<uint32_t>(ib, ib + numIndices));
    } catch (const std::bad_alloc&) {
        return STK_E_OOM;
    }
    return STK_OK;
}

extern "C" STK_API stk_status stk_scene_create_instance(stk_scene* s, uint32_t meshId, int32_t materialId,
                                                        const float transform[16], uint32_t* instId)
{
    if (!s || !instId || !transform || meshId >= s->scene.getMeshes().size() || materialId < -1 ||
        !finite16(transform))
        return STK_E_INVALID;
    try {
        *instId = s->scene.createInstance(meshId, materialId, loadMat4(transform));
    } catch (const std::bad_alloc&) {
        return STK_E_OOM;
    }
    s->instDirty = true;
    return STK_OK;
}

extern "C" STK_API stk_status stk_scene_update_instance_transform(stk_scene* s, uint32_t instId,
                                                                  const float transform[16])
{
    if (!s || !transform || !finite16(transform)) return STK_E_INVALID;
    if (!s->scene.updateInstanceTransform(instId, loadMat4(transform))) return STK_E_INVALID;
    s->instDirty = true;
    return STK_OK;
}

extern "C" STK_API stk_status stk_scene_add_material(stk_scene* s, const StkMaterial* m, uint32_t* materialId)
{
    if (!s || !m || !materialId || m->type > STK_MAT_GLASS || !(m->ior >= 0.0f)) return STK_E_INVALID;
    try {
        *materialId = s->scene.addMaterial(*m);
    } catch (const std::bad_alloc&) {
        return STK_E_OOM;
    }
    return STK_OK;
}

extern "C" STK_API stk_status stk_scene_create_light(stk_scene* s, const StkLightDesc* desc, uint32_t* lightId)
{
    LightDesc d;
    if (!s || !lightId || !toLightDesc(desc, d)) return STK_E_INVALID;
    try {
        *lightId = s->scene.createLight(d);
    } catch (const std::bad_alloc&) {
        return STK_E_OOM;
    }
    s->instDirty = true;
    return STK_OK;
}

extern "C" STK_API stk_status stk_scene_update_light(stk_scene* s, uint32_t lightId, const StkLightDesc* desc)
{
    LightDesc d;
    if (!s || !toLightDesc(desc, d) || !s->scene.updateLight(lightId, d)) return STK_E_INVALID;
    s->instDirty = true;
    return STK_OK;
}

extern "C" STK_API stk_status stk_scene_light_instance(const stk_scene* s, uint32_t lightId, uint32_t* instId)
{
    if (!s || !instId || lightId >= s->scene.getLights().size()) return STK_E_INVALID;
    *instId = s->scene.lightInstance(lightId);
    return STK_OK;
}

extern "C" STK_API stk_status stk_scene_add_camera(stk_scene* s, const StkPathTracerParam* camera, uint32_t* cameraId)
{
    if (!s || !camera || !cameraId || camera->dimension[0] <= 0 || camera->dimension[1] <= 0) return STK_E_INVALID;
    try {
        s->cameras.push_back(*camera);
    } catch (const std::bad_alloc&) {
        return STK_E_OOM;
    }
    *cameraId = (uint32_t)s->cameras.size() - 1;
    return STK_OK;
}

extern "C" STK_API stk_status stk_scene_update_camera(stk_scene* s, uint32_t cameraId,
                                                      const StkPathTracerParam* camera)
{
    if (!s || !camera || cameraId >= s->cameras.size() || camera->dimension[0] <= 0 || camera->dimension[1] <= 0)
        return STK_E_INVALID;
    s->cameras[cameraId] = *camera;
    return STK_OK;
}

extern "C" STK_API uint32_t stk_scene_camera_count(const stk_scene* s) { return s ? (uint32_t)s->cameras.size() : 0; }

extern "C" STK_API stk_status stk_scene_get_camera(const stk_scene* s, uint32_t cameraId, StkPathTracerParam* out)
{
    if (!s || !out || cameraId >= s->cameras.size()) return STK_E_INVALID;
    *out = s->cameras[cameraId];
    return STK_OK;
}

extern "C" STK_API stk_status stk_scene_set_sky(stk_scene* s, float r, float g, float b)
{
    if (!s || !std::isfinite(r) || !std::isfinite(g) || !std::isfinite(b)) return STK_E_INVALID;
    s->sky[0] = r;
    s->sky[1] = g;
    s->sky[2] = b;
    s->sky[3] = 0.0f;
    return STK_OK;
}

extern "C" STK_API stk_status stk_camera_look_at(const float eye[3], const float target[3], const float up[3],
                                                 float vfovDeg, uint32_t width, uint32_t height, float znear,
                                                 float zfar, StkPathTracerParam* out)
{
    if (!eye || !target || !up || !out || !width || !height || !(vfovDeg > 0.0f && vfovDeg < 180.0f) ||
        !(znear > 0.0f && zfar > znear))
        return STK_E_INVALID;
    std::memset(out, 0, sizeof(*out));
    cameraMatrices(ld3(eye), ld3(target), ld3(up), vfovDeg, width, height, znear, zfar, *out);
    return STK_OK;
}

extern "C" STK_API uint32_t stk_pack_normal(float x, float y, float z) { return packNormal(mk3(x, y, z)); }
extern "C" STK_API uint32_t stk_pack_uv(float u, float v) { return packUV(u, v); }
