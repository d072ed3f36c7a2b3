// scene.cpp - see scene.hpp.
#include "scene.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <utility>

namespace stk {

Mat4 Mat4::identity()
{
    Mat4 r{};
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    return r;
}
Mat4 Mat4::translate(float x, float y, float z)
{
    Mat4 r = identity();
    r.m[12] = x;
    r.m[13] = y;
    r.m[14] = z;
    return r;
}
Mat4 Mat4::scale(float x, float y, float z)
{
    Mat4 r = identity();
    r.m[0] = x;
    r.m[5] = y;
    r.m[10] = z;
    return r;
}
Mat4 Mat4::rotateX(float a)
{
    Mat4 r = identity();
    const float c = std::cos(a), s = std::sin(a);
    r.m[5] = c;
    r.m[6] = s;
    r.m[9] = -s;
    r.m[10] = c;
    return r;
}
Mat4 Mat4::operator*(const Mat4& o) const
{
    Mat4 r{};
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row) {
            float s = 0.0f;
            for (int k = 0; k < 4; ++k) s += m[k * 4 + row] * o.m[c * 4 + k];
            r.m[c * 4 + row] = s;
        }
    return r;
}

Mat4 transpose(const Mat4& a)
{
    Mat4 r{};
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row) r.m[c * 4 + row] = a.m[row * 4 + c];
    return r;
}

Mat4 inverse(const Mat4& a)
{
    double m[4][8];
    for (int r = 0; r < 4; ++r) {
        for (int c = 0; c < 4; ++c) m[r][c] = a.m[c * 4 + r];
        for (int c = 0; c < 4; ++c) m[r][4 + c] = (r == c) ? 1.0 : 0.0;
    }
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int r = col + 1; r < 4; ++r)
            if (std::fabs(m[r][col]) > std::fabs(m[piv][col])) piv = r;
        for (int c = 0; c < 8; ++c) std::swap(m[col][c], m[piv][c]);
        const double d = m[col][col];
        if (d == 0.0) return Mat4::identity();
        for (int c = 0; c < 8; ++c) m[col][c] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == col) continue;
            const double f = m[r][col];
            for (int c = 0; c < 8; ++c) m[r][c] -= f * m[col][c];
        }
    }
    Mat4 out{};
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out.m[c * 4 + r] = (float)m[r][4 + c];
    return out;
}

uint32_t packNormal(F3 n)
{
    uint32_t packed = (uint32_t)((n.x + 1.0f) / 2.0f * 511.99999f);
    packed += (uint32_t)((n.y + 1.0f) / 2.0f * 511.99999f) << 10;
    packed += (uint32_t)((n.z + 1.0f) / 2.0f * 511.99999f) << 20;
    return packed;
}

uint32_t packUV(float u, float v)
{
    uint32_t packed = (uint32_t)((u + 10.0f) / 20.0f * 16383.99999f);
    packed += (uint32_t)((v + 10.0f) / 20.0f * 16383.99999f) << 16;
    return packed;
}

uint32_t Scene::createMesh(const std::vector<StkVertex>& vb, const std::vector<uint32_t>& ib)
{
    Mesh mesh;
    mesh.index = (uint32_t)indices_.size();
    mesh.count = (uint32_t)ib.size();
    const uint32_t base = (uint32_t)vertices_.size();
    for (uint32_t i : ib) indices_.push_back(base + i);
    vertices_.insert(vertices_.end(), vb.begin(), vb.end());
    meshes_.push_back(mesh);
    return (uint32_t)meshes_.size() - 1;
}

uint32_t Scene::createInstance(uint32_t meshId, int32_t materialId, const Mat4& transform, int32_t lightId)
{
    Instance inst;
    inst.meshId = meshId;
    inst.materialId = materialId;
    inst.transform = transform;
    inst.lightId = lightId;
    instances_.push_back(inst);
    return (uint32_t)instances_.size() - 1;
}

uint32_t Scene::addMaterial(const StkMaterial& m)
{
    materials_.push_back(m);
    return (uint32_t)materials_.size() - 1;
}

uint32_t Scene::createRectLightMesh()
{
    if (rectLightMesh_ >= 0) return (uint32_t)rectLightMesh_;
    std::vector<StkVertex> vb(4);
    const float pos[4][3] = {{0.5f, 0.5f, 0.0f}, {-0.5f, 0.5f, 0.0f}, {-0.5f, -0.5f, 0.0f}, {0.5f, -0.5f, 0.0f}};
    for (int i = 0; i < 4; ++i) {
        vb[i] = StkVertex{};
        vb[i].position[0] = pos[i][0];
        vb[i].position[1] = pos[i][1];
        vb[i].position[2] = pos[i][2];
        vb[i].normal = packNormal(mk3(1.0f, 0.0f, 0.0f));
    }
    rectLightMesh_ = (int32_t)createMesh(vb, {0, 1, 2, 2, 3, 0});
    return (uint32_t)rectLightMesh_;
}

uint32_t Scene::createDiscLightMesh()
{
    // a 16-segment fan of unit radius in the z = 0 plane (src/scene/scene.cpp:318-360)
    std::vector<StkVertex> vb;
    std::vector<uint32_t> ib;
    StkVertex c{};
    c.normal = packNormal(mk3(0.0f, 0.0f, 1.0f));
    vb.push_back(c);
    const float step = 2.0f * (float)M_PI / 16;
    for (int i = 0; i <= 16; ++i) {
        StkVertex v = c;
        v.position[0] = std::cos(step * (float)i);
        v.position[1] = std::sin(step * (float)i);
        vb.push_back(v);
    }
    for (uint32_t i = 0; i < 16; ++i) {
        ib.push_back(0);
        ib.push_back(i + 1);
        ib.push_back(i + 2);
    }
    return createMesh(vb, ib);
}

uint32_t Scene::createIcosphereMesh(int subdivisions, bool refCompat)
{
    std::vector<StkVertex> vb;
    std::vector<uint32_t> ib;
    makeIcosphere(subdivisions, refCompat, vb, ib);
    return createMesh(vb, ib);
}

// The GPU light record of a description and the transform of its emitting
// mesh (Scene::updateLight, src/scene/scene.cpp:452-511; the mesh transform of
// Scene::createLight, src/scene/scene.cpp:418-450).
static StkLight lightRecord(const LightDesc& desc, Mat4& meshXform)
{
    StkLight l{};
    float intensityPerPoint = desc.intensity;
    if (desc.type == 0) {
        const Mat4 local = desc.xform * Mat4::scale(desc.width, desc.height, 1.0f);
        const float corners[4][3] = {{0.5f, 0.5f, 0.0f}, {-0.5f, 0.5f, 0.0f}, {-0.5f, -0.5f, 0.0f}, {0.5f, -0.5f, 0.0f}};
        for (int i = 0; i < 4; ++i) {
            const F3 p = xform4(local.m, ld3(corners[i]), 1.0f);
            l.points[i][0] = p.x;
            l.points[i][1] = p.y;
            l.points[i][2] = p.z;
            l.points[i][3] = 1.0f;
        }
        l.type = 0;
        intensityPerPoint /= (desc.width * desc.height);
        meshXform = local;
    } else if (desc.type == 1) {
        const Mat4 local = desc.xform * Mat4::scale(desc.radius, desc.radius, desc.radius);
        l.points[0][0] = desc.radius;
        const F3 o = xform4(local.m, mk3(0, 0, 0), 1.0f);
        const F3 ox = xform4(local.m, mk3(1, 0, 0), 0.0f);
        const F3 oy = xform4(local.m, mk3(0, 1, 0), 0.0f);
        const F3 nz = xform4(local.m, mk3(0, 0, 1), 0.0f);
        const F3 v[3] = {o, ox, oy};
        for (int i = 0; i < 3; ++i) {
            l.points[i + 1][0] = v[i].x;
            l.points[i + 1][1] = v[i].y;
            l.points[i + 1][2] = v[i].z;
        }
        l.normal[0] = nz.x;
        l.normal[1] = nz.y;
        l.normal[2] = nz.z;
        l.type = 1;
        intensityPerPoint /= (float)(M_PI * desc.radius * desc.radius);
        meshXform = local;
    } else {
        l.points[0][0] = desc.radius;
        const F3 o = xform4(desc.xform.m, mk3(0, 0, 0), 1.0f);
        l.points[1][0] = o.x;
        l.points[1][1] = o.y;
        l.points[1][2] = o.z;
        l.points[1][3] = 1.0f;
        l.type = 2;
        intensityPerPoint /= (float)(4.0 * M_PI * desc.radius * desc.radius);
        meshXform = desc.xform * Mat4::scale(desc.radius, desc.radius, desc.radius);
    }
    for (int c = 0; c < 3; ++c) l.color[c] = desc.color[c] * intensityPerPoint;
    l.color[3] = 1.0f * intensityPerPoint;
    return l;
}

uint32_t Scene::createLight(const LightDesc& desc)
{
    const uint32_t lightId = (uint32_t)lights_.size();
    Mat4 meshXform = desc.xform;
    const StkLight l = lightRecord(desc, meshXform);
    const uint32_t meshId = desc.type == 0   ? createRectLightMesh()
                            : desc.type == 1 ? createDiscLightMesh()
                                             : createIcosphereMesh(3, false);
    lights_.push_back(l);
    lightDescs_.push_back(desc);
    lightInstance_.push_back(createInstance(meshId, -1, meshXform, (int32_t)lightId));
    return lightId;
}

bool Scene::updateLight(uint32_t lightId, const LightDesc& desc)
{
    if (lightId >= lights_.size() || desc.type != lightDescs_[lightId].type) return false;
    Mat4 meshXform = desc.xform;
    lights_[lightId] = lightRecord(desc, meshXform);
    lightDescs_[lightId] = desc;
    instances_[lightInstance_[lightId]].transform = meshXform;
    return true;
}

bool Scene::updateInstanceTransform(uint32_t instId, const Mat4& transform)
{
    if (instId >= instances_.size()) return false;
    instances_[instId].transform = transform;
    return true;
}

std::vector<StkInstance> Scene::getInstanceConstants() const
{
    std::vector<StkInstance> out(instances_.size());
    for (size_t i = 0; i < instances_.size(); ++i) {
        const Instance& in = instances_[i];
        StkInstance& ic = out[i];
        memcpy(ic.objectToWorld, in.transform.m, sizeof(ic.objectToWorld));
        const Mat4 inv = inverse(in.transform);
        memcpy(ic.worldToObject, inv.m, sizeof(ic.worldToObject));
        const Mat4 nm = inverse(transpose(in.transform));
        memcpy(ic.normalMatrix, nm.m, sizeof(ic.normalMatrix));
        ic.materialId = in.materialId;
        ic.indexOffset = (int32_t)meshes_[in.meshId].index;
        ic.indexCount = (int32_t)meshes_[in.meshId].count;
        ic.lightId = in.lightId;
    }
    return out;
}

namespace {
using Lookup = std::map<std::pair<uint32_t, uint32_t>, uint32_t>;

// vertexForEdge (src/scene/scene.cpp:142-163)
uint32_t vertexForEdge(Lookup& lookup, std::vector<F3>& v, uint32_t first, uint32_t second)
{
    std::pair<uint32_t, uint32_t> key(first, second);
    if (key.first > key.second) std::swap(key.first, key.second);
    auto ins = lookup.insert({key, (uint32_t)v.size()});
    if (ins.second) {
        const F3 a = v[first], b = v[second];
        F3 p = mk3((a.x + b.x) / 2, (a.y + b.y) / 2, (a.z + b.z) / 2);
        const float len = std::sqrt(dot(p, p));
        p = mk3(p.x / len, p.y / len, p.z / len);
        v.push_back(p);
    }
    return ins.first->second;
}

// subdivide (src/scene/scene.cpp:165-195)
std::vector<uint32_t> subdivide(std::vector<F3>& v, const std::vector<uint32_t>& idx)
{
    Lookup lookup;
    std::vector<uint32_t> res;
    for (size_t i = 0; i < idx.size(); i += 3) {
        std::array<uint32_t, 3> mid;
        for (int e = 0; e < 3; ++e) mid[e] = vertexForEdge(lookup, v, idx[i + e], idx[i + (e + 1) % 3]);
        const uint32_t add[12] = {idx[i], mid[0], mid[2], idx[i + 1], mid[1], mid[0],
                                  idx[i + 2], mid[2], mid[1], mid[0], mid[1], mid[2]};
        res.insert(res.end(), add, add + 12);
    }
    return res;
}
} // namespace

void makeIcosphere(int subdivisions, bool refCompat, std::vector<StkVertex>& vb, std::vector<uint32_t>& ib)
{
    // Base icosahedron (src/scene/scene.cpp:219-304).
    std::vector<F3> v(12);
    std::vector<uint32_t> idx;
    const float PI = std::acos(-1.0f);
    const float H = PI / 180 * 72;
    const float V = std::atan(1.0f / 2);
    float h1 = -PI / 2 - H / 2, h2 = -PI / 2;
    const float radius = 1.0f;
    v[0] = mk3(0, 0, radius);
    for (int i = 1; i <= 5; ++i) {
        const int i1 = i, i2 = i + 5;
        const float z = radius * std::sin(V), xy = radius * std::cos(V);
        v[i1] = mk3(xy * std::cos(h1), xy * std::sin(h1), z);
        v[i2] = mk3(xy * std::cos(h2), xy * std::sin(h2), -z);
        const int n1 = (i < 5) ? i1 + 1 : 1;  // next vertex on the upper ring
        const int n2 = (i < 5) ? i1 + 6 : 6;  // next vertex on the lower ring
        const uint32_t f[12] = {0u, (uint32_t)i1, (uint32_t)n1,
                                (uint32_t)i1, (uint32_t)i2, (uint32_t)n1,
                                (uint32_t)i2, (uint32_t)n2, (uint32_t)n1,
                                (uint32_t)i2, 11u, (uint32_t)n2};
        idx.insert(idx.end(), f, f + 12);
        if (i == 5 && refCompat) {
            // the reference's last column: (10,2,1) and (10,11,3)
            idx[idx.size() - 5] = 2;
            idx[idx.size() - 1] = 3;
        }
        h1 += H;
        h2 += H;
    }
    v[11] = mk3(0, 0, -radius);
    for (int s = 0; s < subdivisions; ++s) idx = subdivide(v, idx);
    vb.resize(v.size());
    for (size_t i = 0; i < v.size(); ++i) {
        StkVertex& o = vb[i];
        o = StkVertex{};
        o.position[0] = v[i].x;
        o.position[1] = v[i].y;
        o.position[2] = v[i].z;
        o.normal = packNormal(v[i]);
        o.tangent = packNormal(mk3(1.0f, 0.0f, 0.0f));
    }
    ib = idx;
}

void cameraMatrices(F3 eye, F3 target, F3 up, float vfovDeg, uint32_t width, uint32_t height, float znear,
                    float zfar, StkPathTracerParam& p)
{
    memset(&p, 0, sizeof(p));
    // look-at view (right-handed, camera looks down -z)
    double f[3] = {target.x - eye.x, target.y - eye.y, target.z - eye.z};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& x : f) x /= fl;
    double s[3] = {f[1] * up.z - f[2] * up.y, f[2] * up.x - f[0] * up.z, f[0] * up.y - f[1] * up.x};
    double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& x : s) x /= sl;
    double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    const double e[3] = {eye.x, eye.y, eye.z};
    Mat4 view = Mat4::identity(), v2w = Mat4::identity();
    for (int c = 0; c < 3; ++c) {
        view.m[c * 4 + 0] = (float)s[c];
        view.m[c * 4 + 1] = (float)u[c];
        view.m[c * 4 + 2] = (float)-f[c];
        v2w.m[0 * 4 + c] = (float)s[c];
        v2w.m[1 * 4 + c] = (float)u[c];
        v2w.m[2 * 4 + c] = (float)-f[c];
        v2w.m[3 * 4 + c] = (float)e[c];
    }
    view.m[12] = (float)-(s[0] * e[0] + s[1] * e[1] + s[2] * e[2]);
    view.m[13] = (float)-(u[0] * e[0] + u[1] * e[1] + u[2] * e[2]);
    view.m[14] = (float)(f[0] * e[0] + f[1] * e[1] + f[2] * e[2]);
    // perspective() with near/far swapped for reverse z (src/scene/camera.cpp:61-131)
    const float aspect = (float)width / (float)height;
    const float n = zfar, fz = znear;
    const float focal = 1.0f / std::tan(vfovDeg * (float)M_PI / 180.0f / 2.0f);
    const float x = focal / aspect, y = -focal, A = n / (fz - n), B = fz * A;
    Mat4 proj{}, inv{};
    proj.m[0] = x;
    proj.m[5] = y;
    proj.m[10] = A;
    proj.m[14] = B;
    proj.m[11] = -1.0f;
    inv.m[0] = 1.0f / x;
    inv.m[5] = 1.0f / y;
    inv.m[11] = 1.0f / B; // row 3, col 2
    inv.m[15] = A / B;    // row 3, col 3
    inv.m[14] = -1.0f;    // row 2, col 3
    memcpy(p.viewToWorld, v2w.m, sizeof(p.viewToWorld));
    memcpy(p.clipToView, inv.m, sizeof(p.clipToView));
    memcpy(p.viewToClip, proj.m, sizeof(p.viewToClip));
    memcpy(p.worldToView, view.m, sizeof(p.worldToView));
    p.camPos[0] = eye.x;
    p.camPos[1] = eye.y;
    p.camPos[2] = eye.z;
    p.camPos[3] = 1.0f;
    p.dimension[0] = (int32_t)width;
    p.dimension[1] = (int32_t)height;
    p.invDimension[0] = 1.0f / (float)width;
    p.invDimension[1] = 1.0f / (float)height;
    p.spp = 1;
    p.maxDepth = 8;
}

} // namespace stk
