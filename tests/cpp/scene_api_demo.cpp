// Builds the lit floor (a 40 x 40 Lambert floor of albedo 0.5 under a 1 x 1
// rect light of intensity 10 at y = 2 facing down, a 64 x 64 camera at
// (0, 1, 0) looking down) through the oka::Scene-shaped container of
// include/stk/scene.hpp, checks its arrays byte for byte against the generated
// workload 9, then renders it through stk::PtRender::setScene(const Scene&)
// and writes three frames' RGBA32F result to argv[1].  The generated scene is
// rendered the same way in the same process and must give the same bytes.
// Exit codes: 0 rendered and equal, 2 arrays differ, 3 no usable HIP device,
// 4 frames differ, 1 any other failure.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "stk/ptrender.hpp"
#include "stk/scene.hpp"

namespace {

// column-major 4x4 product, summed in the order k = 0..3 from 0.0f
void mul(const float* a, const float* b, float* out)
{
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            float s = 0.0f;
            for (int k = 0; k < 4; ++k) s += a[k * 4 + r] * b[c * 4 + k];
            out[c * 4 + r] = s;
        }
}

void identity(float* m)
{
    std::memset(m, 0, 16 * sizeof(float));
    m[0] = m[5] = m[10] = m[15] = 1.0f;
}

stk::Scene litFloor()
{
    stk::Scene s;
    StkMaterial m;
    std::memset(&m, 0, sizeof(m));
    m.type = STK_MAT_LAMBERT;
    m.roughness = 1.0f;
    m.albedo[0] = m.albedo[1] = m.albedo[2] = 0.5f;
    m.albedo[3] = 1.0f;
    const uint32_t mat = s.addMaterial(m);

    std::vector<StkVertex> vb(4);
    const float xs[4] = {-20.0f, 20.0f, 20.0f, -20.0f}, zs[4] = {-20.0f, -20.0f, 20.0f, 20.0f};
    for (int k = 0; k < 4; ++k) {
        std::memset(&vb[k], 0, sizeof(StkVertex));
        vb[k].position[0] = xs[k];
        vb[k].position[2] = zs[k];
        vb[k].normal = stk::Scene::packNormal(0.0f, 1.0f, 0.0f);
        vb[k].tangent = stk::Scene::packNormal(1.0f, 0.0f, 0.0f);
    }
    const uint32_t mesh = s.createMesh(vb, {0, 2, 1, 0, 3, 2});
    float eye4[16];
    identity(eye4);
    const float center[3] = {0.0f, 0.0f, 0.0f};
    s.createInstance(mesh, mat, eye4, center);

    // translate(0, 2, 0) * rotateX(-pi/2): the light's emitting side (local -z) faces -y
    float t[16], r[16];
    identity(t);
    t[13] = 2.0f;
    identity(r);
    const float a = -(float)M_PI / 2.0f, c = std::cos(a), sn = std::sin(a);
    r[5] = c;
    r[6] = sn;
    r[9] = -sn;
    r[10] = c;
    StkLightDesc d;
    std::memset(&d, 0, sizeof(d));
    d.type = 0;
    mul(t, r, d.xform);
    d.color[0] = d.color[1] = d.color[2] = 1.0f;
    d.intensity = 10.0f;
    d.width = d.height = 1.0f;
    d.radius = 1.0f;
    s.createLight(d);

    const float eye[3] = {0.0f, 1.0f, 0.0f}, target[3] = {0.0f, 0.0f, 0.0f}, up[3] = {0.0f, 0.0f, -1.0f};
    s.addCamera(stk::Scene::lookAt(eye, target, up, 60.0f, 64, 64, 0.01f, 100.0f));
    return s;
}

template <class T>
bool same(const stk::View<T>& a, const stk::View<T>& b)
{
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

int render(const stk::Scene& s, std::vector<float>& out)
{
    stk::PtRender pt(0);
    const stk_status st = pt.init();
    if (st == STK_E_DEVICE) return 3;
    if (st != STK_OK || pt.setScene(s) != STK_OK) {
        std::printf("setScene: %s\n", pt.lastError().c_str());
        return 1;
    }
    pt.setSetting("render/pt/spp", 4);
    pt.setSetting("render/pt/depth", 4);
    pt.setSetting("render/pt/stratifiedSamplingType", 1);
    const StkPathTracerParam cam = s.getCamera(0);
    out.assign((size_t)cam.dimension[0] * cam.dimension[1] * 4, 0.0f);
    stk::Image img;
    img.data = out.data();
    img.width = (uint32_t)cam.dimension[0];
    img.height = (uint32_t)cam.dimension[1];
    for (int f = 0; f < 3; ++f)
        if (pt.drawFrame(img) != STK_OK) {
            std::printf("drawFrame: %s\n", pt.lastError().c_str());
            return 1;
        }
    return 0;
}

} // namespace

int main(int argc, char** argv)
{
    stk::Scene built = litFloor();
    stk::Scene gen = stk::Scene::generate(9);
    if (!built.valid() || !gen.valid()) return 1;
    const bool equal = same(built.getVertices(), gen.getVertices()) && same(built.getIndices(), gen.getIndices()) &&
                       same(built.getInstances(), gen.getInstances()) && same(built.getLights(), gen.getLights()) &&
                       same(built.getMaterials(), gen.getMaterials()) && built.getCameraCount() == 1 &&
                       gen.getCameraCount() == 1;
    const StkPathTracerParam cb = built.getCamera(0), cg = gen.getCamera(0);
    if (!equal || std::memcmp(&cb, &cg, sizeof(cb)) != 0) {
        std::printf("arrays differ\n");
        return 2;
    }
    std::printf("arrays equal: %zu vertices, %zu instances, %zu lights\n", built.getVertices().size(),
                built.getInstances().size(), built.getLights().size());
    // a failed call reports (uint32_t)-1 like the reference's ids
    if (built.createMesh({}, {}) != stk::Scene::kInvalidId || built.getLightInstance(5) != stk::Scene::kInvalidId)
        return 1;

    std::vector<float> a, b;
    int rc = render(built, a);
    if (rc == 3) {
        std::printf("no device\n");
        return 3;
    }
    if (rc != 0 || (rc = render(gen, b)) != 0) return rc;
    if (a.size() != b.size() || std::memcmp(a.data(), b.data(), a.size() * sizeof(float)) != 0) {
        std::printf("frames differ\n");
        return 4;
    }
    FILE* out = argc > 1 ? std::fopen(argv[1], "wb") : nullptr;
    if (!out) return 1;
    std::fwrite(a.data(), sizeof(float), a.size(), out);
    std::fclose(out);
    std::printf("rendered 64x64, equal to the generated scene\n");
    return 0;
}
