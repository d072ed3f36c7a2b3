// sanitize_driver.cpp - drives the host half of the library (scene container
// and its building calls, workload generators, BVH builders (with and without
// spatial splits) and packers, two-level build, file
// formats, host C-ABI entry points) and the CPU oracle, for a build under
// -fsanitize=address,undefined (tests/test_sanitizers.py).  No device code:
// the HIP sources are not part of this build.  Exits 0 when every check
// passes; a sanitizer report aborts the process.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/stk/stk.h"
#include "../../include/stk/stk_scene.h"
extern "C" {
#include "../../oracle/oracle.h"
}

namespace {

int failures = 0;

void check(bool ok, const char* what)
{
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s\n", what);
        ++failures;
    }
}

struct Scene {
    stk_scene* h = nullptr;
    StkSceneArrays a{};
    explicit Scene(uint32_t config, uint32_t variant = 0)
    {
        check(stk_scene_generate(config, variant, 0, &h) == STK_OK, "stk_scene_generate");
        check(stk_scene_get(h, &a) == STK_OK, "stk_scene_get");
    }
    ~Scene() { stk_scene_free(h); }
};

OrcScene oracleScene(const StkSceneArrays& a, const stk_bvh* b)
{
    OrcScene s;
    std::memset(&s, 0, sizeof(s));
    s.nodes = stk_bvh_nodes(b);
    s.nnodes = stk_bvh_num_nodes(b);
    s.inst = a.inst;
    s.ninst = a.numInstances;
    s.vb = a.vb;
    s.nv = a.numVertices;
    s.ib = a.ib;
    s.ni = a.numIndices;
    s.lights = a.lights;
    s.nlights = a.numLights;
    s.mats = a.mats;
    s.nmats = a.numMaterials;
    s.wide = stk_bvh_wide_nodes(b);
    s.nwide = stk_bvh_num_wide(b);
    s.wideTris = stk_bvh_wide_triangles(b);
    s.nwideTris = stk_bvh_num_triangles(b);
    return s;
}

void flatScene(uint32_t config, uint32_t variant, bool pathTrace)
{
    Scene sc(config, variant);
    const StkSceneArrays& a = sc.a;
    stk_bvh* b = nullptr;
    check(stk_bvh_build_host(a.vb, a.numVertices, a.ib, a.numIndices, a.inst, a.numInstances, nullptr, &b) == STK_OK,
          "stk_bvh_build_host");
    if (!b) return;
    uint32_t tris = 0;
    for (uint32_t i = 0; i < a.numInstances; ++i) tris += (uint32_t)a.inst[i].indexCount / 3;
    check(stk_bvh_num_nodes(b) == 2 * tris - 1, "one triangle per leaf");
    check(stk_bvh_num_triangles(b) == tris, "packed triangles");
    check(stk_bvh_wide_depth(b) >= 1 && stk_bvh_wide_depth(b) <= 16, "wide depth");
    // the packer accepts its own skip-link array
    stk_bvh* again = nullptr;
    check(stk_bvh_from_nodes(a.vb, a.numVertices, a.ib, a.numIndices, a.inst, a.numInstances, stk_bvh_nodes(b),
                             stk_bvh_num_nodes(b), &again) == STK_OK,
          "stk_bvh_from_nodes");
    stk_bvh_free(again);
    // a broken skip link is refused
    std::vector<StkBvhNode> broken(stk_bvh_nodes(b), stk_bvh_nodes(b) + stk_bvh_num_nodes(b));
    if (broken.size() > 2) broken[1].nodeOffset = 1 << 30;
    stk_bvh* bad = nullptr;
    check(stk_bvh_from_nodes(a.vb, a.numVertices, a.ib, a.numIndices, a.inst, a.numInstances, broken.data(),
                             (uint32_t)broken.size(), &bad) != STK_OK,
          "broken skip link refused");
    stk_bvh_free(bad);

    const uint32_t w = (uint32_t)a.camera.dimension[0], h = (uint32_t)a.camera.dimension[1];
    std::vector<StkRay> rays((size_t)w * h);
    check(stk_scene_jittered_rays(sc.h, 0x5354524B00000001ull, rays.data(), (uint32_t)rays.size()) == STK_OK,
          "jittered rays");
    std::vector<StkRay> sub;
    for (size_t i = 0; i < rays.size(); i += 13) sub.push_back(rays[i]);
    OrcScene os = oracleScene(a, b);
    std::vector<StkHit> ref(sub.size()), other(sub.size());
    OrcStats st;
    orc_trace_batch(&os, sub.data(), (uint32_t)sub.size(), ref.data(), ORC_MODE_REF, 4, &st);
    for (int mode : {ORC_MODE_PRUNED, ORC_MODE_WIDE, ORC_MODE_BRUTE}) {
        orc_trace_batch(&os, sub.data(), (uint32_t)sub.size(), other.data(), mode, 4, &st);
        check(std::memcmp(ref.data(), other.data(), ref.size() * sizeof(StkHit)) == 0, "walks agree");
    }
    orc_trace_batch(&os, sub.data(), (uint32_t)sub.size(), other.data(), ORC_MODE_ANY, 4, &st);
    std::vector<StkHit> brute(sub.size());
    orc_trace_batch(&os, sub.data(), (uint32_t)sub.size(), brute.data(), ORC_MODE_ANY_BRUTE, 4, &st);
    for (size_t i = 0; i < sub.size(); ++i)
        check(other[i].instId == brute[i].instId && other[i].primId == brute[i].primId, "any hit = first in preorder");

    if (pathTrace) {
        StkRenderParams rp;
        std::memset(&rp, 0, sizeof(rp));
        rp.pt = a.camera;
        rp.pt.spp = 2;
        rp.pt.maxDepth = 6;
        rp.pt.stratifiedSamplingType = 3;
        rp.rrStartDepth = 3;
        rp.refCompat = 1;
        for (int k = 0; k < 3; ++k) rp.skyColor[k] = a.sky[k];
        const uint32_t cw = w < 16 ? w : 16, ch = h < 8 ? h : 8;
        std::vector<float> samples((size_t)cw * ch * 2 * 3);
        orc_pt_frame(&os, &rp, 0, 0, cw, ch, samples.data(), 2);
        std::vector<float> rgba((size_t)cw * ch * 4), acc(rgba.size()), tm(rgba.size()), hist(rgba.size(), 0.0f);
        orc_reduce(samples.data(), cw * ch, 2, rgba.data());
        orc_accumulate(rgba.data(), hist.data(), cw * ch, 0, acc.data());
        for (uint32_t t = 0; t < 3; ++t) orc_tonemap(acc.data(), cw * ch, t, tm.data());
        bool finite = true;
        for (float v : samples) finite = finite && std::isfinite(v);
        check(finite, "finite path samples");
    }

    // file round trips
    const std::string path = std::string(std::getenv("STK_SAN_TMP") ? std::getenv("STK_SAN_TMP") : "/tmp") +
                             "/san_nodes.stkbvh";
    check(stk_bvh_save_nodes(stk_bvh_nodes(b), stk_bvh_num_nodes(b), path.c_str()) == STK_OK, "save nodes");
    uint32_t n = 0;
    check(stk_bvh_load_nodes(path.c_str(), nullptr, 0, &n) == STK_OK && n == stk_bvh_num_nodes(b), "query nodes");
    std::vector<StkBvhNode> loaded(n);
    check(stk_bvh_load_nodes(path.c_str(), loaded.data(), n, &n) == STK_OK, "load nodes");
    check(std::memcmp(loaded.data(), stk_bvh_nodes(b), (size_t)n * sizeof(StkBvhNode)) == 0, "nodes round trip");
    check(stk_bvh_load_nodes(path.c_str(), loaded.data(), n - 1, &n) != STK_OK, "short capacity refused");
    stk_bvh_free(b);
}

// The spatial-split builder: every walk agrees with the exhaustive test on
// the closed rooms' camera rays (a sample of them).
void spatialSplits(uint32_t config)
{
    Scene sc(config);
    const StkSceneArrays& a = sc.a;
    StkBvhOptions o;
    std::memset(&o, 0, sizeof(o));
    o.maxDepth = 32;
    o.numBins = 32;
    o.traversalCost = o.intersectionCost = 1.0f;
    o.spatialSplitAlpha = 1e-5f;
    stk_bvh* b = nullptr;
    check(stk_bvh_build_host(a.vb, a.numVertices, a.ib, a.numIndices, a.inst, a.numInstances, &o, &b) == STK_OK,
          "spatial-split build");
    if (!b) return;
    uint32_t tris = 0;
    for (uint32_t i = 0; i < a.numInstances; ++i) tris += (uint32_t)a.inst[i].indexCount / 3;
    check(stk_bvh_num_triangles(b) > tris && stk_bvh_num_triangles(b) <= tris + tris * 3 / 10 + 2,
          "split references within the budget");
    const uint32_t w = (uint32_t)a.camera.dimension[0], h = (uint32_t)a.camera.dimension[1];
    std::vector<StkRay> rays((size_t)w * h);
    check(stk_scene_jittered_rays(sc.h, 7, rays.data(), (uint32_t)rays.size()) == STK_OK, "rays");
    std::vector<StkRay> sub;
    for (size_t i = 0; i < rays.size(); i += 17) sub.push_back(rays[i]);
    OrcScene os = oracleScene(a, b);
    std::vector<StkHit> ref(sub.size()), other(sub.size());
    OrcStats st;
    orc_trace_batch(&os, sub.data(), (uint32_t)sub.size(), ref.data(), ORC_MODE_REF, 4, &st);
    for (int mode : {ORC_MODE_PRUNED, ORC_MODE_WIDE, ORC_MODE_BRUTE}) {
        orc_trace_batch(&os, sub.data(), (uint32_t)sub.size(), other.data(), mode, 4, &st);
        check(std::memcmp(ref.data(), other.data(), ref.size() * sizeof(StkHit)) == 0, "split-tree walks agree");
    }
    stk_bvh_free(b);
}

void twoLevel()
{
    Scene sc(5, 60);
    const StkSceneArrays& a = sc.a;
    stk_tlas* t = nullptr;
    check(stk_tlas_build_host(a.vb, a.numVertices, a.ib, a.numIndices, a.inst, a.numInstances, nullptr, &t) ==
              STK_OK,
          "stk_tlas_build_host");
    if (!t) return;
    uint32_t top = 0, mesh = 0;
    stk_tlas_depths(t, &top, &mesh);
    check(top >= 1 && mesh >= 1 && top + mesh <= 16, "two-level depths");
    uint32_t cnt = 0;
    check(stk_tlas_top_wide(t, &cnt) != nullptr && cnt > 0, "instance tree");
    check(stk_tlas_mesh_triangles(t, &cnt) != nullptr && cnt > 0, "mesh triangles");
    stk_tlas_free(t);
}

void images()
{
    const uint32_t w = 7, h = 5;
    std::vector<float> rgba(w * h * 4);
    for (size_t i = 0; i < rgba.size(); ++i) rgba[i] = (i % 4 == 3) ? 1.0f : 0.25f * (float)(i % 13);
    const std::string path = std::string(std::getenv("STK_SAN_TMP") ? std::getenv("STK_SAN_TMP") : "/tmp") +
                             "/san.pfm";
    check(stk_image_write_pfm(path.c_str(), rgba.data(), w, h) == STK_OK, "pfm write");
    uint32_t rw = 0, rh = 0;
    check(stk_image_read_pfm(path.c_str(), nullptr, 0, &rw, &rh) == STK_OK && rw == w && rh == h, "pfm size");
    std::vector<float> back(rgba.size());
    check(stk_image_read_pfm(path.c_str(), back.data(), back.size(), &rw, &rh) == STK_OK, "pfm read");
    check(back == rgba, "pfm round trip");
    check(stk_image_read_pfm(path.c_str(), back.data(), 3, &rw, &rh) != STK_OK, "short image buffer refused");
    const std::string exr = path.substr(0, path.size() - 4) + ".exr";
    check(stk_image_write_exr(exr.c_str(), rgba.data(), w, h) == STK_OK, "exr write");
    FILE* f = std::fopen(exr.c_str(), "rb");
    check(f != nullptr, "exr exists");
    if (f) {
        std::fseek(f, 0, SEEK_END);
        const long size = std::ftell(f);
        std::fclose(f);
        // header + one 8-byte offset per row + per row {y, byte count, 3 float rows}
        check(size > (long)(h * (8 + 8 + 12 * w)), "exr size");
    }
    check(stk_image_write_exr(exr.c_str(), rgba.data(), 0, h) != STK_OK, "empty exr refused");
}

// The scene-building calls: a generated workload re-entered from its own
// arrays gives the same arrays; updates and refused calls leave the container
// consistent.
void sceneApi()
{
    Scene gen(7);
    const StkSceneArrays& g = gen.a;
    stk_scene* s = nullptr;
    check(stk_scene_create(&s) == STK_OK, "stk_scene_create");
    if (!s) return;
    uint32_t id = 0;
    for (uint32_t i = 0; i < g.numMaterials; ++i)
        check(stk_scene_add_material(s, &g.mats[i], &id) == STK_OK && id == i, "add material");
    std::vector<std::pair<int32_t, uint32_t>> meshes; // (indexOffset, mesh id)
    for (uint32_t i = 0; i < g.numInstances; ++i) {
        const StkInstance& in = g.inst[i];
        if (in.lightId >= 0) continue;
        uint32_t mesh = 0xFFFFFFFFu;
        for (auto& m : meshes)
            if (m.first == in.indexOffset) mesh = m.second;
        if (mesh == 0xFFFFFFFFu) {
            uint32_t lo = 0xFFFFFFFFu, hi = 0;
            for (int32_t k = 0; k < in.indexCount; ++k) {
                lo = std::min(lo, g.ib[in.indexOffset + k]);
                hi = std::max(hi, g.ib[in.indexOffset + k]);
            }
            std::vector<uint32_t> local(g.ib + in.indexOffset, g.ib + in.indexOffset + in.indexCount);
            for (uint32_t& x : local) x -= lo;
            check(stk_scene_create_mesh(s, g.vb + lo, hi - lo + 1, local.data(), (uint32_t)local.size(), &mesh) ==
                      STK_OK,
                  "create mesh");
            meshes.push_back({in.indexOffset, mesh});
        }
        check(stk_scene_create_instance(s, mesh, in.materialId, in.objectToWorld, &id) == STK_OK, "create instance");
    }
    for (uint32_t i = 0; i < g.numInstances; ++i) {
        const StkInstance& in = g.inst[i];
        if (in.lightId < 0) continue;
        StkLightDesc d;
        std::memset(&d, 0, sizeof(d));
        std::memcpy(d.xform, in.objectToWorld, sizeof(d.xform));
        d.color[0] = d.color[1] = d.color[2] = 1.0f;
        d.intensity = g.lights[in.lightId].color[0];
        d.width = d.height = d.radius = 1.0f;
        check(stk_scene_create_light(s, &d, &id) == STK_OK, "create light");
    }
    check(stk_scene_add_camera(s, &g.camera, &id) == STK_OK && id == 0, "add camera");
    StkSceneArrays a;
    check(stk_scene_get(s, &a) == STK_OK, "get rebuilt");
    check(a.numVertices == g.numVertices && std::memcmp(a.vb, g.vb, sizeof(StkVertex) * g.numVertices) == 0,
          "rebuilt vertices");
    check(a.numIndices == g.numIndices && std::memcmp(a.ib, g.ib, 4 * (size_t)g.numIndices) == 0, "rebuilt indices");
    check(a.numInstances == g.numInstances &&
              std::memcmp(a.inst, g.inst, sizeof(StkInstance) * g.numInstances) == 0,
          "rebuilt instances");
    check(a.numLights == g.numLights && std::memcmp(a.lights, g.lights, sizeof(StkLight) * g.numLights) == 0,
          "rebuilt lights");
    // updates, then refused calls
    StkLightDesc d;
    std::memset(&d, 0, sizeof(d));
    for (int k = 0; k < 4; ++k) d.xform[k * 5] = 1.0f;
    d.width = d.height = 2.0f;
    d.intensity = 4.0f;
    check(stk_scene_update_light(s, 0, &d) == STK_OK, "update light");
    d.type = 2;
    check(stk_scene_update_light(s, 0, &d) != STK_OK, "light type change refused");
    check(stk_scene_update_instance_transform(s, 0, d.xform) == STK_OK, "update transform");
    check(stk_scene_update_instance_transform(s, 1u << 30, d.xform) != STK_OK, "unknown instance refused");
    const uint32_t bad[3] = {0, 1, 9};
    check(stk_scene_create_mesh(s, g.vb, 3, bad, 3, &id) != STK_OK, "index past the vertices refused");
    check(stk_scene_create_instance(s, 1u << 20, -1, d.xform, &id) != STK_OK, "unknown mesh refused");
    d.type = 1;
    d.radius = 0.5f;
    check(stk_scene_create_light(s, &d, &id) == STK_OK, "disc light");
    d.type = 2;
    check(stk_scene_create_light(s, &d, &id) == STK_OK, "sphere light");
    check(stk_scene_get(s, &a) == STK_OK && a.numLights == g.numLights + 2, "lights after updates");
    StkPathTracerParam cam;
    const float eye[3] = {0, 0, 3}, target[3] = {0, 0, 0}, up[3] = {0, 1, 0};
    check(stk_camera_look_at(eye, target, up, 50.0f, 16, 8, 0.01f, 10.0f, &cam) == STK_OK, "look at");
    check(stk_scene_update_camera(s, 0, &cam) == STK_OK && stk_scene_camera_count(s) == 1, "update camera");
    check(stk_scene_get_camera(s, 3, &cam) != STK_OK, "unknown camera refused");
    stk_scene_free(s);
}

void errorPaths()
{
    stk_bvh* b = nullptr;
    check(stk_bvh_build_host(nullptr, 0, nullptr, 0, nullptr, 0, nullptr, nullptr) != STK_OK, "null out");
    const uint32_t ib[3] = {0, 1, 7};
    StkVertex vb[3];
    std::memset(vb, 0, sizeof(vb));
    StkInstance inst;
    std::memset(&inst, 0, sizeof(inst));
    inst.indexCount = 3;
    inst.lightId = -1;
    for (int k = 0; k < 4; ++k) inst.objectToWorld[k * 5] = 1.0f;
    check(stk_bvh_build_host(vb, 3, ib, 3, &inst, 1, nullptr, &b) != STK_OK, "index past the vertices refused");
    stk_bvh_free(b);
    stk_scene* s = nullptr;
    check(stk_scene_generate(99, 0, 0, &s) != STK_OK && s == nullptr, "unknown config refused");
}

} // namespace

int main()
{
    flatScene(1, 0, false);
    flatScene(6, 0, true);
    flatScene(7, 0, true);
    flatScene(9, 0, true);
    spatialSplits(1);
    spatialSplits(7);
    twoLevel();
    sceneApi();
    images();
    errorPaths();
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("sanitized host build: all checks passed\n");
    return 0;
}
