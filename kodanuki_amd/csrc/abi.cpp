// The C ABI of libkodanuki_hip.so (include/kodanuki.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.h"
#include "kodanuki.h"
#include "scene.h"
#include "tables.h"

struct kdn_scene {
    kdn::Scene scene;
    std::vector<kdn_step> steps;
    std::vector<double> values;
    std::vector<kdn_object> objects;
    uint64_t serial = 0;  // unique per loaded scene in the process
};


namespace {

thread_local std::string g_error;

int fail(const std::string& m, int code) {
    g_error = m;
    return code;
}

int ctxFail(kdn_ctx* ctx, int code) {
    g_error = ctx->c.error;
    return code;
}

void copyError(const std::string& m, char* buf, size_t size) {
    if (!buf || size == 0) return;
    const size_t n = std::min(size - 1, m.size());
    std::memcpy(buf, m.data(), n);
    buf[n] = 0;
}

// The float nearest a double, and which side of it the double lies.
kdn_step toStep(const kdn::Step& s) {
    kdn_step o;
    o.op = s.op;
    o.a = (uint16_t)s.a;
    o.b = (uint16_t)s.b;
    o.c = (uint16_t)s.c;
    o.leaf = s.leaf;
    const float f = (float)s.value;
    o.value = f;
    o.flags = (double)f == s.value ? 0 : ((double)f < s.value ? 1 : 2);
    if (s.op != KDN_CONSTANT && s.op != KDN_ARGUMENT) o.flags = 0;
    return o;
}

void flatten(kdn_scene* s) {
    s->steps.clear();
    s->values.clear();
    s->objects.clear();
    for (size_t i = 0; i < s->scene.programs.size(); ++i) {
        const kdn::Program& p = s->scene.programs[i];
        kdn_object ob;
        ob.first = (uint32_t)s->steps.size();
        ob.count = (uint32_t)p.steps.size();
        ob.result = p.result;
        ob.material = s->scene.programMaterial[i];
        for (const kdn::Step& st : p.steps) {
            s->steps.push_back(toStep(st));
            s->values.push_back(st.value);
        }
        s->objects.push_back(ob);
    }
}

// ---- FP64 intervals for the region search (reference semantics,
// source/sdf/numbers.cpp:198-430, the ops the standard primitives use) ----
struct DI {
    double lo, hi;
};
DI dmul(DI a, DI b) {
    const double p[4] = {a.lo * b.lo, a.lo * b.hi, a.hi * b.lo, a.hi * b.hi};
    return {std::min({p[0], p[1], p[2], p[3]}), std::max({p[0], p[1], p[2], p[3]})};
}
DI devalStep(const kdn::Step& s, const std::vector<DI>& r, const DI box[3]) {
    auto A = [&]() { return r[s.a]; };
    auto B = [&]() { return r[s.b]; };
    switch (s.op) {
        case KDN_CONSTANT: case KDN_ARGUMENT: return {s.value, s.value};
        case KDN_POINT_X: return box[0];
        case KDN_POINT_Y: return box[1];
        case KDN_POINT_Z: return box[2];
        case KDN_ADD: return {A().lo + B().lo, A().hi + B().hi};
        case KDN_SUBTRACT: return {A().lo - B().hi, A().hi - B().lo};
        case KDN_MULTIPLY: return dmul(A(), B());
        case KDN_DIVIDE: {
            DI b = B();
            if (b.lo <= 0 && b.hi >= 0) return {-INFINITY, INFINITY};
            return dmul(A(), DI{1 / b.hi, 1 / b.lo});
        }
        case KDN_NEGATE: return {-A().hi, -A().lo};
        case KDN_SQUARE_ROOT: return {std::sqrt(std::max(A().lo, 0.0)), std::sqrt(std::max(A().hi, 0.0))};
        case KDN_ABSOLUTE: {
            DI a = A();
            if (a.lo >= 0) return a;
            if (a.hi <= 0) return {-a.hi, -a.lo};
            return {0, std::max(-a.lo, a.hi)};
        }
        case KDN_MINIMUM: return {std::min(A().lo, B().lo), std::min(A().hi, B().hi)};
        case KDN_MAXIMUM: return {std::max(A().lo, B().lo), std::max(A().hi, B().hi)};
        case KDN_SMOOTH_MINIMUM: {
            const double rr = r[s.c].lo;
            auto sm = [&](double a, double b) {
                const double blend = std::max(rr - std::abs(a - b), 0.0) / rr;
                return std::min(a, b) - blend * blend * rr * 0.25;
            };
            double lo = sm(A().lo, B().lo), hi = sm(A().hi, B().hi);
            if (std::isnan(lo)) lo = std::min(A().lo, B().lo) - rr / 4;
            if (std::isnan(hi)) hi = std::min(A().hi, B().hi);
            return {lo, hi};
        }
        case KDN_POSITIVE_SQUARE: {
            const double l = std::max(A().lo, 0.0), h = std::max(A().hi, 0.0);
            return {l * l, h * h};
        }
        default: return {-INFINITY, INFINITY};  // anything else: no claim
    }
}
DI devalProgram(const kdn::Program& p, const DI box[3], std::vector<DI>& r) {
    r.resize(p.steps.size());
    for (size_t i = 0; i < p.steps.size(); ++i) r[i] = devalStep(p.steps[i], r, box);
    return r[p.result];
}

void searchRegion(const kdn::Scene& scene, const DI box[3], const std::vector<uint32_t>& objects, int level, int maxLevel,
                  double lo[3], double hi[3], std::vector<DI>& r) {
    // only cells that may hold the surface count: none Empty for every
    // object, none Filled by one (a filled cell holds no surface)
    std::vector<uint32_t> near;
    for (uint32_t o : objects) {
        const DI v = devalProgram(scene.programs[o], box, r);
        if (v.hi < 0) return;
        if (!(v.lo > 0)) near.push_back(o);
    }
    if (near.empty()) return;
    if (level == maxLevel) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], box[k].lo);
            hi[k] = std::max(hi[k], box[k].hi);
        }
        return;
    }
    for (int oct = 0; oct < 8; ++oct) {
        DI child[3];
        for (int k = 0; k < 3; ++k) {
            const double mid = (box[k].lo + box[k].hi) / 2;
            child[k] = ((oct >> k) & 1) ? DI{mid, box[k].hi} : DI{box[k].lo, mid};
        }
        searchRegion(scene, child, near, level + 1, maxLevel, lo, hi, r);
    }
}

}  // namespace

namespace kdn {

int setGlobalError(const std::string& m, int code) { return fail(m, code); }

bool ensure(DevBuf& b, size_t bytes, std::string& err, bool keep) {
    if (b.bytes >= bytes) return true;
    size_t n = b.bytes ? b.bytes : (1u << 16);
    while (n < bytes) n *= 2;
    void* p = nullptr;
    if (hipMalloc(&p, n) != hipSuccess) {
        err = "hipMalloc of " + std::to_string(n) + " bytes failed";
        return false;
    }
    if (keep && b.ptr) hipMemcpy(p, b.ptr, b.bytes, hipMemcpyDeviceToDevice);
    if (b.ptr) hipFree(b.ptr);
    b.ptr = p;
    b.bytes = n;
    return true;
}

static void freeBuf(DevBuf& b) {
    if (b.ptr) hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

uint64_t sceneSerial(const kdn_scene* scene) { return scene->serial; }

void releaseScratch(Ctx& c) {
    for (DevBuf& b : c.scratch) freeBuf(b);
}

uint64_t residentBytes(const Ctx& c) {
    const SceneDev& s = c.scene;
    uint64_t bytes = (uint64_t)s.n_steps * sizeof(kdn_step) + (uint64_t)s.n_objects * sizeof(kdn_object) +
                     (uint64_t)s.n_materials * sizeof(kdn_material) + (uint64_t)s.n_lights * sizeof(kdn_light) +
                     (uint64_t)s.n_leaves * 4;
    bytes += c.tree.nodes.bytes + c.tree.entries.bytes + c.tree.pool.bytes + c.tree.spool.bytes + c.tree.mpool.bytes;
    for (const DevBuf& b : c.scratch) bytes += b.bytes;
    bytes += c.paths_bytes + c.round_bytes + c.hguides.bytes + c.sort_tmp.bytes + guideRoundBytes(c);
    if (c.accum) bytes += (uint64_t)c.width * c.height * (6 * sizeof(double) + sizeof(uint32_t));
    return bytes;
}

static void freeScene(SceneDev& s) {
    if (s.steps) hipFree(s.steps);
    if (s.objects) hipFree(s.objects);
    if (s.materials) hipFree(s.materials);
    if (s.lights) hipFree(s.lights);
    if (s.leaf_material) hipFree(s.leaf_material);
    s = SceneDev{};
}

void releaseAll(Ctx& c) {
    freeScene(c.scene);
    freeBuf(c.tree.nodes);
    freeBuf(c.tree.entries);
    freeBuf(c.tree.pool);
    freeBuf(c.tree.spool);
    freeBuf(c.tree.mpool);
    freeBuf(c.job_counters);
    if (c.tree.level_first_dev) hipFree(c.tree.level_first_dev);
    c.tree = OctreeDev{};
    for (DevBuf& b : c.scratch) freeBuf(b);
    if (c.accum) hipFree(c.accum);
    if (c.accum_sq) hipFree(c.accum_sq);
    if (c.accum_n) hipFree(c.accum_n);
    c.accum = c.accum_sq = nullptr;
    c.accum_n = nullptr;
    freeIntegrator(c);
    freeGuide(c);
    freeBuf(c.hguides);
    if (c.leaf_steps_dev) hipFree(c.leaf_steps_dev);
    if (c.leaf_progs_dev) hipFree(c.leaf_progs_dev);
    c.leaf_steps_dev = c.leaf_progs_dev = nullptr;
    c.n_leaf_progs = 0;
    freeBuf(c.sort_tmp);
    for (DevBuf* b : {&c.stop_wanted, &c.stop_spreads, &c.stop_allotted, &c.stop_counts, &c.stop_offsets, &c.stop_tmp, &c.tm_history[0],
                      &c.tm_history[1], &c.tm_gbuffer})
        freeBuf(*b);
    if (c.firsts) hipFree(c.firsts);
    c.firsts = nullptr;
    c.firsts_width = c.firsts_height = 0;
    c.tm_width = c.tm_height = 0;
    c.stop_width = c.stop_height = 0;
    c.stop_total = 0;
    c.hguide_width = c.hguide_height = 0;
    c.hguide_deep = 0;
    if (c.materials_dev) hipFree(c.materials_dev);
    if (c.lamps_dev) hipFree(c.lamps_dev);
    c.materials_dev = c.lamps_dev = nullptr;
    c.integrator_ready = false;
}

}  // namespace kdn

extern "C" {

const char* kdn_global_error(void) { return g_error.c_str(); }

int kdn_scene_load(const char* text, const char* base_dir, kdn_scene** out, char* err, size_t err_size) {
    if (!text || !out) return fail("kdn_scene_load: null argument", KDN_E_ARGUMENT);
    kdn_scene* s = new kdn_scene();
    static std::atomic<uint64_t> loaded{0};
    s->serial = ++loaded;
    std::string e;
    try {
        if (!kdn::loadScene(text, base_dir ? base_dir : "", s->scene, e)) {
            copyError(e, err, err_size);
            delete s;
            return fail(e, KDN_E_SCENE);
        }
        flatten(s);
    } catch (const std::exception& x) {
        copyError(x.what(), err, err_size);
        delete s;
        return fail(x.what(), KDN_E_SCENE);
    }
    *out = s;
    return KDN_OK;
}

void kdn_scene_free(kdn_scene* scene) { delete scene; }

int kdn_scene_describe(const kdn_scene* s, kdn_scene_info* info) {
    if (!s || !info) return fail("kdn_scene_describe: null argument", KDN_E_ARGUMENT);
    info->steps = (uint32_t)s->steps.size();
    info->objects = (uint32_t)s->objects.size();
    info->leaves = (uint32_t)s->scene.leaves.size();
    info->nodes = (uint32_t)s->scene.nodes.size();
    info->top_nodes = (uint32_t)s->scene.objects.size();
    uint32_t children = 0;
    for (const kdn::Node& n : s->scene.nodes) children += (uint32_t)n.children.size();
    info->children = children;
    info->materials = (uint32_t)s->scene.materials.size();
    info->lights = (uint32_t)s->scene.lights.size();
    return KDN_OK;
}

int kdn_scene_tape(const kdn_scene* s, kdn_step* steps, kdn_object* objects) {
    if (!s) return fail("kdn_scene_tape: null scene", KDN_E_ARGUMENT);
    if (steps) std::memcpy(steps, s->steps.data(), s->steps.size() * sizeof(kdn_step));
    if (objects) std::memcpy(objects, s->objects.data(), s->objects.size() * sizeof(kdn_object));
    return KDN_OK;
}

int kdn_scene_tape_values(const kdn_scene* s, double* values) {
    if (!s || !values) return fail("kdn_scene_tape_values: null argument", KDN_E_ARGUMENT);
    std::memcpy(values, s->values.data(), s->values.size() * sizeof(double));
    return KDN_OK;
}

int kdn_scene_leaves(const kdn_scene* s, kdn_leaf* leaves) {
    if (!s || !leaves) return fail("kdn_scene_leaves: null argument", KDN_E_ARGUMENT);
    for (size_t i = 0; i < s->scene.leaves.size(); ++i) {
        const kdn::Leaf& l = s->scene.leaves[i];
        kdn_leaf& o = leaves[i];
        o.primitive = l.primitive;
        o.material = l.material;
        std::memcpy(o.args, l.args, sizeof(o.args));
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) o.linear[3 * r + c] = l.linear[r][c];
        std::memcpy(o.origin, l.origin, sizeof(o.origin));
        o.stretch = l.stretch;
    }
    return KDN_OK;
}

int kdn_scene_leaf_program(const kdn_scene* s, uint32_t leaf, kdn_step* steps, double* values, uint32_t cap, uint32_t* count,
                           uint32_t* result) {
    if (!s || !count || !result) return fail("kdn_scene_leaf_program: null argument", KDN_E_ARGUMENT);
    if (leaf >= s->scene.leaves.size()) return fail("kdn_scene_leaf_program: no such leaf", KDN_E_ARGUMENT);
    const kdn::Program prog = kdn::leafProgram(s->scene, leaf);
    *count = (uint32_t)prog.steps.size();
    *result = prog.result;
    if (!steps && !values) return KDN_OK;
    if (prog.steps.size() > cap) return fail("kdn_scene_leaf_program: more steps than cap", KDN_E_ARGUMENT);
    for (size_t i = 0; i < prog.steps.size(); ++i) {
        const kdn::Step& st = prog.steps[i];
        if (steps) steps[i] = kdn_step{st.op, 0, (uint16_t)st.a, (uint16_t)st.b, (uint16_t)st.c, (float)st.value, st.leaf};
        if (values) values[i] = st.value;
    }
    return KDN_OK;
}

int kdn_scene_nodes(const kdn_scene* s, kdn_node* nodes, uint32_t* top_nodes, uint32_t* children) {
    if (!s) return fail("kdn_scene_nodes: null scene", KDN_E_ARGUMENT);
    uint32_t at = 0;
    for (size_t i = 0; i < s->scene.nodes.size(); ++i) {
        const kdn::Node& n = s->scene.nodes[i];
        if (nodes) {
            nodes[i].join = n.join;
            nodes[i].radius = n.radius;
            nodes[i].leaf = n.leaf;
            nodes[i].first_child = at;
            nodes[i].child_count = (uint32_t)n.children.size();
        }
        for (uint32_t ch : n.children) {
            if (children) children[at] = ch;
            ++at;
        }
    }
    if (top_nodes) std::memcpy(top_nodes, s->scene.objects.data(), s->scene.objects.size() * 4);
    return KDN_OK;
}

int kdn_scene_materials(const kdn_scene* s, kdn_material* m) {
    if (!s || !m) return fail("kdn_scene_materials: null argument", KDN_E_ARGUMENT);
    std::memcpy(m, s->scene.materials.data(), s->scene.materials.size() * sizeof(kdn_material));
    return KDN_OK;
}

int kdn_scene_lights(const kdn_scene* s, kdn_light* l) {
    if (!s || !l) return fail("kdn_scene_lights: null argument", KDN_E_ARGUMENT);
    std::memcpy(l, s->scene.lights.data(), s->scene.lights.size() * sizeof(kdn_light));
    return KDN_OK;
}

int kdn_scene_settings(const kdn_scene* s, kdn_camera* camera, kdn_render_settings* render) {
    if (!s) return fail("kdn_scene_settings: null scene", KDN_E_ARGUMENT);
    if (camera) *camera = s->scene.camera;
    if (render) *render = s->scene.render;
    return KDN_OK;
}

int kdn_scene_region(const kdn_scene* s, int depth_cap, float* lower3, float* size3) {
    if (!s || !lower3 || !size3) return fail("kdn_scene_region: null argument", KDN_E_ARGUMENT);
    if (depth_cap < 1 || depth_cap > 14) return fail("kdn_scene_region: depth cap out of 1..14", KDN_E_ARGUMENT);
    const kdn::Scene& sc = s->scene;
    const double far_ = sc.render.far_;
    DI box[3];
    for (int k = 0; k < 3; ++k) box[k] = DI{sc.camera.target[k] - far_, sc.camera.target[k] + far_};
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    std::vector<uint32_t> all(sc.programs.size());
    for (size_t i = 0; i < all.size(); ++i) all[i] = (uint32_t)i;
    std::vector<DI> regs;
    searchRegion(sc, box, all, 0, 5, lo, hi, regs);
    if (!(lo[0] <= hi[0])) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = box[k].lo;
            hi[k] = box[k].hi;
        }
    }
    // per axis: the least power of two that covers the extent, its corner
    // on the finest grid
    for (int k = 0; k < 3; ++k) {
        double sz = std::ldexp(1.0, (int)std::ceil(std::log2(std::max(hi[k] - lo[k], 1.0 / 16))));
        for (;;) {
            const double finest = std::ldexp(sz, -depth_cap);
            const double mid = (lo[k] + hi[k]) / 2;
            const double corner = std::floor((mid - sz / 2) / finest) * finest;
            if (corner <= lo[k] && corner + sz >= hi[k]) {
                lower3[k] = (float)corner;
                size3[k] = (float)sz;
                break;
            }
            sz *= 2;
        }
    }
    return KDN_OK;
}

int kdn_open(int device, kdn_ctx** out) {
    if (!out) return fail("kdn_open: null argument", KDN_E_ARGUMENT);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail("kdn_open: no HIP device", KDN_E_DEVICE);
    if (device < 0 || device >= count) return fail("kdn_open: no such device", KDN_E_ARGUMENT);
    if (hipSetDevice(device) != hipSuccess) return fail("kdn_open: hipSetDevice failed", KDN_E_DEVICE);
    kdn_ctx* ctx = new kdn_ctx();
    ctx->c.device = device;
    if (hipStreamCreate(&ctx->c.stream) != hipSuccess) {
        delete ctx;
        return fail("kdn_open: hipStreamCreate failed", KDN_E_DEVICE);
    }
    if (!kdn::denoiseStaticScratch(ctx->c)) {
        const std::string why = ctx->c.error;
        hipStreamDestroy(ctx->c.stream);
        delete ctx;
        return fail("kdn_open: " + why, KDN_E_DEVICE);
    }
    *out = ctx;
    return KDN_OK;
}

void kdn_close(kdn_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->c.device);
    hipStreamSynchronize(ctx->c.stream);
    for (hipStream_t& js : ctx->c.job_streams) {
        if (js) (void)hipStreamSynchronize(js);
    }
    kdn::releaseAll(ctx->c);
    for (hipEvent_t& e : ctx->c.job_events) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    for (hipStream_t& js : ctx->c.job_streams) {
        if (js) (void)hipStreamDestroy(js);
        js = nullptr;
    }
    hipStreamDestroy(ctx->c.stream);
    delete ctx;
}

int kdn_last_error(const kdn_ctx* ctx, char* buf, size_t size) {
    copyError(ctx ? ctx->c.error : g_error, buf, size);
    return KDN_OK;
}

int kdn_upload_scene(kdn_ctx* ctx, const kdn_scene* s) {
    if (!ctx || !s) return fail("kdn_upload_scene: null argument", KDN_E_ARGUMENT);
    kdn::Ctx& c = ctx->c;
    hipSetDevice(c.device);
    kdn::SceneDev& d = c.scene;
    if (d.steps) hipFree(d.steps);
    if (d.objects) hipFree(d.objects);
    if (d.materials) hipFree(d.materials);
    if (d.lights) hipFree(d.lights);
    if (d.leaf_material) hipFree(d.leaf_material);
    d = kdn::SceneDev{};
    c.tree.built = false;
    kdn::freeGuide(c);  // the guides belong to one scene
    c.hguide_width = c.hguide_height = 0;
    c.hguide_deep = 0;
    std::vector<kdn_material> mats = s->scene.materials;
    const uint32_t fallback = (uint32_t)mats.size();
    kdn_material grey{};
    grey.kind = KDN_MATTE;
    grey.albedo[0] = grey.albedo[1] = grey.albedo[2] = 0.5f;
    grey.roughness = 1;
    grey.strength = 1;
    mats.push_back(grey);  // the material of a leaf painted with none
    std::vector<uint32_t> leafMat;
    for (const kdn::Leaf& l : s->scene.leaves) leafMat.push_back(l.material == KDN_NO_MATERIAL ? fallback : l.material);
    if (leafMat.empty()) leafMat.push_back(fallback);
    std::vector<kdn_light> lights = s->scene.lights;
    auto up = [&](void** dst, const void* src, size_t bytes) -> bool {
        if (hipMalloc(dst, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
        return bytes == 0 || hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!up((void**)&d.steps, s->steps.data(), s->steps.size() * sizeof(kdn_step)) ||
        !up((void**)&d.objects, s->objects.data(), s->objects.size() * sizeof(kdn_object)) ||
        !up((void**)&d.materials, mats.data(), mats.size() * sizeof(kdn_material)) ||
        !up((void**)&d.lights, lights.data(), lights.size() * sizeof(kdn_light)) ||
        !up((void**)&d.leaf_material, leafMat.data(), leafMat.size() * 4)) {
        c.error = "kdn_upload_scene: device allocation or copy failed";
        return ctxFail(ctx, KDN_E_MEMORY);
    }
    // the integrator's materials (the fallback last) and lamps
    const std::vector<kdn::MaterialDev> md = kdn::materialTable(mats.data(), mats.size());
    const std::vector<kdn::LampDev> lamps = kdn::lampTable(lights.data(), lights.size());
    if (c.materials_dev) hipFree(c.materials_dev);
    if (c.lamps_dev) hipFree(c.lamps_dev);
    c.materials_dev = c.lamps_dev = nullptr;
    if (!up(&c.materials_dev, md.data(), md.size() * sizeof(kdn::MaterialDev)) ||
        !up(&c.lamps_dev, lamps.data(), lamps.size() * sizeof(kdn::LampDev))) {
        c.error = "kdn_upload_scene: device allocation or copy failed";
        return ctxFail(ctx, KDN_E_MEMORY);
    }
    // every leaf's own program, the survey's chains walk them (FP64)
    {
        std::vector<kdn::LeafStepDev> lsteps;
        std::vector<kdn::LeafProgDev> lprogs;
        for (uint32_t leaf = 0; leaf < (uint32_t)s->scene.leaves.size(); ++leaf) {
            const kdn::Program prog = kdn::leafProgram(s->scene, leaf);
            lprogs.push_back(kdn::LeafProgDev{(uint32_t)lsteps.size(), (uint32_t)prog.steps.size(), prog.result, 0});
            for (const kdn::Step& st : prog.steps) lsteps.push_back(kdn::LeafStepDev{st.op, st.a, st.b, st.c, st.value});
        }
        if (c.leaf_steps_dev) hipFree(c.leaf_steps_dev);
        if (c.leaf_progs_dev) hipFree(c.leaf_progs_dev);
        c.leaf_steps_dev = c.leaf_progs_dev = nullptr;
        if (!up(&c.leaf_steps_dev, lsteps.data(), lsteps.size() * sizeof(kdn::LeafStepDev)) ||
            !up(&c.leaf_progs_dev, lprogs.data(), lprogs.size() * sizeof(kdn::LeafProgDev))) {
            c.error = "kdn_upload_scene: device allocation or copy failed";
            return ctxFail(ctx, KDN_E_MEMORY);
        }
        c.n_leaf_progs = (uint32_t)lprogs.size();
    }
    c.scene_has_glass = false;
    for (uint32_t m : leafMat) c.scene_has_glass = c.scene_has_glass || (m < mats.size() && mats[m].kind == KDN_GLASS);
    c.n_materials_dev = (uint32_t)md.size();
    c.n_lamps = (uint32_t)lamps.size();
    c.integrator_ready = true;
    d.n_steps = (uint32_t)s->steps.size();
    d.n_objects = (uint32_t)s->objects.size();
    d.n_materials = (uint32_t)mats.size();
    d.n_lights = (uint32_t)lights.size();
    d.n_leaves = (uint32_t)leafMat.size();
    d.camera = s->scene.camera;
    d.render = s->scene.render;
    return KDN_OK;
}

int kdn_build_octree(kdn_ctx* ctx, const float* lower3, const float* size3, int depth_cap, kdn_octree_stats* stats) {
    if (!ctx || !lower3 || !size3) return fail("kdn_build_octree: null argument", KDN_E_ARGUMENT);
    if (depth_cap < 0 || depth_cap > kdn::kMaxDepth) return fail("kdn_build_octree: depth cap out of 0..11", KDN_E_ARGUMENT);
    // each side a power of two and each corner on the finest grid, so that
    // every cell box is exact in FP32
    for (int k = 0; k < 3; ++k) {
        int e = 0;
        if (!(size3[k] > 0) || std::frexp((double)size3[k], &e) != 0.5) {
            return fail("kdn_build_octree: region sizes must be powers of two", KDN_E_ARGUMENT);
        }
        const double q = (double)lower3[k] / std::ldexp((double)size3[k], -depth_cap);
        if (q != std::floor(q) || std::abs(q) > 16777216.0) {
            return fail("kdn_build_octree: region corner must be a multiple of the finest cell", KDN_E_ARGUMENT);
        }
    }
    kdn::Ctx& c = ctx->c;
    hipSetDevice(c.device);
    if (c.scene.n_objects == 0) {
        c.error = "kdn_build_octree: no scene uploaded";
        return ctxFail(ctx, KDN_E_STATE);
    }
    if (!kdn::buildOctree(c, lower3, size3, depth_cap, stats)) return ctxFail(ctx, KDN_E_DEVICE);
    return KDN_OK;
}

int kdn_octree_sizes(kdn_ctx* ctx, uint64_t* nodes, uint64_t* entries, uint64_t* pool_steps) {
    if (!ctx) return fail("kdn_octree_sizes: null context", KDN_E_ARGUMENT);
    if (!ctx->c.tree.built) {
        ctx->c.error = "kdn_octree_sizes: no octree";
        return ctxFail(ctx, KDN_E_STATE);
    }
    if (nodes) *nodes = ctx->c.tree.n_nodes;
    if (entries) *entries = ctx->c.tree.n_entries;
    if (pool_steps) *pool_steps = ctx->c.tree.n_pool;
    return KDN_OK;
}

int kdn_octree_download(kdn_ctx* ctx, kdn_node_record* nodes, kdn_entry_record* entries, kdn_step* pool) {
    if (!ctx) return fail("kdn_octree_download: null context", KDN_E_ARGUMENT);
    kdn::Ctx& c = ctx->c;
    if (!c.tree.built) {
        c.error = "kdn_octree_download: no octree";
        return ctxFail(ctx, KDN_E_STATE);
    }
    hipSetDevice(c.device);
    static_assert(sizeof(kdn::NodeDev) == sizeof(kdn_node_record), "node layout");
    static_assert(sizeof(kdn::EntryDev) == sizeof(kdn_entry_record), "entry layout");
    if ((nodes && hipMemcpy(nodes, c.tree.nodes.ptr, c.tree.n_nodes * sizeof(kdn_node_record), hipMemcpyDeviceToHost) != hipSuccess) ||
        (entries && hipMemcpy(entries, c.tree.entries.ptr, c.tree.n_entries * sizeof(kdn_entry_record), hipMemcpyDeviceToHost) != hipSuccess) ||
        (pool && hipMemcpy(pool, c.tree.pool.ptr, c.tree.n_pool * sizeof(kdn_step), hipMemcpyDeviceToHost) != hipSuccess)) {
        c.error = "kdn_octree_download: copy failed";
        return ctxFail(ctx, KDN_E_DEVICE);
    }
    return KDN_OK;
}

int kdn_extend(kdn_ctx* ctx, const kdn_ray* rays, uint32_t count, kdn_hit* hits) {
    if (!ctx || (count && (!rays || !hits))) return fail("kdn_extend: null argument", KDN_E_ARGUMENT);
    kdn::Ctx& c = ctx->c;
    hipSetDevice(c.device);
    if (count == 0) return KDN_OK;
    if (!kdn::ensure(c.scratch[6], (size_t)count * sizeof(kdn_ray), c.error, false) ||
        !kdn::ensure(c.scratch[7], (size_t)count * sizeof(kdn_hit), c.error, false)) {
        return ctxFail(ctx, KDN_E_MEMORY);
    }
    if (hipMemcpyAsync(c.scratch[6].ptr, rays, (size_t)count * sizeof(kdn_ray), hipMemcpyHostToDevice, c.stream) != hipSuccess) {
        c.error = "kdn_extend: upload failed";
        return ctxFail(ctx, KDN_E_DEVICE);
    }
    if (!kdn::extendRays(c, (const kdn_ray*)c.scratch[6].ptr, count, (kdn_hit*)c.scratch[7].ptr, c.stream)) {
        return ctxFail(ctx, KDN_E_STATE);
    }
    if (hipMemcpyAsync(hits, c.scratch[7].ptr, (size_t)count * sizeof(kdn_hit), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess) {
        c.error = std::string("kdn_extend: ") + hipGetErrorString(hipGetLastError());
        return ctxFail(ctx, KDN_E_DEVICE);
    }
    return KDN_OK;
}

int kdn_extend_device(kdn_ctx* ctx, const void* rays, uint32_t count, void* hits, void* stream) {
    if (!ctx || (count && (!rays || !hits))) return fail("kdn_extend_device: null argument", KDN_E_ARGUMENT);
    kdn::Ctx& c = ctx->c;
    hipSetDevice(c.device);
    if (!kdn::extendRays(c, (const kdn_ray*)rays, count, (kdn_hit*)hits, (hipStream_t)stream)) return ctxFail(ctx, KDN_E_STATE);
    return KDN_OK;
}

int kdn_render_pass(kdn_ctx* ctx, const kdn_pass_desc* desc, kdn_pass_stats* stats) {
    if (!ctx || !desc) return fail("kdn_render_pass: null argument", KDN_E_ARGUMENT);
    kdn::Ctx& c = ctx->c;
    hipSetDevice(c.device);
    if (desc->width <= 0 || desc->height <= 0 || desc->width > 16384 || desc->height > 16384) {
        return fail("kdn_render_pass: image size out of 1..16384", KDN_E_ARGUMENT);
    }
    if (!c.tree.built) {
        c.error = "kdn_render_pass: no octree";
        return ctxFail(ctx, KDN_E_STATE);
    }
    if (c.scene_has_glass && !KDN_GLASS_ON_DEVICE) {
        c.error = "kdn_render_pass: glass scattering is held off the device (renders with glass faulted it; the cause is "
                  "not found): paint the scene without glass";
        return ctxFail(ctx, KDN_E_STATE);
    }
    if (!kdn::renderPass(c, *desc, stats)) return ctxFail(ctx, KDN_E_DEVICE);
    return KDN_OK;
}

int kdn_read_accumulation(kdn_ctx* ctx, double* rgb_sum, double* sq_sum, uint32_t* count) {
    if (!ctx) return fail("kdn_read_accumulation: null context", KDN_E_ARGUMENT);
    kdn::Ctx& c = ctx->c;
    hipSetDevice(c.device);
    if (!c.accum) {
        c.error = "kdn_read_accumulation: nothing rendered";
        return ctxFail(ctx, KDN_E_STATE);
    }
    const size_t px = (size_t)c.width * c.height;
    hipStreamSynchronize(c.stream);
    if ((rgb_sum && hipMemcpy(rgb_sum, c.accum, px * 3 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
        (sq_sum && hipMemcpy(sq_sum, c.accum_sq, px * 3 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
        (count && hipMemcpy(count, c.accum_n, px * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)) {
        c.error = "kdn_read_accumulation: copy failed";
        return ctxFail(ctx, KDN_E_DEVICE);
    }
    return KDN_OK;
}

int kdn_read_estimates(kdn_ctx* ctx, kdn_estimate* estimates) {
    if (!ctx || !estimates) return fail("kdn_read_estimates: null argument", KDN_E_ARGUMENT);
    kdn::Ctx& c = ctx->c;
    if (!c.estimates) {
        c.error = "kdn_read_estimates: no pass kept estimates (kdn_pass_desc.flags bit 6)";
        return ctxFail(ctx, KDN_E_STATE);
    }
    hipSetDevice(c.device);
    const size_t px = (size_t)c.width * c.height;
    if (hipStreamSynchronize(c.stream) != hipSuccess ||
        hipMemcpy(estimates, c.estimates, px * sizeof(kdn_estimate), hipMemcpyDeviceToHost) != hipSuccess) {
        c.error = "kdn_read_estimates: copy failed";
        return ctxFail(ctx, KDN_E_DEVICE);
    }
    return KDN_OK;
}

int kdn_fold_samples(kdn_ctx* ctx, const double* colours, uint32_t pixels, uint32_t spp, uint32_t batch, kdn_estimate* estimates) {
    if (!ctx || (pixels && spp && (!colours || !estimates))) return fail("kdn_fold_samples: null argument", KDN_E_ARGUMENT);
    if ((uint64_t)pixels * spp > (1ull << 28)) return fail("kdn_fold_samples: more than 2^28 samples", KDN_E_ARGUMENT);
    if (pixels == 0 || spp == 0) return KDN_OK;
    kdn::Ctx& c = ctx->c;
    hipSetDevice(c.device);
    if (!kdn::foldSamples(c, colours, pixels, spp, batch, estimates)) return ctxFail(ctx, KDN_E_DEVICE);
    return KDN_OK;
}

int kdn_stopping_effort(double tolerance, int32_t* samples, int32_t* batch) {
    if (!samples || !batch) return fail("kdn_stopping_effort: null argument", KDN_E_ARGUMENT);
    kdn::stoppingEffort(tolerance, samples, batch);
    return KDN_OK;
}

int kdn_stopping_trial(const kdn_stopping* stopping, int32_t* samples) {
    if (!stopping || !samples) return fail("kdn_stopping_trial: null argument", KDN_E_ARGUMENT);
    std::string err;
    if (!kdn::stoppingCheck(*stopping, err)) return fail("kdn_stopping_trial: " + err, KDN_E_ARGUMENT);
    *samples = kdn::stoppingTrial(*stopping);
    return KDN_OK;
}

int kdn_stopping_allot(kdn_ctx* ctx, const kdn_stopping* stopping, int32_t trial, uint64_t* total) {
    if (!ctx || !stopping) return fail("kdn_stopping_allot: null argument", KDN_E_ARGUMENT);
    kdn::Ctx& c = ctx->c;
    std::string err;
    if (!kdn::stoppingCheck(*stopping, err)) {
        c.error = "kdn_stopping_allot: " + err;
        return ctxFail(ctx, KDN_E_ARGUMENT);
    }
    hipSetDevice(c.device);
    if (!kdn::stoppingAllot(c, *stopping, trial, total)) return ctxFail(ctx, c.estimates ? KDN_E_ARGUMENT : KDN_E_STATE);
    return KDN_OK;
}

int kdn_read_stopping(kdn_ctx* ctx, float* wanted, float* spreads, int32_t* allotted) {
    if (!ctx) return fail("kdn_read_stopping: null argument", KDN_E_ARGUMENT);
    kdn::Ctx& c = ctx->c;
    hipSetDevice(c.device);
    if (!kdn::readStopping(c, wanted, spreads, allotted)) return ctxFail(ctx, KDN_E_STATE);
    return KDN_OK;
}

int kdn_sequence_numbers(kdn_ctx* ctx, uint64_t stream, const uint32_t* indices, uint32_t count, uint32_t coordinate, int32_t sobol,
                         double* out) {
    if (!ctx || (count && (!indices || !out))) return fail("kdn_sequence_numbers: null argument", KDN_E_ARGUMENT);
    if (count > (1u << 24)) return fail("kdn_sequence_numbers: more than 2^24 numbers", KDN_E_ARGUMENT);
    if (count == 0) return KDN_OK;
    kdn::Ctx& c = ctx->c;
    hipSetDevice(c.device);
    if (!kdn::sequenceNumbers(c, stream, indices, count, coordinate, sobol != 0, out)) return ctxFail(ctx, KDN_E_DEVICE);
    return KDN_OK;
}

int kdn_set_camera(kdn_ctx* ctx, const kdn_camera* camera) {
    if (!ctx || !camera) return fail("kdn_set_camera: null argument", KDN_E_ARGUMENT);
    kdn::Ctx& c = ctx->c;
    if (!c.integrator_ready) {
        c.error = "kdn_set_camera: no scene on the device";
        return ctxFail(ctx, KDN_E_STATE);
    }
    c.scene.camera = *camera;
    return KDN_OK;
}

int kdn_temporal_frame(kdn_ctx* ctx, const kdn_temporal_desc* desc, float* shown, float* denoised, kdn_temporal_stats* stats) {
    if (!ctx || !desc) return fail("kdn_temporal_frame: null argument", KDN_E_ARGUMENT);
    kdn::Ctx& c = ctx->c;
    hipSetDevice(c.device);
    if (!kdn::temporalFrame(c, *desc, shown, denoised, stats)) return ctxFail(ctx, c.firsts ? KDN_E_DEVICE : KDN_E_STATE);
    return KDN_OK;
}

int kdn_temporal_history(kdn_ctx* ctx, kdn_denoise_pixel* pixels) {
    if (!ctx || !pixels) return fail("kdn_temporal_history: null argument", KDN_E_ARGUMENT);
    kdn::Ctx& c = ctx->c;
    hipSetDevice(c.device);
    if (!kdn::temporalHistory(c, pixels)) return ctxFail(ctx, KDN_E_STATE);
    return KDN_OK;
}

int kdn_temporal_step(kdn_ctx* ctx, int32_t width, int32_t height, const kdn_lens* lens, const kdn_lens* before,
                      const kdn_temporal_desc* desc, const double* parts, const kdn_surface* record, const kdn_surface* gbuffer,
                      const kdn_denoise_pixel* before_pixels, int32_t before_width, int32_t before_height,
                      kdn_denoise_pixel* pixels, float* shown) {
    if (!ctx || !lens || !before || !desc || !parts || !record || !gbuffer || !pixels || !shown)
        return fail("kdn_temporal_step: null argument", KDN_E_ARGUMENT);
    if (width <= 0 || height <= 0 || (size_t)width * (size_t)height > ((size_t)1 << 26))
        return fail("kdn_temporal_step: an image of 1 to 2^26 pixels", KDN_E_ARGUMENT);
    if (before_pixels && (before_width <= 0 || before_height <= 0))
        return fail("kdn_temporal_step: a history needs a positive size", KDN_E_ARGUMENT);
    kdn::Ctx& c = ctx->c;
    hipSetDevice(c.device);
    if (!kdn::temporalStep(c, width, height, *lens, *before, *desc, parts, record, gbuffer, before_pixels, before_width, before_height,
                           pixels, shown))
        return ctxFail(ctx, KDN_E_DEVICE);
    return KDN_OK;
}

int kdn_read_image(kdn_ctx* ctx, float* rgb) {
    if (!ctx || !rgb) return fail("kdn_read_image: null argument", KDN_E_ARGUMENT);
    kdn::Ctx& c = ctx->c;
    const size_t px = (size_t)c.width * c.height;
    std::vector<double> sum(px * 3);
    std::vector<uint32_t> n(px);
    int rc = kdn_read_accumulation(ctx, sum.data(), nullptr, n.data());
    if (rc) return rc;
    for (size_t i = 0; i < px; ++i) {
        for (int k = 0; k < 3; ++k) rgb[3 * i + k] = n[i] ? (float)(sum[3 * i + k] / n[i]) : 0.0f;
    }
    return KDN_OK;
}

}  // extern "C"
