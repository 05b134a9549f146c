// Style table from a reference scene file (akari_scenegraph JSON). Each requested material's
// shader graph (akari_scenegraph/src/shader.rs:24-28) is walked from its output node, and its
// sty_* nodes (shader.rs:201-263) are lowered the way the SVM compiler lowers them
// (akari_render/src/svm/compiler.rs:337-477). Any other surface node (diffuse, principled, null,
// ...) ends a chain as the identity (the default apply_stylization, svm/surface/mod.rs:275-285).
// Identical subtrees are shared, so scenes whose materials repeat one chain need few nodes.
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#include "common.hpp"
#include "json.hpp"
#include "style.hpp"

namespace nlpf {

namespace {

struct M4 {
    double m[4][4];
};
M4 mat_identity() {
    M4 r{};
    for (int i = 0; i < 4; ++i) r.m[i][i] = 1;
    return r;
}
M4 mat_mul(const M4& a, const M4& b) {
    M4 r{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k) r.m[i][j] += a.m[i][k] * b.m[k][j];
    return r;
}
// rotation about a coordinate axis (0 = x, 1 = y, 2 = z), right-handed
M4 mat_rot(int axis, double a) {
    M4 r = mat_identity();
    const double c = std::cos(a), s = std::sin(a);
    const int i = (axis + 1) % 3, j = (axis + 2) % 3;
    r.m[i][i] = c;
    r.m[i][j] = -s;
    r.m[j][i] = s;
    r.m[j][j] = c;
    return r;
}
M4 mat_translate(double x, double y, double z) {
    M4 r = mat_identity();
    r.m[0][3] = x;
    r.m[1][3] = y;
    r.m[2][3] = z;
    return r;
}
M4 mat_scale(double x, double y, double z) {
    M4 r{};
    r.m[0][0] = x;
    r.m[1][1] = y;
    r.m[2][2] = z;
    r.m[3][3] = 1;
    return r;
}

struct Lowering {
    const JVal* nodes = nullptr;  // the material's shader graph nodes
    std::string material;
    std::vector<nlpf_style_node>* table = nullptr;
    uint32_t cap = 0;
    std::map<std::string, int> memo;
    std::string err;
    int depth = 0;

    bool fail(const std::string& m) {
        if (err.empty()) err = "material '" + material + "': " + m;
        return false;
    }
    const JVal* node(const std::string& id) { return jget(*nodes, id.c_str()); }
    static const char* type_of(const JVal& n) {
        const JVal* t = jget(n, "type");
        return t && t->kind == JVal::Str ? t->str.c_str() : "";
    }
    bool ref_id(const JVal& n, const char* field, std::string& id) {
        const JVal* r = jget(n, field);
        const JVal* i = r ? jget(*r, "id") : nullptr;
        if (!i || i->kind != JVal::Str) return fail(std::string("missing node reference '") + field + "'");
        id = i->str;
        return true;
    }
    // a constant scalar input: a "float" node (svm/compiler.rs:128-131)
    bool scalar(const JVal& n, const char* field, float& out) {
        std::string id;
        if (!ref_id(n, field, id)) return false;
        const JVal* c = node(id);
        if (!c) return fail("unknown node '" + id + "'");
        const JVal* v = jget(*c, "value");
        if (strcmp(type_of(*c), "float") || !v || v->kind != JVal::Num)
            return fail(std::string("input '") + field + "' must be a constant float node");
        out = (float)v->num;
        return true;
    }
    // a constant colour input: "rgb" / "float3" (compiler.rs:132-143), possibly through
    // "spectral_uplift", which is the identity for an RGB colour pipeline
    bool color(const JVal& n, const char* field, float out[3]) {
        std::string id;
        if (!ref_id(n, field, id)) return false;
        const JVal* c = node(id);
        for (int hop = 0; c && !strcmp(type_of(*c), "spectral_uplift") && hop < 4; ++hop) {
            if (!ref_id(*c, "rgb", id)) return false;
            c = node(id);
        }
        if (!c) return fail("unknown node '" + id + "'");
        const JVal* v = jget(*c, "value");
        const char* t = type_of(*c);
        if ((strcmp(t, "rgb") && strcmp(t, "float3")) || !v || v->kind != JVal::Arr || v->arr.size() != 3)
            return fail(std::string("input '") + field + "' must be a constant rgb or float3 node");
        for (int k = 0; k < 3; ++k) {
            if (v->arr[k].kind != JVal::Num) return fail(std::string("input '") + field + "' must be numeric");
            out[k] = (float)v->arr[k].num;
        }
        return true;
    }
    bool uint_field(const JVal& n, const char* field, uint32_t& out) {
        const JVal* v = jget(n, field);
        if (!v || v->kind != JVal::Num || v->num < 0 || std::floor(v->num) != v->num || v->num > 4294967295.0)
            return fail(std::string("field '") + field + "' must be a non-negative integer");
        out = (uint32_t)v->num;
        return true;
    }
    // does any node reachable from n (through node references) carry a sty_* type?
    bool reaches_style(const JVal& n, int hops) {
        if (hops > 64) return false;
        if (n.kind == JVal::Obj) {
            const JVal* i = jget(n, "id");
            if (i && i->kind == JVal::Str && n.obj.size() == 1) {
                const JVal* c = node(i->str);
                if (!c) return false;
                if (!strncmp(type_of(*c), "sty_", 4)) return true;
                return reaches_style(*c, hops + 1);
            }
            for (const auto& kv : n.obj)
                if (reaches_style(kv.second, hops)) return true;
        } else if (n.kind == JVal::Arr) {
            for (const auto& e : n.arr)
                if (reaches_style(e, hops)) return true;
        }
        return false;
    }
    // shares an identical node already in the table, else appends
    int intern(const nlpf_style_node& s) {
        for (size_t i = 0; i < table->size(); ++i)
            if (!memcmp(&(*table)[i], &s, sizeof(s))) return (int)i;
        if (table->size() >= cap) {
            fail("style table capacity (" + std::to_string(cap) + " nodes) exceeded");
            return -2;
        }
        table->push_back(s);
        return (int)table->size() - 1;
    }

    // lowers node `id`; returns its table index, -1 for the identity, -2 on error
    int lower(const std::string& id) {
        auto it = memo.find(id);
        if (it != memo.end()) return it->second;
        if (++depth > NLPF_STYLE_STACK) {
            fail("stylization chain deeper than 16 nodes (or cyclic)");
            return -2;
        }
        const int r = lower_uncached(id);
        --depth;
        if (r >= -1) memo[id] = r;
        return r;
    }
    int lower_uncached(const std::string& id) {
        const JVal* n = node(id);
        if (!n) {
            fail("unknown node '" + id + "'");
            return -2;
        }
        const std::string t = type_of(*n);
        if (t == "output") {
            std::string in;
            if (!ref_id(*n, "node", in)) return -2;
            return lower(in);
        }
        if (t.compare(0, 4, "sty_")) {
            // a base surface: identity for stylization, and nothing stylized may hide below it
            if (reaches_style(*n, 0)) {
                fail("node '" + id + "' (" + t + ") has stylization nodes below it; not supported");
                return -2;
            }
            return -1;
        }
        nlpf_style_node s;
        memset(&s, 0, sizeof(s));
        s.inner = -1;
        s.else_ = -1;
        s.depth_mask = 0xFFFFFFFFu;
        std::string in;
        if (!ref_id(*n, "surface", in)) return -2;
        const int inner = lower(in);
        if (inner < -1) return -2;
        s.inner = inner;
        bool ok = true;
        if (t == "sty_toon") {  // compiler.rs:337-347
            s.type = NLPF_STY_TOON;
            ok = scalar(*n, "color_band_width", s.p[0]);
            if (ok && !(s.p[0] > 0)) ok = fail("toon color_band_width must be > 0");
        } else if (t == "sty_8bit") {  // compiler.rs:348-358
            s.type = NLPF_STY_EIGHT_BIT;
            ok = scalar(*n, "indirect_scale", s.p[0]);
        } else if (t == "sty_half_tone") {  // compiler.rs:359-396
            s.type = NLPF_STY_HALFTONE;
            ok = scalar(*n, "indirect_scale", s.p[0]) && scalar(*n, "radius_scale", s.p[1]) &&
                 scalar(*n, "radius", s.p[2]) && scalar(*n, "grid_distance", s.p[3]) &&
                 scalar(*n, "blend", s.p[4]) && color(*n, "dark_tone", s.p + 5) &&
                 color(*n, "light_tone", s.p + 8) && uint_field(*n, "max_depth", s.max_depth) &&
                 trs(*n, s.p + 11);
        } else if (t == "sty_saturation_adjust") {  // compiler.rs:397-414
            s.type = NLPF_STY_SATURATION;
            ok = scalar(*n, "power", s.p[0]) && scalar(*n, "indirect_scale", s.p[1]) &&
                 uint_field(*n, "max_depth", s.max_depth);
        } else if (t == "sty_saturation_adjust_rex") {  // compiler.rs:415-419
            s.type = NLPF_STY_REX;
            ok = scalar(*n, "blend", s.p[0]);
        } else if (t == "sty_color_palette") {  // compiler.rs:420-453
            s.type = NLPF_STY_PALETTE;
            ok = scalar(*n, "indirect_scale", s.p[0]) && scalar(*n, "gamma", s.p[1]) && palette(*n, s);
        } else if (t == "sty_hue_shift") {  // shader.rs:218-223 (parsed, never compiled by the reference)
            s.type = NLPF_STY_HUE_SHIFT;
            ok = scalar(*n, "shift", s.p[0]) && scalar(*n, "indirect_scale", s.p[1]);
        } else if (t == "sty_set_idx") {  // compiler.rs:454-458
            s.type = NLPF_STY_SET_IDX;
            ok = uint_field(*n, "idx", s.idx);
        } else if (t == "sty_enable_if") {  // compiler.rs:459-476 (no else: the null surface)
            s.type = NLPF_STY_ENABLE_IF;
            ok = uint_field(*n, "idx", s.idx);
            const JVal* e = jget(*n, "else");
            if (ok && e && e->kind != JVal::Null) {
                std::string eid;
                if (!ref_id(*n, "else", eid)) return -2;
                const int el = lower(eid);
                if (el < -1) return -2;
                s.else_ = el;
            }
        } else {
            fail("unknown stylization node type '" + t + "'");
            return -2;
        }
        if (!ok) return -2;
        if ((s.type == NLPF_STY_SET_IDX || s.type == NLPF_STY_ENABLE_IF) && s.idx > 3) {
            fail("style index must be 0..3");
            return -2;
        }
        return intern(s);
    }
    // palette entries "0xRRGGBB" -> rgb / 255 (compiler.rs:428-440)
    bool palette(const JVal& n, nlpf_style_node& s) {
        const JVal* p = jget(n, "palette");
        if (!p || p->kind != JVal::Arr) return fail("palette must be a list of hex strings");
        if (p->arr.size() > 18) return fail("palette holds at most 18 entries");
        s.count = (uint32_t)p->arr.size();
        for (size_t k = 0; k < p->arr.size(); ++k) {
            const JVal& e = p->arr[k];
            if (e.kind != JVal::Str) return fail("palette entries must be strings");
            const char* h = e.str.c_str();
            while (!strncmp(h, "0x", 2)) h += 2;  // trim_start_matches("0x")
            char* end = nullptr;
            const unsigned long x = strtoul(h, &end, 16);
            if (!*h || *end || x > 0xFFFFFFFFul) return fail("bad palette entry '" + e.str + "'");
            s.p[2 + 3 * k] = (float)((x >> 16) & 0xFF) / 255.0f;
            s.p[3 + 3 * k] = (float)((x >> 8) & 0xFF) / 255.0f;
            s.p[4 + 3 * k] = (float)(x & 0xFF) / 255.0f;
        }
        return true;
    }
    // TRS::to_matrix(false) (akari_scenegraph/src/scene.rs:31-61), first three rows
    bool trs(const JVal& n, float* out) {
        const JVal* t = jget(n, "trs");
        if (!t) return fail("half tone needs 'trs'");
        double v[3][3];
        const char* names[3] = {"translation", "rotation", "scale"};
        for (int f = 0; f < 3; ++f) {
            const JVal* a = jget(*t, names[f]);
            if (!a || a->kind != JVal::Arr || a->arr.size() != 3) return fail(std::string("trs.") + names[f]);
            for (int k = 0; k < 3; ++k) {
                if (a->arr[k].kind != JVal::Num) return fail(std::string("trs.") + names[f]);
                v[f][k] = (double)(float)a->arr[k].num;
            }
        }
        const JVal* cs = jget(*t, "coordinate_system");
        const std::string sys = cs && cs->kind == JVal::Str ? cs->str : "";
        M4 m = mat_scale(v[2][0], v[2][1], v[2][2]);
        if (sys == "Akari") {
            m = mat_mul(mat_rot(2, v[1][2]), m);
            m = mat_mul(mat_rot(0, v[1][0]), m);
            m = mat_mul(mat_rot(1, v[1][1]), m);
            m = mat_mul(mat_translate(v[0][0], v[0][1], v[0][2]), m);
        } else if (sys == "Blender") {
            m = mat_mul(mat_rot(0, v[1][0]), m);
            m = mat_mul(mat_rot(2, -v[1][1]), m);
            m = mat_mul(mat_rot(1, v[1][2]), m);
            m = mat_mul(mat_translate(v[0][0], v[0][2], -v[0][1]), m);
        } else {
            return fail("trs.coordinate_system must be \"Akari\" or \"Blender\"");
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 4; ++j) out[4 * i + j] = (float)m.m[i][j];
        return true;
    }
};

}  // namespace

}  // namespace nlpf

using namespace nlpf;

static int style_from_scene(const char* scene_json, const char* const* materials, uint32_t n_materials,
                            nlpf_style_node* nodes, uint32_t* n_nodes, int32_t* roots) {
    if (!scene_json || (n_materials && (!materials || !roots)) || !n_nodes || (*n_nodes && !nodes)) {
        set_global_error("nlpf_style_from_scene: NULL argument");
        return NLPF_E_INVALID_ARG;
    }
    JParser p{scene_json, ""};
    JVal root;
    if (!p.parse(root)) {
        set_global_error("scene JSON: " + p.err);
        return NLPF_E_INVALID_ARG;
    }
    const JVal* mats = jget(root, "materials");
    if (!mats || mats->kind != JVal::Obj) {
        set_global_error("scene JSON: no \"materials\" object");
        return NLPF_E_INVALID_ARG;
    }
    std::vector<nlpf_style_node> table;
    for (uint32_t m = 0; m < n_materials; ++m) {
        if (!materials[m]) {
            set_global_error("nlpf_style_from_scene: NULL material name");
            return NLPF_E_INVALID_ARG;
        }
        const JVal* mat = jget(*mats, materials[m]);
        const JVal* shader = mat ? jget(*mat, "shader") : nullptr;
        const JVal* gnodes = shader ? jget(*shader, "nodes") : nullptr;
        const JVal* out = shader ? jget(*shader, "output") : nullptr;
        const JVal* out_id = out ? jget(*out, "id") : nullptr;
        if (!gnodes || gnodes->kind != JVal::Obj || !out_id || out_id->kind != JVal::Str) {
            set_global_error(std::string("scene JSON: material '") + materials[m] + "' has no shader graph");
            return NLPF_E_INVALID_ARG;
        }
        Lowering lw;
        lw.nodes = gnodes;
        lw.material = materials[m];
        lw.table = &table;
        lw.cap = *n_nodes;
        const int r = lw.lower(out_id->str);
        if (r < -1) {
            set_global_error(lw.err);
            return NLPF_E_INVALID_ARG;
        }
        roots[m] = r;
    }
    if (!table.empty()) memcpy(nodes, table.data(), sizeof(nlpf_style_node) * table.size());
    *n_nodes = (uint32_t)table.size();
    return NLPF_OK;
}

extern "C" int nlpf_style_from_scene(const char* scene_json, const char* const* materials, uint32_t n_materials,
                                     nlpf_style_node* nodes, uint32_t* n_nodes, int32_t* roots) {
    try {
        return style_from_scene(scene_json, materials, n_materials, nodes, n_nodes, roots);
    } catch (const std::bad_alloc&) {
        set_global_error("nlpf_style_from_scene: out of host memory");
        return NLPF_E_OOM;
    }
}

// BSDF of each named material: the base surface below the output node, reached through the sty_*
// nodes' "surface" inputs (what their closures' evaluate delegates to, svm/surface/mod.rs:451-880)
static int bsdfs_from_scene(const char* scene_json, const char* const* materials, uint32_t n_materials,
                            nlpf_bsdf* out) {
    if (!scene_json || (n_materials && (!materials || !out))) {
        set_global_error("nlpf_bsdfs_from_scene: NULL argument");
        return NLPF_E_INVALID_ARG;
    }
    JParser p{scene_json, ""};
    JVal root;
    if (!p.parse(root)) {
        set_global_error("scene JSON: " + p.err);
        return NLPF_E_INVALID_ARG;
    }
    const JVal* mats = jget(root, "materials");
    if (!mats || mats->kind != JVal::Obj) {
        set_global_error("scene JSON: no \"materials\" object");
        return NLPF_E_INVALID_ARG;
    }
    for (uint32_t m = 0; m < n_materials; ++m) {
        if (!materials[m]) {
            set_global_error("nlpf_bsdfs_from_scene: NULL material name");
            return NLPF_E_INVALID_ARG;
        }
        const JVal* mat = jget(*mats, materials[m]);
        const JVal* shader = mat ? jget(*mat, "shader") : nullptr;
        const JVal* gnodes = shader ? jget(*shader, "nodes") : nullptr;
        const JVal* outn = shader ? jget(*shader, "output") : nullptr;
        const JVal* out_id = outn ? jget(*outn, "id") : nullptr;
        if (!gnodes || gnodes->kind != JVal::Obj || !out_id || out_id->kind != JVal::Str) {
            set_global_error(std::string("scene JSON: material '") + materials[m] + "' has no shader graph");
            return NLPF_E_INVALID_ARG;
        }
        Lowering lw;
        lw.nodes = gnodes;
        lw.material = materials[m];
        auto fail = [&](const std::string& msg, int code) {
            lw.fail(msg);
            set_global_error(lw.err);
            return code;
        };
        std::string id = out_id->str;
        const JVal* n = lw.node(id);
        for (int hop = 0; n; ++hop) {
            const char* t = Lowering::type_of(*n);
            const char* field = !strcmp(t, "output") ? "node" : (!strncmp(t, "sty_", 4) ? "surface" : nullptr);
            if (!field) break;
            if (hop > 64) return fail("surface chain deeper than 64 nodes (or cyclic)", NLPF_E_INVALID_ARG);
            if (!lw.ref_id(*n, field, id)) return fail("", NLPF_E_INVALID_ARG);
            n = lw.node(id);
        }
        if (!n) return fail("unknown node '" + id + "'", NLPF_E_INVALID_ARG);
        nlpf_bsdf b;
        memset(&b, 0, sizeof(b));
        const std::string t = Lowering::type_of(*n);
        if (t == "diffuse") {
            b.kind = NLPF_BSDF_DIFFUSE;
        } else if (t == "emission") {
            b.kind = NLPF_BSDF_EMISSION;
        } else if (t == "principled") {
            b.kind = NLPF_BSDF_PRINCIPLED;
            if (!lw.scalar(*n, "metallic", b.metallic) || !lw.scalar(*n, "roughness", b.roughness) ||
                !lw.color(*n, "specular_tint", b.specular_tint) ||
                !lw.scalar(*n, "specular_ior_level", b.specular_ior_level) ||
                !lw.scalar(*n, "transmission_weight", b.transmission) ||
                !lw.scalar(*n, "coat_weight", b.coat_weight))
                return fail("", NLPF_E_UNSUPPORTED);
            if (b.coat_weight != 0.0f)
                return fail("principled coat (coat_weight != 0) needs the reference's precomputed albedo table",
                            NLPF_E_UNSUPPORTED);
            if (b.metallic < 1.0f - 1e-4f && (b.specular_ior_level != 0.0f || b.transmission != 0.0f))
                return fail("principled with metallic < 1 needs specular_ior_level 0 and transmission 0",
                            NLPF_E_UNSUPPORTED);
            // normal_map with a zero normal keeps the shading frame (principled.rs:313-321,
            // mod.rs:2469-2506); anything else is a normal map
            float nrm[3];
            if (!lw.color(*n, "normal", nrm)) return fail("", NLPF_E_UNSUPPORTED);
            if (nrm[0] != 0.0f || nrm[1] != 0.0f || nrm[2] != 0.0f)
                return fail("principled normal map is not supported", NLPF_E_UNSUPPORTED);
        } else {
            return fail("surface '" + t + "' is not supported by the world filter", NLPF_E_UNSUPPORTED);
        }
        out[m] = b;
    }
    return NLPF_OK;
}

extern "C" int nlpf_bsdfs_from_scene(const char* scene_json, const char* const* materials, uint32_t n_materials,
                                     nlpf_bsdf* out) {
    try {
        return bsdfs_from_scene(scene_json, materials, n_materials, out);
    } catch (const std::bad_alloc&) {
        set_global_error("nlpf_bsdfs_from_scene: out of host memory");
        return NLPF_E_OOM;
    }
}
