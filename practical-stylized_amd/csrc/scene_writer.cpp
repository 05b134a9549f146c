// Writes the procedural box scene of nlpf_generate_world_records as a reference scene file:
// akari_scenegraph JSON (akari_scenegraph/src/scene.rs:121-129, 368-383) plus one binary buffer
// next to it, which MmapScene::open resolves relative to the JSON's directory
// (scene.rs:641-690). Geometry is triangle meshes: the walls and the light are exact quads,
// and the spheres are UV spheres with their analytic normals at the vertices. Materials are
// diffuse with the generator's reflectances, and the light is an emission node (colour 1,
// strength 15). The style table is lifted back into sty_* shader nodes, the inverse of
// nlpf_style_from_scene. The camera is the generator's pinhole, expressed in the reference's
// PerspectiveCamera conventions (camera/mod.rs:130-152).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <map>
#include <string>
#include <vector>
#include "common.hpp"
#include "scene.hpp"

namespace nlpf {

namespace {

std::string fnum(double v) {
    char b[40];
    snprintf(b, sizeof b, "%.9g", (double)(float)v);
    return b;
}
std::string jstr(const std::string& s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o + "\"";
}
std::string jref(const std::string& id) { return "{\"id\": " + jstr(id) + "}"; }
std::string vec3s(const float* v) { return "[" + fnum(v[0]) + ", " + fnum(v[1]) + ", " + fnum(v[2]) + "]"; }

// one mesh: positions, normals, triangles (u32 x3), all material slot 0
struct MeshData {
    std::vector<float> pos, nrm;
    std::vector<uint32_t> tri;
    uint32_t vert(float x, float y, float z, float nx, float ny, float nz) {
        pos.insert(pos.end(), {x, y, z});
        nrm.insert(nrm.end(), {nx, ny, nz});
        return (uint32_t)(pos.size() / 3 - 1);
    }
    // triangle whose geometric normal (v1 - v0) x (v2 - v0) points along n
    void face(uint32_t a, uint32_t b, uint32_t c, const float n[3]) {
        const float* p = pos.data();
        float e1[3], e2[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = p[3 * b + k] - p[3 * a + k];
            e2[k] = p[3 * c + k] - p[3 * a + k];
        }
        const float cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                    cz = e1[0] * e2[1] - e1[1] * e2[0];
        if (cx * n[0] + cy * n[1] + cz * n[2] >= 0.0f)
            tri.insert(tri.end(), {a, b, c});
        else
            tri.insert(tri.end(), {a, c, b});
    }
};

// axis-aligned quad on plane coordinate[axis] = c with the in-plane half extents, facing n
MeshData quad(int axis, float c, float h0, float h1, const float n[3]) {
    MeshData m;
    const int a0 = axis == 0 ? 1 : 0, a1 = axis == 2 ? 1 : 2;
    uint32_t id[4];
    for (int k = 0; k < 4; ++k) {
        float p[3];
        p[axis] = c;
        p[a0] = (k == 1 || k == 2) ? h0 : -h0;
        p[a1] = (k >= 2) ? h1 : -h1;
        id[k] = m.vert(p[0], p[1], p[2], n[0], n[1], n[2]);
    }
    m.face(id[0], id[1], id[2], n);
    m.face(id[0], id[2], id[3], n);
    return m;
}

// UV sphere: `seg` segments around, seg/2 rings, outward normals
MeshData uv_sphere(float cx, float cy, float cz, float r, uint32_t seg) {
    MeshData m;
    const uint32_t rings = seg / 2;
    const double pi = 3.14159265358979323846;
    std::vector<uint32_t> grid((rings + 1) * seg);
    for (uint32_t i = 0; i <= rings; ++i) {
        const double th = pi * i / rings;
        for (uint32_t j = 0; j < seg; ++j) {
            const double ph = 2 * pi * j / seg;
            const double nx = std::sin(th) * std::cos(ph), ny = std::cos(th), nz = std::sin(th) * std::sin(ph);
            grid[i * seg + j] = (i == 0 || i == rings) && j > 0
                                    ? grid[i * seg]  // one vertex per pole
                                    : m.vert((float)(cx + r * nx), (float)(cy + r * ny), (float)(cz + r * nz),
                                             (float)nx, (float)ny, (float)nz);
        }
    }
    for (uint32_t i = 0; i < rings; ++i)
        for (uint32_t j = 0; j < seg; ++j) {
            const uint32_t a = grid[i * seg + j], b = grid[i * seg + (j + 1) % seg];
            const uint32_t c = grid[(i + 1) * seg + j], d = grid[(i + 1) * seg + (j + 1) % seg];
            // outward direction at the quad's centre
            const double th = pi * (i + 0.5) / rings, ph = 2 * pi * (j + 0.5) / seg;
            const float n[3] = {(float)(std::sin(th) * std::cos(ph)), (float)std::cos(th),
                                (float)(std::sin(th) * std::sin(ph))};
            if (i > 0) m.face(a, b, d, n);
            if (i + 1 < rings) m.face(a, d, c, n);
        }
    return m;
}

struct Lift {
    const nlpf_style_node* nodes;
    uint32_t n_nodes;
    std::string err;
    std::vector<std::pair<std::string, std::string>> out;  // node id, JSON
    std::map<std::pair<int, bool>, std::string> memo;
    int consts = 0;

    std::string constant(float v) {
        const std::string id = "c" + std::to_string(consts++);
        out.emplace_back(id, "{\"type\": \"float\", \"value\": " + fnum(v) + "}");
        return id;
    }
    std::string color(const float* v) {
        const std::string id = "c" + std::to_string(consts++);
        out.emplace_back(id, "{\"type\": \"rgb\", \"value\": " + vec3s(v) + ", \"colorspace\": \"srgb\"}");
        out.emplace_back(id + "_s", "{\"type\": \"spectral_uplift\", \"rgb\": " + jref(id) + "}");
        return id + "_s";
    }
    // node i of the table as a shader node id; the bottom of a chain is the material's BSDF,
    // except below an enable_if's else branch, whose surface is never evaluated as a BSDF
    // (svm/surface/mod.rs:866-940 use inner only), which gets the null surface like the
    // reference's scenes
    bool lift(int i, bool in_else, std::string& id, int depth) {
        if (i < 0) {
            id = in_else ? "null" : "bsdf";
            return true;
        }
        if (depth > NLPF_STYLE_STACK || (uint32_t)i >= n_nodes) {
            err = "style table: bad node index or chain deeper than 16";
            return false;
        }
        auto it = memo.find({i, in_else});
        if (it != memo.end()) {
            id = it->second;
            return true;
        }
        const nlpf_style_node& n = nodes[i];
        if (n.depth_mask != 0xFFFFFFFFu) {
            err = "node " + std::to_string(i) + ": depth gating has no counterpart in the reference's scene format";
            return false;
        }
        std::string inner;
        if (!lift(n.inner, in_else, inner, depth + 1)) return false;
        std::string j;
        const float* p = n.p;
        switch (n.type) {
            case NLPF_STY_BASE: id = inner; return true;
            case NLPF_STY_TOON:
                j = "{\"type\": \"sty_toon\", \"surface\": " + jref(inner) + ", \"color_band_width\": " +
                    jref(constant(p[0])) + "}";
                break;
            case NLPF_STY_EIGHT_BIT:
                j = "{\"type\": \"sty_8bit\", \"surface\": " + jref(inner) + ", \"indirect_scale\": " +
                    jref(constant(p[0])) + "}";
                break;
            case NLPF_STY_SATURATION:
                j = "{\"type\": \"sty_saturation_adjust\", \"surface\": " + jref(inner) + ", \"power\": " +
                    jref(constant(p[0])) + ", \"indirect_scale\": " + jref(constant(p[1])) +
                    ", \"max_depth\": " + std::to_string(n.max_depth) + "}";
                break;
            case NLPF_STY_REX:
                j = "{\"type\": \"sty_saturation_adjust_rex\", \"surface\": " + jref(inner) + ", \"blend\": " +
                    jref(constant(p[0])) + "}";
                break;
            case NLPF_STY_PALETTE: {
                std::string pal;
                for (uint32_t k = 0; k < n.count; ++k) {
                    uint32_t x = 0;
                    for (int c = 0; c < 3; ++c) {
                        const float v = p[2 + 3 * k + c];
                        const long q = std::lround((double)v * 255.0);
                        if (q < 0 || q > 255 || (float)q / 255.0f != v) {
                            err = "node " + std::to_string(i) + ": palette entry is not a multiple of 1/255";
                            return false;
                        }
                        x = (x << 8) | (uint32_t)q;
                    }
                    char b[16];
                    snprintf(b, sizeof b, "\"0x%06x\"", x);
                    pal += (k ? ", " : "") + std::string(b);
                }
                j = "{\"type\": \"sty_color_palette\", \"surface\": " + jref(inner) + ", \"palette\": [" + pal +
                    "], \"gamma\": " + jref(constant(p[1])) + ", \"indirect_scale\": " + jref(constant(p[0])) + "}";
                break;
            }
            case NLPF_STY_HALFTONE: {
                std::string trs;
                if (!to_trs(p + 11, trs)) {
                    err = "node " + std::to_string(i) + ": " + err;
                    return false;
                }
                j = "{\"type\": \"sty_half_tone\", \"surface\": " + jref(inner) + ", \"trs\": " + trs +
                    ", \"radius_scale\": " + jref(constant(p[1])) + ", \"radius\": " + jref(constant(p[2])) +
                    ", \"indirect_scale\": " + jref(constant(p[0])) + ", \"grid_distance\": " +
                    jref(constant(p[3])) + ", \"max_depth\": " + std::to_string(n.max_depth) +
                    ", \"blend\": " + jref(constant(p[4])) + ", \"light_tone\": " + jref(color(p + 8)) +
                    ", \"dark_tone\": " + jref(color(p + 5)) + "}";
                break;
            }
            case NLPF_STY_SET_IDX:
                j = "{\"type\": \"sty_set_idx\", \"surface\": " + jref(inner) + ", \"idx\": " + std::to_string(n.idx) +
                    "}";
                break;
            case NLPF_STY_ENABLE_IF: {
                j = "{\"type\": \"sty_enable_if\", \"surface\": " + jref(inner) + ", \"idx\": " + std::to_string(n.idx);
                if (n.else_ >= 0) {
                    std::string e;
                    if (!lift(n.else_, true, e, depth + 1)) return false;
                    j += ", \"else\": " + jref(e);
                }
                j += "}";
                break;
            }
            default:
                err = "node " + std::to_string(i) +
                      ": hue shift and colour ramp have no compiled counterpart in the reference (svm/compiler.rs:478)";
                return false;
        }
        id = "sty_" + std::to_string(i) + (in_else ? "_e" : "");
        out.emplace_back(id, j);
        memo[{i, in_else}] = id;
        return true;
    }
    // affine 3x4 = T R S with R = Ry(b) Rx(a) Rz(c): the Akari TRS of scene.rs:42-47
    bool to_trs(const float* a, std::string& out_trs) {
        double M[3][3], s[3], R[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) M[r][c] = a[4 * r + c];
        for (int c = 0; c < 3; ++c) {
            s[c] = std::sqrt(M[0][c] * M[0][c] + M[1][c] * M[1][c] + M[2][c] * M[2][c]);
            if (!(s[c] > 0)) {
                err = "half tone affine is singular";
                return false;
            }
            for (int r = 0; r < 3; ++r) R[r][c] = M[r][c] / s[c];
        }
        for (int c0 = 0; c0 < 3; ++c0)
            for (int c1 = c0; c1 < 3; ++c1) {
                double d = 0;
                for (int r = 0; r < 3; ++r) d += R[r][c0] * R[r][c1];
                if (std::fabs(d - (c0 == c1 ? 1.0 : 0.0)) > 1e-5) {
                    err = "half tone affine is not translation * rotation * scale";
                    return false;
                }
            }
        const double det = R[0][0] * (R[1][1] * R[2][2] - R[1][2] * R[2][1]) -
                           R[0][1] * (R[1][0] * R[2][2] - R[1][2] * R[2][0]) +
                           R[0][2] * (R[1][0] * R[2][1] - R[1][1] * R[2][0]);
        if (det < 0) {
            err = "half tone affine mirrors (negative scale)";
            return false;
        }
        const double ra = std::asin(std::fmax(-1.0, std::fmin(1.0, -R[1][2])));
        const double rc = std::atan2(R[1][0], R[1][1]);
        const double rb = std::atan2(R[0][2], R[2][2]);
        const float t[3] = {a[3], a[7], a[11]}, rot[3] = {(float)ra, (float)rb, (float)rc},
                    sc[3] = {(float)s[0], (float)s[1], (float)s[2]};
        out_trs = "{\"translation\": " + vec3s(t) + ", \"rotation\": " + vec3s(rot) + ", \"scale\": " + vec3s(sc) +
                  ", \"coordinate_system\": \"Akari\"}";
        return true;
    }
};

const char* kMatNames[NLPF_SCENE_MATERIALS] = {"floor", "ceiling", "back", "left", "right", "sphere0", "sphere1",
                                               "light"};

}  // namespace

}  // namespace nlpf

using namespace nlpf;

static int write_world_scene(const char* path, uint32_t width, uint32_t height, const nlpf_gen_desc* gen,
                             const nlpf_style_node* nodes, uint32_t n_nodes, const int32_t* roots,
                             uint32_t n_materials, uint32_t sphere_segments, const nlpf_bsdf* bsdfs,
                             uint32_t n_bsdfs) {
    auto fail = [](int code, const std::string& m) {
        set_global_error("nlpf_write_world_scene: " + m);
        return code;
    };
    if (!path || !*path || !gen || (n_nodes && !nodes) || (n_materials && !roots) || (n_bsdfs && !bsdfs))
        return fail(NLPF_E_INVALID_ARG, "NULL argument");
    for (uint32_t m = 0; m < n_bsdfs && m < (uint32_t)NLPF_SCENE_MATERIALS; ++m) {
        const nlpf_bsdf& b = bsdfs[m];
        if (b.kind > NLPF_BSDF_EMISSION) return fail(NLPF_E_INVALID_ARG, "unknown BSDF kind");
        if (m == NLPF_MAT_LIGHT && b.kind != NLPF_BSDF_EMISSION && b.kind != NLPF_BSDF_DIFFUSE)
            return fail(NLPF_E_INVALID_ARG, "the light material stays an emission surface");
    }
    if (width == 0 || height == 0) return fail(NLPF_E_INVALID_ARG, "width and height must be >= 1");
    if (sphere_segments < 4 || sphere_segments > 4096 || sphere_segments % 2)
        return fail(NLPF_E_INVALID_ARG, "sphere_segments must be even, 4..4096");
    typedef SceneConst C;
    const std::filesystem::path jp(path);
    const std::string stem = jp.stem().string();
    const std::filesystem::path bp = jp.parent_path() / (stem + ".bin");

    // geometry, in material order
    std::vector<MeshData> meshes;
    const float up[3] = {0, 1, 0}, dn[3] = {0, -1, 0}, fw[3] = {0, 0, 1}, px[3] = {1, 0, 0}, nx[3] = {-1, 0, 0};
    meshes.push_back(quad(1, -C::h, C::h, C::h, up));    // floor, facing +y (into the box)
    meshes.push_back(quad(1, C::h, C::h, C::h, dn));     // ceiling
    meshes.push_back(quad(2, -C::h, C::h, C::h, fw));    // back
    meshes.push_back(quad(0, -C::h, C::h, C::h, px));    // left
    meshes.push_back(quad(0, C::h, C::h, C::h, nx));     // right
    meshes.push_back(uv_sphere(C::s0x, C::s0y, C::s0z, C::s0r, sphere_segments));
    meshes.push_back(uv_sphere(C::s1x, C::s1y, C::s1z, C::s1r, sphere_segments));
    meshes.push_back(quad(1, C::ly, C::lh, C::lh, dn));  // light, emitting downwards

    // one buffer: per mesh vertices | normals | indices | per-triangle material slots
    std::vector<uint8_t> blob;
    std::string views, geoms, insts;
    auto add_view = [&](const std::string& id, const void* data, size_t bytes) {
        const size_t off = blob.size();
        blob.insert(blob.end(), (const uint8_t*)data, (const uint8_t*)data + bytes);
        views += (views.empty() ? "" : ",\n") + jstr(id) + ": {\"buffer\": " + jref(stem) +
                 ", \"offset\": " + std::to_string(off) + ", \"length\": " + std::to_string(bytes) + "}";
    };
    for (int m = 0; m < NLPF_SCENE_MATERIALS; ++m) {
        const MeshData& md = meshes[m];
        const std::string g = "geom" + std::to_string(m) + "_" + kMatNames[m];
        std::vector<uint32_t> slots(md.tri.size() / 3, 0u);
        add_view(g + "_vertices", md.pos.data(), md.pos.size() * 4);
        add_view(g + "_normals", md.nrm.data(), md.nrm.size() * 4);
        add_view(g + "_indices", md.tri.data(), md.tri.size() * 4);
        add_view(g + "_materials", slots.data(), slots.size() * 4);
        geoms += (geoms.empty() ? "" : ",\n") + jstr(g) + ": {\"type\": \"mesh\", \"vertices\": " +
                 jref(g + "_vertices") + ", \"indices\": " + jref(g + "_indices") + ", \"normals\": " +
                 jref(g + "_normals") + ", \"uvs\": null, \"tangents\": null, \"materials\": " +
                 jref(g + "_materials") + "}";
        insts += (insts.empty() ? "" : ",\n") + jstr("inst" + std::to_string(m) + "_" + kMatNames[m]) +
                 ": {\"geometry\": " + jref(g) +
                 ", \"transform\": {\"type\": \"matrix\", \"data\": [[1.0, 0.0, 0.0, 0.0], [0.0, 1.0, 0.0, 0.0], "
                 "[0.0, 0.0, 1.0, 0.0], [0.0, 0.0, 0.0, 1.0]]}, \"materials\": [" +
                 jref("mat" + std::to_string(m) + "_" + kMatNames[m]) + "]}";
    }

    // materials: names sort in material-id order, so the importer's default order matches
    std::string mats;
    const uint64_t key = splitmix64(gen->seed);
    for (int m = 0; m < NLPF_SCENE_MATERIALS; ++m) {
        Lift lf{nodes, n_nodes, "", {}, {}, 0};
        std::string body;
        if (m == NLPF_MAT_LIGHT) {
            const float one[3] = {1, 1, 1};
            lf.out.emplace_back("ecol", "{\"type\": \"rgb\", \"value\": " + vec3s(one) + ", \"colorspace\": \"srgb\"}");
            lf.out.emplace_back("ecol_s", "{\"type\": \"spectral_uplift\", \"rgb\": " + jref("ecol") + "}");
            lf.out.emplace_back("estr", "{\"type\": \"float\", \"value\": " + fnum(NLPF_LIGHT_EMISSION) + "}");
            lf.out.emplace_back("bsdf", "{\"type\": \"emission\", \"color\": " + jref("ecol_s") + ", \"strength\": " +
                                            jref("estr") + "}");
        } else {
            const v3<float> R = material_albedo(m, key, gen->albedo_lo, gen->albedo_hi);
            const float a[3] = {R.x, R.y, R.z};
            lf.out.emplace_back("albedo", "{\"type\": \"rgb\", \"value\": " + vec3s(a) + ", \"colorspace\": \"srgb\"}");
            lf.out.emplace_back("albedo_s", "{\"type\": \"spectral_uplift\", \"rgb\": " + jref("albedo") + "}");
            const nlpf_bsdf* b = (uint32_t)m < n_bsdfs ? &bsdfs[m] : nullptr;
            if (b && b->kind == NLPF_BSDF_PRINCIPLED) {
                // every input of the reference's principled node (shader.rs), the ones the
                // closure does not read at Blender's defaults
                auto flt = [&](const char* id, float v) {
                    lf.out.emplace_back(id, "{\"type\": \"float\", \"value\": " + fnum(v) + "}");
                };
                auto col = [&](const char* id, const float c[3]) {
                    const std::string raw = std::string(id) + "_rgb";
                    lf.out.emplace_back(raw, "{\"type\": \"rgb\", \"value\": " + vec3s(c) + ", \"colorspace\": \"srgb\"}");
                    lf.out.emplace_back(id, "{\"type\": \"spectral_uplift\", \"rgb\": " + jref(raw) + "}");
                };
                auto f3 = [&](const char* id, const float c[3]) {
                    lf.out.emplace_back(id, "{\"type\": \"float3\", \"value\": " + vec3s(c) + "}");
                };
                const float one3[3] = {1, 1, 1}, zero3[3] = {0, 0, 0}, ssr[3] = {1.0f, 0.2f, 0.1f};
                flt("p_metallic", b->metallic);
                flt("p_roughness", b->roughness);
                col("p_spec_tint", b->specular_tint);
                flt("p_spec_level", b->specular_ior_level);
                flt("p_transmission", b->transmission);
                flt("p_coat_weight", b->coat_weight);
                f3("p_zero3", zero3);
                flt("p_zero", 0.0f);
                flt("p_one", 1.0f);
                flt("p_ior", 1.45f);
                flt("p_coat_ior", 1.5f);
                flt("p_coat_rough", 0.03f);
                flt("p_ss_scale", 0.05f);
                col("p_white", one3);
                f3("p_ss_radius", ssr);
                const auto r = [&](const char* id) { return jref(id); };
                lf.out.emplace_back(
                    "bsdf", std::string("{\"type\": \"principled\", \"preference\": \"mix\", \"base_color\": ") +
                                r("albedo_s") + ", \"metallic\": " + r("p_metallic") + ", \"roughness\": " +
                                r("p_roughness") + ", \"specular_tint\": " + r("p_spec_tint") +
                                ", \"specular_ior_level\": " + r("p_spec_level") + ", \"transmission_weight\": " +
                                r("p_transmission") + ", \"coat_weight\": " + r("p_coat_weight") +
                                ", \"normal\": " + r("p_zero3") + ", \"coat_normal\": " + r("p_zero3") +
                                ", \"tangent\": " + r("p_zero3") + ", \"alpha\": " + r("p_one") + ", \"ior\": " +
                                r("p_ior") + ", \"coat_ior\": " + r("p_coat_ior") + ", \"coat_roughness\": " +
                                r("p_coat_rough") + ", \"coat_tint\": " + r("p_white") + ", \"emission_color\": " +
                                r("p_white") + ", \"emission_strength\": " + r("p_zero") + ", \"sheen_weight\": " +
                                r("p_zero") + ", \"sheen_tint\": " + r("p_white") + ", \"subsurface_weight\": " +
                                r("p_zero") + ", \"subsurface_radius\": " + r("p_ss_radius") +
                                ", \"subsurface_scale\": " + r("p_ss_scale") + ", \"subsurface_anisotropy\": " +
                                r("p_zero") + ", \"anisotropic\": " + r("p_zero") + ", \"anisotropic_rotation\": " +
                                r("p_zero") + "}");
            } else if (b && b->kind == NLPF_BSDF_EMISSION) {
                // an emission-only surface of strength 0: it reflects nothing and emits nothing
                lf.out.emplace_back("estr", "{\"type\": \"float\", \"value\": 0.0}");
                lf.out.emplace_back("bsdf", "{\"type\": \"emission\", \"color\": " + jref("albedo_s") +
                                                ", \"strength\": " + jref("estr") + "}");
            } else {
                lf.out.emplace_back("bsdf", "{\"type\": \"diffuse\", \"color\": " + jref("albedo_s") + "}");
            }
        }
        lf.out.emplace_back("null", "{\"type\": \"null\"}");
        const int root = (uint32_t)m < n_materials ? roots[m] : -1;
        std::string top;
        if (!lf.lift(root, false, top, 0))
            return fail(NLPF_E_UNSUPPORTED, std::string("material ") + kMatNames[m] + ": " + lf.err);
        lf.out.emplace_back("out", "{\"type\": \"output\", \"node\": " + jref(top) + "}");
        for (size_t k = 0; k < lf.out.size(); ++k)
            body += (k ? ",\n" : "") + jstr(lf.out[k].first) + ": " + lf.out[k].second;
        mats += (mats.empty() ? "" : ",\n") + jstr("mat" + std::to_string(m) + "_" + kMatNames[m]) +
                ": {\"shader\": {\"nodes\": {\n" + body + "}, \"output\": " + jref("out") + ", \"kind\": \"surface\"}}";
    }

    // pinhole at (0, 0, cam_z) looking down -z; the reference scales the larger image side by
    // tan(fov / 2) (camera/mod.rs:145-150) where the generator scales the vertical side
    const double tan_half = C::tan_half_fov * (width > height ? (double)width / height : 1.0);
    const double fov_deg = 2.0 * std::atan(tan_half) * 180.0 / 3.14159265358979323846;
    const float cam_t[3] = {0.0f, 0.0f, C::cam_z}, zero[3] = {0, 0, 0}, one[3] = {1, 1, 1};
    std::string js = "{\n\"camera\": {\"type\": \"perspective\", \"data\": {\"transform\": {\"type\": \"trs\", "
                     "\"data\": {\"translation\": " + vec3s(cam_t) + ", \"rotation\": " + vec3s(zero) +
                     ", \"scale\": " + vec3s(one) + ", \"coordinate_system\": \"Akari\"}}, \"fov\": " + fnum(fov_deg) +
                     ", \"focal_distance\": 1.0, \"fstop\": 2.8, \"sensor_width\": " + std::to_string(width) +
                     ", \"sensor_height\": " + std::to_string(height) + ", \"depth_of_field\": false}},\n";
    js += "\"instances\": {\n" + insts + "},\n";
    js += "\"geometries\": {\n" + geoms + "},\n";
    js += "\"materials\": {\n" + mats + "},\n";
    js += "\"lights\": {},\n";
    js += "\"buffers\": {" + jstr(stem) + ": {\"type\": \"path\", \"path\": " + jstr(stem + ".bin") +
          ", \"length\": " + std::to_string(blob.size()) + "}},\n";
    js += "\"buffer_views\": {\n" + views + "}\n}\n";

    std::error_code ec;
    if (!jp.parent_path().empty()) std::filesystem::create_directories(jp.parent_path(), ec);
    if (ec) return fail(NLPF_E_IO, "cannot create " + jp.parent_path().string());
    for (int f = 0; f < 2; ++f) {
        const std::string fp = f ? jp.string() : bp.string();
        FILE* fh = fopen(fp.c_str(), "wb");
        if (!fh) return fail(NLPF_E_IO, "cannot open " + fp);
        const size_t want = f ? js.size() : blob.size();
        const size_t w = fwrite(f ? (const void*)js.data() : (const void*)blob.data(), 1, want, fh);
        if (fclose(fh) != 0 || w != want) return fail(NLPF_E_IO, "short write to " + fp);
    }
    return NLPF_OK;
}

extern "C" int nlpf_write_world_scene(const char* path, uint32_t width, uint32_t height, const nlpf_gen_desc* gen,
                                      const nlpf_style_node* nodes, uint32_t n_nodes, const int32_t* roots,
                                      uint32_t n_materials, uint32_t sphere_segments, const nlpf_bsdf* bsdfs,
                                      uint32_t n_bsdfs) {
    try {
        return write_world_scene(path, width, height, gen, nodes, n_nodes, roots, n_materials, sphere_segments, bsdfs,
                                 n_bsdfs);
    } catch (const std::bad_alloc&) {
        set_global_error("nlpf_write_world_scene: out of host memory");
        return NLPF_E_OOM;
    } catch (const std::filesystem::filesystem_error& e) {
        set_global_error(std::string("nlpf_write_world_scene: ") + e.what());
        return NLPF_E_IO;
    }
}
