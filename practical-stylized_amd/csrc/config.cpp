// Parser for the `"type":"pf"` method JSON (serde schema of
// crates/akari_integrator/src/lm_path_filtering.rs:264-298, #[serde(default)] on every field).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <sstream>
#include "common.hpp"
#include "json.hpp"

namespace nlpf {
namespace {

bool get_uint(const JVal& v, const char* key, uint64_t& out, std::string& err) {
    if (v.kind != JVal::Num || v.num < 0 || std::floor(v.num) != v.num) {
        err = std::string("field '") + key + "' must be a non-negative integer";
        return false;
    }
    out = (uint64_t)v.num;
    return true;
}
bool get_bool(const JVal& v, const char* key, bool& out, std::string& err) {
    if (v.kind != JVal::Bool) {
        err = std::string("field '") + key + "' must be a boolean";
        return false;
    }
    out = v.b;
    return true;
}
bool get_float(const JVal& v, const char* key, float& out, std::string& err) {
    if (v.kind != JVal::Num) {
        err = std::string("field '") + key + "' must be a number";
        return false;
    }
    out = (float)v.num;
    return true;
}

}  // namespace

bool parse_config(const char* json, Config& cfg, std::string& err) {
    if (!json) json = "{}";
    JParser p{json, ""};
    JVal root;
    if (!p.parse(root)) {
        err = "config JSON: " + p.err;
        return false;
    }
    p.ws();
    if (*p.s) {
        err = "config JSON: trailing characters";
        return false;
    }
    if (root.kind != JVal::Obj) {
        err = "config JSON: top level must be an object";
        return false;
    }
    Config c;
    for (auto& kv : root.obj) {
        const std::string& k = kv.first;
        const JVal& v = kv.second;
        uint64_t u = 0;
        if (k == "type") {
            if (v.kind != JVal::Str || v.str != "pf") {
                err = "field 'type' must be \"pf\"";
                return false;
            }
        } else if (k == "max_depth") {
            if (!get_uint(v, "max_depth", u, err)) return false;
            if (u > 30) {
                err = "max_depth must be <= 30";
                return false;
            }
            c.max_depth = (uint32_t)u;
        } else if (k == "repeats") {
            if (!get_uint(v, "repeats", u, err)) return false;
            c.repeats = (uint32_t)u;
        } else if (k == "branch") {
            if (!get_uint(v, "branch", u, err)) return false;
            if (u == 0) {
                err = "branch must be >= 1";
                return false;
            }
            c.branch = (uint32_t)u;
        } else if (k == "radius_scale") {
            if (v.kind != JVal::Arr || v.arr.empty()) {
                err = "field 'radius_scale' must be a non-empty array of numbers";
                return false;
            }
            c.radius_scale.clear();
            for (auto& e : v.arr) {
                float f;
                if (!get_float(e, "radius_scale", f, err)) return false;
                if (!(f > 0)) {
                    err = "radius_scale entries must be > 0";
                    return false;
                }
                c.radius_scale.push_back(f);
            }
        } else if (k == "reuse") {
            if (v.kind != JVal::Arr || v.arr.empty()) {
                err = "field 'reuse' must be a non-empty array of integers";
                return false;
            }
            c.reuse.clear();
            for (auto& e : v.arr) {
                if (!get_uint(e, "reuse", u, err)) return false;
                if (u == 0 || u > (1u << 20)) {
                    err = "reuse entries must be in [1, 2^20]";
                    return false;
                }
                c.reuse.push_back((uint32_t)u);
            }
        } else if (k == "vram") {
            if (!get_uint(v, "vram", u, err)) return false;
            if (u == 0) {
                err = "vram must be >= 1 (MB)";
                return false;
            }
            c.vram = u;
        } else if (k == "smis") {
            if (!get_bool(v, "smis", c.smis, err)) return false;
        } else if (k == "no_first_hit_stylization") {
            if (!get_bool(v, "no_first_hit_stylization", c.no_first_hit_stylization, err)) return false;
        } else if (k == "disable_flag_check") {
            if (!get_bool(v, "disable_flag_check", c.disable_flag_check, err)) return false;
        } else if (k == "cluster") {
            if (!get_bool(v, "cluster", c.cluster, err)) return false;
        } else if (k == "scene_bounds") {
            if (v.kind == JVal::Null) {
                c.has_scene_bounds = false;
            } else {
                if (v.kind != JVal::Arr || v.arr.size() != 2) {
                    err = "scene_bounds must be null or [[x,y,z],[x,y,z]]";
                    return false;
                }
                for (int i = 0; i < 2; ++i) {
                    const JVal& a = v.arr[i];
                    if (a.kind != JVal::Arr || a.arr.size() != 3) {
                        err = "scene_bounds must be null or [[x,y,z],[x,y,z]]";
                        return false;
                    }
                    for (int j = 0; j < 3; ++j)
                        if (!get_float(a.arr[j], "scene_bounds", c.scene_bounds[i][j], err)) return false;
                }
                for (int j = 0; j < 3; ++j)
                    if (!(c.scene_bounds[1][j] > c.scene_bounds[0][j])) {
                        err = "scene_bounds max must exceed min on every axis";
                        return false;
                    }
                c.has_scene_bounds = true;
            }
        } else if (k == "mode") {
            if (v.kind != JVal::Str || (v.str != "world" && v.str != "screen")) {
                err = "mode must be \"world\" or \"screen\"";
                return false;
            }
            c.mode = v.str;
        } else if (k == "pixel_radius") {
            if (!get_uint(v, "pixel_radius", u, err)) return false;
            if (u < 1 || u > 7) {
                err = "pixel_radius must be in [1, 7]";
                return false;
            }
            c.pixel_radius = (int)u;
        } else if (k == "sigma_pos" || k == "sigma_normal" || k == "sigma_lum") {
            float f;
            if (!get_float(v, k.c_str(), f, err)) return false;
            if (!(f > 0) || !std::isfinite(f)) {
                err = k + " must be a positive finite number";
                return false;
            }
            (k == "sigma_pos" ? c.sigma_pos : k == "sigma_normal" ? c.sigma_normal : c.sigma_lum) = f;
        } else if (k == "precision") {
            if (v.kind != JVal::Str) {
                err = "precision must be \"fp32\", \"fp16\" or \"fp64\"";
                return false;
            }
            if (v.str == "fp32") c.precision = NLPF_FP32;
            else if (v.str == "fp16") c.precision = NLPF_FP16;
            else if (v.str == "fp64") c.precision = NLPF_FP64;
            else {
                err = "precision must be \"fp32\", \"fp16\" or \"fp64\"";
                return false;
            }
        } else if (k == "seed") {
            if (!get_uint(v, "seed", u, err)) return false;
            c.seed = u;
        } else if (k == "hash_table_size") {
            if (!get_uint(v, "hash_table_size", u, err)) return false;
            if (u < 1 || u > (1ull << 30)) {
                err = "hash_table_size must be in [1, 2^30]";
                return false;
            }
            c.hash_table_size = u;
        } else if (k == "world_kernel") {
            if (v.kind != JVal::Str || (v.str != "auto" && v.str != "thread" && v.str != "wave")) {
                err = "world_kernel must be \"auto\", \"thread\" or \"wave\"";
                return false;
            }
            c.world_kernel = v.str;
        } else if (k == "world_grid") {
            if (v.kind != JVal::Str || (v.str != "auto" && v.str != "sort")) {
                err = "world_grid must be \"auto\" or \"sort\"";
                return false;
            }
            c.world_grid = v.str;
        } else if (k == "screen_kernel") {
            if (v.kind != JVal::Str || (v.str != "auto" && v.str != "tiles" && v.str != "strips")) {
                err = "screen_kernel must be \"auto\", \"tiles\" or \"strips\"";
                return false;
            }
            c.screen_kernel = v.str;
        }
        // other keys (e.g. "final_gather" in configs/teaser/nl-pf.json) are ignored like serde does
    }
    cfg = c;
    return true;
}

std::string config_to_json(const Config& c) {
    std::ostringstream o;
    o.precision(9);
    o << "{\"type\":\"pf\",\"max_depth\":" << c.max_depth << ",\"repeats\":" << c.repeats
      << ",\"branch\":" << c.branch << ",\"radius_scale\":[";
    for (size_t i = 0; i < c.radius_scale.size(); ++i) o << (i ? "," : "") << c.radius_scale[i];
    o << "],\"reuse\":[";
    for (size_t i = 0; i < c.reuse.size(); ++i) o << (i ? "," : "") << c.reuse[i];
    o << "],\"vram\":" << c.vram << ",\"smis\":" << (c.smis ? "true" : "false")
      << ",\"no_first_hit_stylization\":" << (c.no_first_hit_stylization ? "true" : "false")
      << ",\"disable_flag_check\":" << (c.disable_flag_check ? "true" : "false")
      << ",\"cluster\":" << (c.cluster ? "true" : "false") << ",\"scene_bounds\":";
    if (c.has_scene_bounds) {
        o << "[[" << c.scene_bounds[0][0] << "," << c.scene_bounds[0][1] << "," << c.scene_bounds[0][2]
          << "],[" << c.scene_bounds[1][0] << "," << c.scene_bounds[1][1] << "," << c.scene_bounds[1][2] << "]]";
    } else {
        o << "null";
    }
    o << ",\"mode\":\"" << c.mode << "\",\"pixel_radius\":" << c.pixel_radius << ",\"sigma_pos\":" << c.sigma_pos
      << ",\"sigma_normal\":" << c.sigma_normal << ",\"sigma_lum\":" << c.sigma_lum << ",\"precision\":\""
      << (c.precision == NLPF_FP16 ? "fp16" : c.precision == NLPF_FP64 ? "fp64" : "fp32")
      << "\",\"seed\":" << c.seed << ",\"hash_table_size\":" << c.hash_table_size << ",\"world_kernel\":\""
      << c.world_kernel << "\",\"world_grid\":\"" << c.world_grid << "\",\"screen_kernel\":\"" << c.screen_kernel
      << "\"}";
    return o.str();
}

}  // namespace nlpf
