// Handle state shared by the C-ABI entry points and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "../../include/nlpf.h"

namespace nlpf {

// `"type":"pf"` method config (crates/akari_integrator/src/lm_path_filtering.rs:264-298) with
// the reference defaults (:280-298), plus the screen-mode fields of this library.
struct Config {
    uint32_t max_depth = 7;
    uint32_t repeats = 1;
    uint32_t branch = 1;
    std::vector<float> radius_scale{0.1f, 0.05f};
    std::vector<uint32_t> reuse{128u, 64u};
    uint64_t vram = 2048;  // MB
    bool smis = true;
    bool no_first_hit_stylization = false;
    bool disable_flag_check = false;
    bool cluster = true;
    bool has_scene_bounds = false;
    float scene_bounds[2][3] = {{0, 0, 0}, {0, 0, 0}};
    // screen mode
    std::string mode = "world";
    int pixel_radius = 7;
    float sigma_pos = 0.05f, sigma_normal = 0.9f, sigma_lum = 0.5f;
    int precision = NLPF_FP32;
    uint64_t seed = 0;
    // hash-grid bucket count; the reference fixes 64 Mi heads (lm_path_filtering.rs:1667)
    uint64_t hash_table_size = 64ull * 1024 * 1024;
    // world-mode filter kernel: "auto" (by the expected neighbour count), "thread" (one thread
    // per vertex) or "wave" (one wavefront per vertex); the results are the same
    std::string world_kernel = "auto";
    // world-mode grid build: "auto" (counting build on the bucket table, radix sort when a bucket
    // exceeds 1024 entries) or "sort" (always the radix sort); the results are identical
    std::string world_grid = "auto";
    // screen-mode level kernel: "auto" (the fastest per window radius), "tiles" (64 x 8..16
    // output tiles, each staged with its halo) or "strips" (column-strip segments with an LDS
    // ring of rows; the tiles where a launch does not qualify); the results agree within FP32
    // summation order
    std::string screen_kernel = "auto";

    // per-depth vectors repeat their last element (lm_path_filtering.rs:1337-1347, 1584-1586)
    float radius_scale_at(uint32_t d) const {
        return d < radius_scale.size() ? radius_scale[d] : radius_scale.back();
    }
    uint32_t reuse_at(uint32_t d) const { return d < reuse.size() ? reuse[d] : reuse.back(); }
};

bool parse_config(const char* json, Config& cfg, std::string& err);
std::string config_to_json(const Config& cfg);

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
};

}  // namespace nlpf

struct nlpf_mlp_state;
struct nlpf_world_state;

struct nlpf_handle {
    nlpf::Config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream of the cache pass (work that overlaps the handle's stream) and the events
    // that order the two; created on first use, default priority
    hipStream_t aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // identity permutation of an image's pixels, for the cache pass's training picks (a partial
    // shuffle that is undone after each pick)
    std::vector<uint32_t> pick_perm;
    std::string err;
    // stylization table
    nlpf_style_node* d_nodes = nullptr;
    int32_t* d_roots = nullptr;
    std::vector<nlpf_style_node> h_nodes;
    std::vector<int32_t> h_roots;
    // world-mode BSDF table (nlpf_set_bsdfs): the validated entries and their device form
    std::vector<nlpf_bsdf> h_bsdf;
    void* d_bsdf = nullptr;
    int bsdf_nondiffuse = 0;  // some entry is not NLPF_BSDF_DIFFUSE
    // timing
    std::vector<hipEvent_t> events;
    nlpf_stats stats{};
    // reusable device scratch
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // this device's static scratch of nlpf_screen_run_noalloc (not owned: process lifetime)
    void* noalloc_scratch = nullptr;
    // learned-cache and world-mode state, created on first use (owned by the handle, so
    // different handles never share state)
    nlpf_mlp_state* mlp = nullptr;
    nlpf_world_state* world = nullptr;
};

#define NLPF_FAIL(h, code, msg)          \
    do {                                 \
        (h)->err = (msg);                \
        return (code);                   \
    } while (0)

#define NLPF_HIP(h, expr)                                                                     \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            (h)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                     \
            return e_ == hipErrorOutOfMemory ? NLPF_E_OOM : NLPF_E_HIP;                       \
        }                                                                                     \
    } while (0)

namespace nlpf {
// error message of a call that has no handle, read by nlpf_last_error(NULL)
void set_global_error(const std::string& msg);
// grow-only scratch owned by the handle
int ensure_scratch(nlpf_handle* h, size_t bytes);
// record the last kernel failure (launch errors are sticky per stream)
int check_launch(nlpf_handle* h, const char* what);
int record_event(nlpf_handle* h, size_t i);
double event_ms(nlpf_handle* h, size_t i, size_t j);

inline size_t elem_size(int precision) {
    return precision == NLPF_FP16 ? 2 : (precision == NLPF_FP64 ? 8 : 4);
}
}  // namespace nlpf
