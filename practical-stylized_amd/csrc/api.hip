// C-ABI entry points of libnlpf (declared in include/nlpf.h).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include "common.hpp"
#include "screen_impl.hpp"

namespace nlpf {
hipError_t screen_level_fp32(const ScreenArgs<float>& a, int radius, int variant, hipStream_t st);
hipError_t screen_level_fp16(const ScreenArgs<float>& a, int radius, int variant, hipStream_t st);
hipError_t screen_level_fp64(const ScreenArgs<double>& a, int radius, int variant, hipStream_t st);
int screen_variants_fp32();

// level-kernel variant of a handle's screen passes: the config's screen_kernel ("auto": each
// radius's fastest, "tiles", "strips"); NLPF_SCREEN_VARIANT (0..4) forces one variant for
// every radius (for tuning)
int screen_variant_of(const Config& cfg) {
    static int forced = -2;
    if (forced == -2) {
        const char* e = getenv("NLPF_SCREEN_VARIANT");
        int v = e ? atoi(e) : -1;
        if (v < 0 || v >= screen_variants_fp32()) v = -1;
        forced = v;
    }
    if (forced >= 0) return forced;
    if (cfg.screen_kernel == "tiles") return kScreenTiles;
    if (cfg.screen_kernel == "strips") return kScreenStripsAuto;
    return kScreenAuto;
}
hipError_t screen_film_fp32(const float* out, uint32_t S, uint32_t out_levels, uint32_t slot0, uint64_t plane,
                            float* film, hipStream_t st, uint32_t s_total = 0, int accumulate = 0, int finalize = 1);
hipError_t screen_film_fp64(const double* out, uint32_t S, uint32_t out_levels, uint32_t slot0, uint64_t plane,
                            double* film, hipStream_t st, uint32_t s_total = 0, int accumulate = 0, int finalize = 1);
hipError_t launch_gen_screen(float* out, uint32_t W, uint32_t H, uint32_t L, uint32_t s_begin, uint32_t s_count,
                             const nlpf_gen_desc& gd, hipStream_t st);
hipError_t launch_pack_fp16(const float* in, __half* out, float* anchors, uint32_t W, uint32_t H, uint32_t nslices,
                            hipStream_t st);
hipError_t launch_stylize(const float* in, float* out, uint64_t n, const uint32_t* depth, uint32_t depth0,
                          const uint8_t* flags, const uint16_t* material, const float* pos,
                          const nlpf_style_node* nodes, const int32_t* roots, uint32_t n_materials, hipStream_t st);

void world_release(nlpf_handle* h);
void mlp_release(nlpf_handle* h);

int ensure_scratch(nlpf_handle* h, size_t bytes) {
    if (h->scratch_bytes >= bytes) return NLPF_OK;
    if (h->scratch) {
        NLPF_HIP(h, hipStreamSynchronize(h->stream));
        NLPF_HIP(h, hipFree(h->scratch));
        h->scratch = nullptr;
        h->scratch_bytes = 0;
    }
    NLPF_HIP(h, hipMalloc(&h->scratch, bytes));
    h->scratch_bytes = bytes;
    return NLPF_OK;
}

int check_launch(nlpf_handle* h, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        h->err = std::string(what) + ": " + hipGetErrorString(e);
        return NLPF_E_HIP;
    }
    return NLPF_OK;
}

int record_event(nlpf_handle* h, size_t i) {
    while (h->events.size() <= i) {
        hipEvent_t ev;
        NLPF_HIP(h, hipEventCreate(&ev));
        h->events.push_back(ev);
    }
    NLPF_HIP(h, hipEventRecord(h->events[i], h->stream));
    return NLPF_OK;
}

double event_ms(nlpf_handle* h, size_t i, size_t j) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, h->events[i], h->events[j]) != hipSuccess) return -1.0;
    return (double)ms;
}

static bool tree_has_type(const std::vector<nlpf_style_node>& nodes, int root, int type, int depth = 0) {
    if (root < 0 || depth > NLPF_STYLE_STACK) return false;
    const nlpf_style_node& n = nodes[root];
    if (n.type == type) return true;
    if (tree_has_type(nodes, n.inner, type, depth + 1)) return true;
    return n.type == NLPF_STY_ENABLE_IF && tree_has_type(nodes, n.else_, type, depth + 1);
}

// every branch path of the tree must end within NLPF_STYLE_STACK steps (this also rules out cycles)
static bool tree_ok(const std::vector<nlpf_style_node>& nodes, int root, int depth, std::string& err) {
    if (root < 0) return true;
    if (depth >= NLPF_STYLE_STACK) {
        err = "stylization tree deeper than 16 nodes (or cyclic)";
        return false;
    }
    const nlpf_style_node& n = nodes[root];
    if (!tree_ok(nodes, n.inner, depth + 1, err)) return false;
    if (n.type == NLPF_STY_ENABLE_IF && !tree_ok(nodes, n.else_, depth + 1, err)) return false;
    return true;
}

}  // namespace nlpf

using namespace nlpf;

static std::mutex g_create_err_mu;
static std::string g_create_err;

// Scratch of nlpf_screen_run_noalloc: every GPU the library runs on holds its own copy until
// the process ends.  In host code `scratch` names a host-side stand-in, not device memory:
// host code uses only the per-device address that hipGetSymbolAddress returns.
static __device__ __align__(256) unsigned char scratch[NLPF_NOALLOC_SCRATCH_BYTES];
static std::mutex g_noalloc_mu;
static std::vector<void*> g_noalloc_ptr;  // by device ordinal

// address of this device's copy of `scratch` (the current device), resolved once per device
static hipError_t noalloc_scratch_address(int device, void** out) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (device < 0) return hipErrorInvalidDevice;
    if ((size_t)device >= g_noalloc_ptr.size()) g_noalloc_ptr.resize((size_t)device + 1, nullptr);
    if (!g_noalloc_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_noalloc_ptr[device] = p;
    }
    *out = g_noalloc_ptr[device];
    return hipSuccess;
}

namespace nlpf {
void set_global_error(const std::string& msg) {
    std::lock_guard<std::mutex> lk(g_create_err_mu);
    g_create_err = msg;
}
}  // namespace nlpf

extern "C" {

int nlpf_create(const char* config_json, nlpf_handle** out) {
    if (!out) {
        std::lock_guard<std::mutex> lk(g_create_err_mu);
        g_create_err = "nlpf_create: out is NULL";
        return NLPF_E_INVALID_ARG;
    }
    *out = nullptr;
    Config cfg;
    std::string err;
    if (!parse_config(config_json, cfg, err)) {
        std::lock_guard<std::mutex> lk(g_create_err_mu);
        g_create_err = err;
        return NLPF_E_INVALID_CONFIG;
    }
    nlpf_handle* h = new nlpf_handle();
    h->cfg = cfg;
    hipError_t e = hipGetDevice(&h->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> lk(g_create_err_mu);
        g_create_err = std::string("HIP initialisation failed: ") + hipGetErrorString(e);
        delete h;
        return NLPF_E_HIP;
    }
    // the no-allocation entry point's static scratch on this device; a failure here only
    // disables that entry point (it reports the missing address)
    if (noalloc_scratch_address(h->device, &h->noalloc_scratch) != hipSuccess) {
        (void)hipGetLastError();
        h->noalloc_scratch = nullptr;
    }
    *out = h;
    return NLPF_OK;
}

int nlpf_destroy(nlpf_handle* h) {
    if (!h) return NLPF_E_INVALID_ARG;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    world_release(h);
    mlp_release(h);
    if (h->d_nodes) (void)hipFree(h->d_nodes);
    if (h->d_roots) (void)hipFree(h->d_roots);
    if (h->d_bsdf) (void)hipFree(h->d_bsdf);
    if (h->scratch) (void)hipFree(h->scratch);
    for (hipEvent_t ev : h->events) (void)hipEventDestroy(ev);
    if (h->aux) {
        (void)hipStreamSynchronize(h->aux);
        (void)hipStreamDestroy(h->aux);
    }
    if (h->fork) (void)hipEventDestroy(h->fork);
    if (h->join) (void)hipEventDestroy(h->join);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return NLPF_OK;
}

const char* nlpf_last_error(const nlpf_handle* h) {
    if (!h) {
        std::lock_guard<std::mutex> lk(g_create_err_mu);
        return g_create_err.c_str();
    }
    return h->err.c_str();
}

int nlpf_get_config_json(const nlpf_handle* h, char* buf, size_t cap) {
    if (!h || !buf || cap == 0) return NLPF_E_INVALID_ARG;
    std::string s = config_to_json(h->cfg);
    if (s.size() + 1 > cap) return NLPF_E_INVALID_ARG;
    memcpy(buf, s.c_str(), s.size() + 1);
    return NLPF_OK;
}

int nlpf_set_sigmas(nlpf_handle* h, float sp, float sn, float sl) {
    if (!h) return NLPF_E_INVALID_ARG;
    if (!(sp > 0) || !(sn > 0) || !(sl > 0) || !std::isfinite(sp) || !std::isfinite(sn) || !std::isfinite(sl))
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "sigmas must be positive and finite");
    h->cfg.sigma_pos = sp;
    h->cfg.sigma_normal = sn;
    h->cfg.sigma_lum = sl;
    return NLPF_OK;
}

int nlpf_alloc(nlpf_handle* h, size_t bytes, void** dptr) {
    if (!h || !dptr) return NLPF_E_INVALID_ARG;
    *dptr = nullptr;
    NLPF_HIP(h, hipMalloc(dptr, bytes ? bytes : 1));
    return NLPF_OK;
}
int nlpf_free(nlpf_handle* h, void* dptr) {
    if (!h) return NLPF_E_INVALID_ARG;
    if (!dptr) return NLPF_OK;
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    NLPF_HIP(h, hipFree(dptr));
    return NLPF_OK;
}
int nlpf_upload(nlpf_handle* h, void* dst, const void* src, size_t bytes) {
    if (!h || (!dst && bytes) || (!src && bytes)) return NLPF_E_INVALID_ARG;
    NLPF_HIP(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    return NLPF_OK;
}
int nlpf_download(nlpf_handle* h, void* dst, const void* src, size_t bytes) {
    if (!h || (!dst && bytes) || (!src && bytes)) return NLPF_E_INVALID_ARG;
    NLPF_HIP(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    return NLPF_OK;
}
int nlpf_memset(nlpf_handle* h, void* dst, int value, size_t bytes) {
    if (!h || (!dst && bytes)) return NLPF_E_INVALID_ARG;
    NLPF_HIP(h, hipMemsetAsync(dst, value, bytes, h->stream));
    return NLPF_OK;
}
int nlpf_copy(nlpf_handle* h, void* dst, const void* src, size_t bytes) {
    if (!h || (!dst && bytes) || (!src && bytes)) return NLPF_E_INVALID_ARG;
    NLPF_HIP(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, h->stream));
    return NLPF_OK;
}
int nlpf_sync(nlpf_handle* h) {
    if (!h) return NLPF_E_INVALID_ARG;
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    return NLPF_OK;
}

int nlpf_set_style(nlpf_handle* h, const nlpf_style_node* nodes, uint32_t n_nodes, const int32_t* roots,
                   uint32_t n_materials) {
    if (!h) return NLPF_E_INVALID_ARG;
    if (n_nodes > NLPF_STY_MAX_NODES) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "too many stylization nodes (max 1024)");
    if (n_materials > NLPF_STY_MAX_MATERIALS) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "too many materials (max 65536)");
    if ((n_nodes && !nodes) || (n_materials && !roots)) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "NULL style table");
    std::vector<nlpf_style_node> nv(nodes, nodes + n_nodes);
    for (uint32_t i = 0; i < n_nodes; ++i) {
        const nlpf_style_node& n = nv[i];
        if (n.type < NLPF_STY_BASE || n.type > NLPF_STY_SET_IDX)
            NLPF_FAIL(h, NLPF_E_INVALID_ARG, "unknown stylization node type at node " + std::to_string(i));
        if (n.inner < -1 || n.inner >= (int)n_nodes || n.else_ < -1 || n.else_ >= (int)n_nodes)
            NLPF_FAIL(h, NLPF_E_INVALID_ARG, "node index out of range at node " + std::to_string(i));
        if ((n.type == NLPF_STY_ENABLE_IF || n.type == NLPF_STY_SET_IDX) && n.idx > 3)
            NLPF_FAIL(h, NLPF_E_INVALID_ARG, "style index must be 0..3 at node " + std::to_string(i));
        if (n.type == NLPF_STY_PALETTE && n.count > 18)
            NLPF_FAIL(h, NLPF_E_INVALID_ARG, "palette holds at most 18 entries");
        if (n.type == NLPF_STY_RAMP) {
            if (n.count == 0 || n.count > 16) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "ramp needs 1..16 stops");
            for (uint32_t k = 1; k < n.count; ++k)
                if (!(n.p[k] >= n.p[k - 1])) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "ramp stop positions must ascend");
        }
        if (n.type == NLPF_STY_TOON && !(n.p[0] > 0)) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "toon band width must be > 0");
    }
    std::vector<int32_t> rv(roots, roots + n_materials);
    for (uint32_t m = 0; m < n_materials; ++m) {
        if (rv[m] < -1 || rv[m] >= (int)n_nodes) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "material root out of range");
        std::string err;
        if (!tree_ok(nv, rv[m], 0, err)) NLPF_FAIL(h, NLPF_E_INVALID_ARG, err);
    }
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    if (h->d_nodes) NLPF_HIP(h, hipFree(h->d_nodes));
    if (h->d_roots) NLPF_HIP(h, hipFree(h->d_roots));
    h->d_nodes = nullptr;
    h->d_roots = nullptr;
    NLPF_HIP(h, hipMalloc(&h->d_nodes, sizeof(nlpf_style_node) * (n_nodes ? n_nodes : 1)));
    NLPF_HIP(h, hipMalloc(&h->d_roots, sizeof(int32_t) * (n_materials ? n_materials : 1)));
    if (n_nodes)
        NLPF_HIP(h, hipMemcpy(h->d_nodes, nv.data(), sizeof(nlpf_style_node) * n_nodes, hipMemcpyHostToDevice));
    if (n_materials)
        NLPF_HIP(h, hipMemcpy(h->d_roots, rv.data(), sizeof(int32_t) * n_materials, hipMemcpyHostToDevice));
    h->h_nodes = nv;
    h->h_roots = rv;
    return NLPF_OK;
}

int nlpf_stylize(nlpf_handle* h, const float* in, float* out, uint64_t n, const uint32_t* depth, uint32_t depth0,
                 const uint8_t* flags, const uint16_t* material, const float* p) {
    if (!h) return NLPF_E_INVALID_ARG;
    if (n == 0) return NLPF_OK;
    if (!in || !out) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_stylize: NULL rgb buffer");
    if (h->h_roots.empty()) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_stylize: no style table set");
    NLPF_HIP(h, launch_stylize(in, out, n, depth, depth0, flags, material, p, h->d_nodes, h->d_roots,
                               (uint32_t)h->h_roots.size(), h->stream));
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    return NLPF_OK;
}

static int validate_screen_desc(nlpf_handle* h, const nlpf_screen_desc* d) {
    if (!d) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "screen desc is NULL");
    if (d->width == 0 || d->height == 0 || d->spp == 0 || d->levels == 0)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "screen desc: width, height, spp and levels must be >= 1");
    if (d->levels > 31) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "screen desc: at most 31 levels");
    if (d->radius < 1 || d->radius > 7) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "screen desc: radius must be 1..7");
    if (d->precision != NLPF_FP32 && d->precision != NLPF_FP16 && d->precision != NLPF_FP64)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "screen desc: unknown precision");
    if ((uint64_t)d->width * d->height * d->spp >= (1ull << 32))
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "screen desc: width*height*spp must be < 2^32");
    return NLPF_OK;
}

// one level launch of the screen pass over slices [s0, s0 + S) of desc's records, writing
// slot `slot` of out ([S][out_levels][3][H][W]) and reading the child level from slot `cslot`
static int screen_level_launch(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records,
                               const float* anchors, void* out, uint32_t out_levels, uint32_t s0, uint32_t S,
                               int d, const uint16_t* material = nullptr, const uint8_t* flags = nullptr) {
    const uint64_t plane = (uint64_t)desc->width * desc->height;
    const uint32_t L = desc->levels;
    const uint32_t slot = out_levels == L ? (uint32_t)d : (uint32_t)(d & 1);
    const uint32_t cslot = out_levels == L ? (uint32_t)(d + 1) : (uint32_t)((d + 1) & 1);
    const int root = h->h_roots.empty() ? -1 : h->h_roots[0];
    int need_pos = root >= 0 && tree_has_type(h->h_nodes, root, NLPF_STY_HALFTONE);
    if (material)  // any material's tree may be the one a vertex takes
        for (const int32_t r : h->h_roots) need_pos |= r >= 0 && tree_has_type(h->h_nodes, r, NLPF_STY_HALFTONE);
    // per-vertex material and flags planes [s][L][H][W], from this launch's first slice
    const size_t shade_off = (size_t)s0 * L * plane;
    const uint16_t* mat = material ? material + shade_off : nullptr;
    const uint8_t* flg = material && flags ? flags + shade_off : nullptr;
    const uint32_t n_mat = (uint32_t)h->h_roots.size();
    const uint32_t lds_nodes = h->h_nodes.size() <= kShadeLdsNodes ? (uint32_t)h->h_nodes.size() : 0u;
    const double log2e = 1.4426950408889634;
    const double kp = sqrt(log2e / (2.0 * (double)h->cfg.sigma_pos * h->cfg.sigma_pos));
    const double kn = sqrt(log2e / (2.0 * (double)h->cfg.sigma_normal * h->cfg.sigma_normal));
    const double kl = sqrt(log2e / (2.0 * (double)h->cfg.sigma_lum * h->cfg.sigma_lum));
    // records [s][L][15][H][W]; FP16 anchors [s][L][ceil(H/8)][ceil(W/8)][3]
    const uint32_t abx = (desc->width + 7) / 8, aby = (desc->height + 7) / 8;
    const char* rec = (const char*)records + (size_t)s0 * L * NLPF_SCREEN_PLANES * plane * elem_size(desc->precision);
    hipError_t e;
    if (desc->precision == NLPF_FP64) {
        ScreenArgs<double> a{};
        a.rec = rec;
        a.anchors = nullptr;
        a.out = (double*)out;
        a.nodes = h->d_nodes;
        a.root = root;
        a.chain = make_chain(h->h_nodes.data(), root, 0u);
        a.need_pos = need_pos;
        a.material = mat;
        a.flags = flg;
        a.roots = h->d_roots;
        a.n_materials = n_mat;
        a.lds_nodes = lds_nodes;
        a.W = desc->width;
        a.H = desc->height;
        a.S = S;
        a.L = L;
        a.d = (uint32_t)d;
        a.out_levels = out_levels;
        a.out_slot = slot;
        a.child_slot = cslot;
        a.has_child = d + 1 < (int)L;
        a.plane = plane;
        a.kp = kp;
        a.kn = kn;
        a.kl = kl;
        a.rscale = 1.0;
        e = screen_level_fp64(a, desc->radius, 0, h->stream);
    } else {
        ScreenArgs<float> a{};
        a.rec = rec;
        a.anchors = desc->precision == NLPF_FP16 ? anchors + (size_t)s0 * L * abx * aby * 3 : nullptr;
        a.out = (float*)out;
        a.nodes = h->d_nodes;
        a.root = root;
        a.chain = make_chain(h->h_nodes.data(), root, 0u);
        a.need_pos = need_pos;
        a.material = mat;
        a.flags = flg;
        a.roots = h->d_roots;
        a.n_materials = n_mat;
        a.lds_nodes = lds_nodes;
        a.W = desc->width;
        a.H = desc->height;
        a.S = S;
        a.L = L;
        a.d = (uint32_t)d;
        a.out_levels = out_levels;
        a.out_slot = slot;
        a.child_slot = cslot;
        a.has_child = d + 1 < (int)L;
        a.plane = plane;
        a.abx = abx;
        a.aby = aby;
        a.kp = (float)kp;
        a.kn = (float)kn;
        a.kl = (float)kl;
        a.rscale = desc->precision == NLPF_FP16 ? NLPF_FP16_RADIANCE_SCALE : 1.0f;
        e = desc->precision == NLPF_FP16 ? screen_level_fp16(a, desc->radius, screen_variant_of(h->cfg), h->stream)
                                         : screen_level_fp32(a, desc->radius, screen_variant_of(h->cfg), h->stream);
    }
    if (e != hipSuccess) NLPF_FAIL(h, NLPF_E_HIP, std::string("screen level launch: ") + hipGetErrorString(e));
    return NLPF_OK;
}

// algorithmic figures of a screen pass (timings left to the caller)
static nlpf_stats screen_pass_stats(const nlpf_screen_desc* desc) {
    nlpf_stats st{};
    const uint64_t plane = (uint64_t)desc->width * desc->height;
    const uint32_t L = desc->levels, S = desc->spp;
    const double out_b = 3.0 * (desc->precision == NLPF_FP64 ? 8 : 4);
    st.vertices = plane * S * L;
    const double rec_b = (double)elem_size(desc->precision) * NLPF_SCREEN_PLANES;
    // every record read once, every level output written once and (below the top) read once
    st.bytes_algorithmic = (uint64_t)((double)st.vertices * (rec_b + out_b) + (double)plane * S * (L - 1) * out_b);
    // per neighbour pair: |dp|^2 8 (3 sub, mul, 2 FMA), normal term 8 (4 FMA), luminance 3
    // (sub, FMA), exp 1, accumulate 7 (weight-sum add, 3 FMA)
    const double K = 2.0 * desc->radius + 1.0;
    st.flops_algorithmic = (double)st.vertices * K * K * 27.0;
    return st;
}

// screen pass over desc->spp slices; with s_total > desc->spp the slices are spp slices
// [slice0, slice0 + desc->spp) of an s_total-slice frame and film carries the frame's running
// sums between calls (accumulate: continue them; finalize: divide by s_total)
static int screen_run_impl(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records, const float* anchors,
                           void* film, void* level_out, uint32_t s_total, int accumulate, int finalize,
                           const char* name, const uint16_t* material = nullptr, const uint8_t* flags = nullptr) {
    int rc = validate_screen_desc(h, desc);
    if (rc) return rc;
    if (!records || !film) NLPF_FAIL(h, NLPF_E_INVALID_ARG, std::string(name) + ": NULL records or film");
    if (desc->precision == NLPF_FP16 && !anchors)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, std::string(name) + ": FP16 records need position anchors");
    const bool f64 = desc->precision == NLPF_FP64;
    const size_t asz = f64 ? 8 : 4;
    const uint64_t plane = (uint64_t)desc->width * desc->height;
    const uint32_t L = desc->levels, S = desc->spp;
    uint32_t out_levels = L;
    void* outbuf = level_out;
    if (!outbuf) {
        out_levels = L > 1 ? 2 : 1;
        rc = ensure_scratch(h, (size_t)S * out_levels * 3 * plane * asz);
        if (rc) return rc;
        outbuf = h->scratch;
    }
    rc = record_event(h, 0);
    if (rc) return rc;
    for (int d = (int)L - 1; d >= 0; --d) {
        rc = screen_level_launch(h, desc, records, anchors, outbuf, out_levels, 0, S, d, material, flags);
        if (rc) return rc;
        rc = record_event(h, 1 + (L - 1 - d));
        if (rc) return rc;
    }
    hipError_t e = f64 ? screen_film_fp64((const double*)outbuf, S, out_levels, 0, plane, (double*)film, h->stream,
                                          s_total, accumulate, finalize)
                       : screen_film_fp32((const float*)outbuf, S, out_levels, 0, plane, (float*)film, h->stream,
                                          s_total, accumulate, finalize);
    if (e != hipSuccess) NLPF_FAIL(h, NLPF_E_HIP, std::string("screen film launch: ") + hipGetErrorString(e));
    rc = record_event(h, 1 + L);
    if (rc) return rc;
    NLPF_HIP(h, hipStreamSynchronize(h->stream));

    nlpf_stats st = screen_pass_stats(desc);
    st.total_ms = event_ms(h, 0, 1 + L);
    for (uint32_t i = 0; i < L && i < 32; ++i) {
        const uint32_t d = L - 1 - i;
        st.level_ms[d] = event_ms(h, i, i + 1);
    }
    st.film_ms = event_ms(h, L, L + 1);
    h->stats = st;
    return NLPF_OK;
}

int nlpf_screen_run(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records, const float* anchors,
                    void* film, void* level_out) {
    if (!h) return NLPF_E_INVALID_ARG;
    return screen_run_impl(h, desc, records, anchors, film, level_out, desc ? desc->spp : 0, 0, 1,
                           "nlpf_screen_run");
}

int nlpf_screen_run_materials(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records,
                              const float* anchors, const uint16_t* material, const uint8_t* flags, void* film,
                              void* level_out) {
    if (!h) return NLPF_E_INVALID_ARG;
    if (!material) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_screen_run_materials: NULL material plane");
    return screen_run_impl(h, desc, records, anchors, film, level_out, desc ? desc->spp : 0, 0, 1,
                           "nlpf_screen_run_materials", material, flags);
}

int nlpf_screen_run_chunk(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records, const float* anchors,
                          void* film, void* level_out, uint32_t slice0, uint32_t total_spp) {
    if (!h) return NLPF_E_INVALID_ARG;
    if (!desc) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "screen desc is NULL");
    if (desc->spp == 0 || (uint64_t)slice0 + desc->spp > total_spp)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_screen_run_chunk: slices [slice0, slice0 + spp) must lie in total_spp");
    return screen_run_impl(h, desc, records, anchors, film, level_out, total_spp, slice0 > 0,
                           slice0 + desc->spp == total_spp, "nlpf_screen_run_chunk");
}

int nlpf_screen_run_noalloc(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records,
                            const float* anchors, void* film) {
    if (!h) return NLPF_E_INVALID_ARG;
    int rc = validate_screen_desc(h, desc);
    if (rc) return rc;
    if (!records || !film) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_screen_run_noalloc: NULL records or film");
    if (desc->precision == NLPF_FP16 && !anchors)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_screen_run_noalloc: FP16 records need position anchors");
    if (!h->noalloc_scratch)
        NLPF_FAIL(h, NLPF_E_HIP, "nlpf_screen_run_noalloc: the static scratch of this device is not resolved");
    const bool f64 = desc->precision == NLPF_FP64;
    const uint64_t plane = (uint64_t)desc->width * desc->height;
    const uint32_t L = desc->levels, S = desc->spp;
    const uint32_t out_levels = L > 1 ? 2 : 1;
    // two level images per slice (one for a single level); as many slices per group as fit
    const uint64_t per_slice = (uint64_t)out_levels * 3 * plane * (f64 ? 8 : 4);
    const uint64_t group = NLPF_NOALLOC_SCRATCH_BYTES / per_slice;
    if (group == 0)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG,
                  "nlpf_screen_run_noalloc: one slice needs " + std::to_string(per_slice) +
                      " bytes of scratch, more than the static " + std::to_string(NLPF_NOALLOC_SCRATCH_BYTES) +
                      " (use nlpf_screen_run)");
    for (uint32_t s0 = 0; s0 < S; s0 += (uint32_t)group) {
        const uint32_t cnt = (uint32_t)std::min<uint64_t>(group, S - s0);
        for (int d = (int)L - 1; d >= 0; --d) {
            rc = screen_level_launch(h, desc, records, anchors, h->noalloc_scratch, out_levels, s0, cnt, d);
            if (rc) return rc;
        }
        const int acc = s0 > 0, fin = s0 + cnt == S;
        hipError_t e =
            f64 ? screen_film_fp64((const double*)h->noalloc_scratch, cnt, out_levels, 0, plane, (double*)film,
                                   h->stream, S, acc, fin)
                : screen_film_fp32((const float*)h->noalloc_scratch, cnt, out_levels, 0, plane, (float*)film,
                                   h->stream, S, acc, fin);
        if (e != hipSuccess) NLPF_FAIL(h, NLPF_E_HIP, std::string("screen film launch: ") + hipGetErrorString(e));
    }
    h->stats = screen_pass_stats(desc);
    return NLPF_OK;
}

int nlpf_get_stats(const nlpf_handle* h, nlpf_stats* out) {
    if (!h || !out) return NLPF_E_INVALID_ARG;
    *out = h->stats;
    return NLPF_OK;
}

int nlpf_generate_screen_records(nlpf_handle* h, const nlpf_screen_desc* desc, const nlpf_gen_desc* gen,
                                 void* records, float* anchors) {
    if (!h) return NLPF_E_INVALID_ARG;
    int rc = validate_screen_desc(h, desc);
    if (rc) return rc;
    if (!gen || !records) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_generate_screen_records: NULL argument");
    if (desc->precision == NLPF_FP64) NLPF_FAIL(h, NLPF_E_UNSUPPORTED, "generator writes FP32 or FP16 records");
    const uint64_t plane = (uint64_t)desc->width * desc->height;
    if (desc->precision == NLPF_FP32) {
        NLPF_HIP(h, launch_gen_screen((float*)records, desc->width, desc->height, desc->levels, 0, desc->spp, *gen,
                                      h->stream));
    } else {
        if (!anchors) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "FP16 records need an anchor buffer");
        // one slice at a time through an FP32 staging buffer
        const size_t slice_bytes = (size_t)desc->levels * NLPF_SCREEN_PLANES * plane * sizeof(float);
        rc = ensure_scratch(h, slice_bytes);
        if (rc) return rc;
        const uint64_t abn = (uint64_t)((desc->width + 7) / 8) * ((desc->height + 7) / 8);
        for (uint32_t s = 0; s < desc->spp; ++s) {
            NLPF_HIP(h, launch_gen_screen((float*)h->scratch, desc->width, desc->height, desc->levels, s, 1, *gen,
                                          h->stream));
            __half* dst = (__half*)records + (size_t)s * desc->levels * NLPF_SCREEN_PLANES * plane;
            NLPF_HIP(h, launch_pack_fp16((const float*)h->scratch, dst, anchors + (size_t)s * desc->levels * abn * 3,
                                         desc->width, desc->height, desc->levels, h->stream));
        }
    }
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    return NLPF_OK;
}

int nlpf_convert_records_fp16(nlpf_handle* h, const nlpf_screen_desc* desc, const float* in, void* out,
                              float* anchors) {
    if (!h) return NLPF_E_INVALID_ARG;
    int rc = validate_screen_desc(h, desc);
    if (rc) return rc;
    if (!in || !out || !anchors) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_convert_records_fp16: NULL argument");
    NLPF_HIP(h, launch_pack_fp16(in, (__half*)out, anchors, desc->width, desc->height, desc->spp * desc->levels,
                                 h->stream));
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    return NLPF_OK;
}

}  // extern "C"
