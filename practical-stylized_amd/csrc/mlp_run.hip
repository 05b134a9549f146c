// Screen-mode NL-PF with the learned inner-expectation cache at depth k.
//
// Per frame (one call):
//   1. audit slice  levels D..k of one spp slice filtered exactly (plain NL-PF); the unstylized
//                   depth-k values F_k are the training targets; levels k-1..0 from the exact
//                   depth-k output give the slice's exact images (kept as its final result);
//   2. train        `train_records` audit vertices, Adam steps of `batch` on BF16 MFMA;
//   3. infer        the network estimates F_k for every slice (levels > k are not read),
//                   stylized at depth k like a filtered value;
//   4. audit        levels k-1..0 of the audit slice rerun on its estimate (in a separate
//                   image) give P; per 64x8 tile, the relative L2 of P against the exact
//                   level-0 image E is the error the cache puts into the image;
//   5. fallback     tiles above the threshold are filtered exactly at depths D..k in every
//                   other slice: a level-0 tile depends on depth-d tiles up to d tiles away
//                   (window radius <= 7 < tile height 8), so the depth-d mask is the level-0
//                   mask dilated d times;
//   6. levels k-1..0 plain NL-PF on the mixed depth-k output of the other slices, then the
//      film over all slices.
// A second stream of the handle takes the audit slice's exact levels k-1..0 (beside training
// and inference) and then its rerun (beside the other slices' inference); the handle's stream
// waits for both before the per-tile errors.
#include "common.hpp"
#include "screen_impl.hpp"
#include "scene.hpp"

namespace nlpf {
hipError_t screen_level_fp32(const ScreenArgs<float>& a, int radius, int variant, hipStream_t st);
hipError_t screen_level_fp16(const ScreenArgs<float>& a, int radius, int variant, hipStream_t st);
hipError_t screen_film_fp32(const float* out, uint32_t S, uint32_t out_levels, uint32_t slot0, uint64_t plane,
                            float* film, hipStream_t st, uint32_t s_total = 0, int accumulate = 0, int finalize = 1);
int mlp_infer_planar(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records, const float* anchors,
                     uint32_t depth, uint64_t sample0, uint64_t n, float* planar_out, uint32_t out_levels,
                     uint32_t out_slot, uint32_t out_slice0);
int mlp_train_enqueue(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records, const float* anchors,
                      uint32_t depth, const uint32_t* idx, const float* feats, const float* target, uint32_t batch,
                      float lr, int apply_update, float* grad);
const float* mlp_loss_device(nlpf_handle* h);

__global__ void gather_targets_kernel(const float* F, uint64_t plane, const uint32_t* idx, uint32_t n, float* t) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t px = idx[i];
    t[3 * i] = F[px];
    t[3 * i + 1] = F[plane + px];
    t[3 * i + 2] = F[2 * plane + px];
}

// per-channel mean of a planar [3][plane] image in a fixed order: workgroup b of channel c
// sums the b-th of 256 contiguous chunks (each lane a strided share, then a fixed LDS tree),
// then one lane per channel adds the 256 chunk sums in order
__global__ void __launch_bounds__(256) plane_partial_kernel(const float* F, uint64_t plane, double* part) {
    __shared__ double red[256];
    const uint32_t c = blockIdx.y, b = blockIdx.x, t = threadIdx.x;
    const uint64_t chunk = (plane + 255) / 256, lo = b * chunk, hi = lo + chunk < plane ? lo + chunk : plane;
    double acc = 0.0;
    for (uint64_t i = lo + t; i < hi; i += 256) acc += (double)F[c * plane + i];
    red[t] = acc;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) part[c * 256 + b] = red[0];
}
__global__ void plane_mean_kernel(const double* part, uint64_t plane, float* mean3) {
    const uint32_t c = threadIdx.x;
    if (c >= 3) return;
    double acc = 0.0;
    for (int b = 0; b < 256; ++b) acc += part[c * 256 + b];
    mean3[c] = (float)(acc / (double)plane);
}
// slot `slot` of slices [s0, s0 + n) of planar level data ([S][out_levels][3][plane]) set to the
// constant mean3 (zero when mean3 is NULL)
__global__ void fill_slot_kernel(float* io, uint32_t n, uint64_t plane, uint32_t out_levels, uint32_t slot,
                                 const float* mean3) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= plane * n) return;
    const uint64_t s = i / plane, px = i - s * plane;
    float* o = io + (s * out_levels + slot) * 3 * plane + px;
    o[0] = mean3 ? mean3[0] : 0.0f;
    o[plane] = mean3 ? mean3[1] : 0.0f;
    o[2 * plane] = mean3 ? mean3[2] : 0.0f;
}

// out = S_d(in) on planar [S][3][plane] level data (flags 0, material 0); the chain comes by
// value, its ramp stops through LDS (as in the level kernel)
template <typename S>
__global__ void __launch_bounds__(256) style_planar_kernel(float* io, uint32_t Sn, uint64_t plane,
                                                           uint32_t out_levels, uint32_t slot, StyleChain chain,
                                                           uint32_t d, const S* rec, const float* anchors, uint32_t W,
                                                           uint32_t H, uint32_t L, int need_pos) {
    __shared__ float tabs[NLPF_CHAIN_MAX * 32];
    if (threadIdx.x < NLPF_CHAIN_MAX * 32 && (int)(threadIdx.x >> 5) < chain.n)
        tabs[threadIdx.x] = chain.node[threadIdx.x >> 5].p[threadIdx.x & 31];
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= plane * Sn) return;
    const uint64_t s = i / plane, px = i - s * plane;
    float* o = io + (s * out_levels + slot) * 3 * plane + px;
    v3<float> p{0.0f, 0.0f, 0.0f};
    if (need_pos) {
        const S* r = rec + (s * L + d) * NLPF_SCREEN_PLANES * plane + px;
        p = v3<float>{(float)r[0], (float)r[plane], (float)r[2 * plane]};
        if (anchors) {
            const uint32_t abx = (W + 7) / 8, aby = (H + 7) / 8, x = (uint32_t)(px % W), y = (uint32_t)(px / W);
            const float* q = anchors + ((s * L + d) * abx * aby + (uint64_t)(y >> 3) * abx + (x >> 3)) * 3;
            p.x += q[0];
            p.y += q[1];
            p.z += q[2];
        }
    }
    v3<float> v{o[0], o[plane], o[2 * plane]};
    v = apply_chain<float>(chain, d, v, p, tabs);
    o[0] = v.x;
    o[plane] = v.y;
    o[2 * plane] = v.z;
}

// squared error and squared norm of the prediction per 64x8 tile of the audit slice
// (one workgroup per tile, fixed-order reduction)
__global__ void tile_error_kernel(const float* pred, const float* audit, uint32_t W, uint32_t H, uint32_t tiles_x,
                                  double* num_den) {
    __shared__ double sn[256], sd[256];
    const uint32_t tile = blockIdx.x;
    const uint32_t x0 = (tile % tiles_x) * 64, y0 = (tile / tiles_x) * 8;
    const uint64_t plane = (uint64_t)W * H;
    double num = 0.0, den = 0.0;
    for (uint32_t k = threadIdx.x; k < 512; k += blockDim.x) {
        const uint32_t x = x0 + (k & 63), y = y0 + (k >> 6);
        if (x < W && y < H) {
            const uint64_t o = (uint64_t)y * W + x;
            for (int c = 0; c < 3; ++c) {
                const double a = pred[c * plane + o], b = audit[c * plane + o];
                num += (a - b) * (a - b);
                den += b * b;
            }
        }
    }
    sn[threadIdx.x] = num;
    sd[threadIdx.x] = den;
    __syncthreads();
    for (uint32_t off = blockDim.x / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sn[threadIdx.x] += sn[threadIdx.x + off];
            sd[threadIdx.x] += sd[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        num_den[2 * tile] = sn[0];
        num_den[2 * tile + 1] = sd[0];
    }
}

static bool tree_has(const std::vector<nlpf_style_node>& nodes, int root, int type, int depth = 0) {
    if (root < 0 || depth > NLPF_STYLE_STACK) return false;
    const nlpf_style_node& n = nodes[root];
    if (n.type == type) return true;
    if (tree_has(nodes, n.inner, type, depth + 1)) return true;
    return n.type == NLPF_STY_ENABLE_IF && tree_has(nodes, n.else_, type, depth + 1);
}

}  // namespace nlpf

using namespace nlpf;

static void dilate_tiles(std::vector<uint8_t>& m, uint32_t tx, uint32_t ty) {
    std::vector<uint8_t> o(m.size(), 0);
    for (uint32_t y = 0; y < ty; ++y)
        for (uint32_t x = 0; x < tx; ++x) {
            uint8_t v = 0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int xx = (int)x + dx, yy = (int)y + dy;
                    if (xx >= 0 && yy >= 0 && xx < (int)tx && yy < (int)ty) v |= m[(size_t)yy * tx + xx];
                }
            o[(size_t)y * tx + x] = v;
        }
    m.swap(o);
}

extern "C" int nlpf_screen_run_mlp(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records,
                                   const float* anchors, const nlpf_mlp_run_desc* run, float* film, float* level_out,
                                   nlpf_mlp_run_stats* stats) {
    if (!h) return NLPF_E_INVALID_ARG;
    if (!desc || !records || !run || !film) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_screen_run_mlp: NULL argument");
    if (desc->precision != NLPF_FP32 && desc->precision != NLPF_FP16)
        NLPF_FAIL(h, NLPF_E_UNSUPPORTED, "the learned cache runs on FP32 or FP16 records");
    if (desc->precision == NLPF_FP16 && !anchors) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "FP16 records need anchors");
    const uint32_t W = desc->width, H = desc->height, S = desc->spp, L = desc->levels, k = run->depth;
    if (W == 0 || H == 0 || S == 0 || L == 0 || L > 31 || desc->radius < 1 || desc->radius > 7)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_screen_run_mlp: bad screen desc");
    if (k < 1 || k >= L) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "cache depth must be in [1, levels - 1]");
    if (run->audit_slice >= S) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "audit slice out of range");
    if (run->batch == 0 || run->batch % 128 != 0 || run->train_records % run->batch != 0)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "batch must be a multiple of 128 dividing train_records");
    const uint64_t plane = (uint64_t)W * H;
    if (run->train_records > plane) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "more training records than audit vertices");
    if (plane * S >= (1ull << 32)) NLPF_FAIL(h, NLPF_E_INVALID_ARG, "too many vertices");
    const uint32_t est = run->estimator;
    if (est != NLPF_CACHE_NETWORK && est != NLPF_CACHE_ZERO && est != NLPF_CACHE_MEAN)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "unknown cache estimator");
    const int fp16 = desc->precision == NLPF_FP16;
    const uint32_t tiles_x = (W + 63) / 64, tiles_y = (H + 7) / 8, ntiles = tiles_x * tiles_y;
    const uint32_t sj = run->audit_slice;
    const size_t esz = fp16 ? 2 : 4;
    const uint32_t abx = (W + 7) / 8, aby = (H + 7) / 8;

    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t lv_bytes = level_out ? 0 : (size_t)S * L * 3 * plane * sizeof(float);
    const size_t f_bytes = 3 * plane * sizeof(float);
    const size_t tgt_bytes = (size_t)run->train_records * 3 * sizeof(float);
    const size_t idx_bytes = (size_t)run->train_records * sizeof(uint32_t);
    const size_t mask_bytes = (size_t)L * ntiles;
    const size_t nd_bytes = (size_t)ntiles * 2 * sizeof(double);
    const size_t p_bytes = (size_t)(k + 1) * f_bytes;
    const size_t part_bytes = 3 * 256 * sizeof(double) + 4 * sizeof(float);
    int rc = ensure_scratch(h, al(lv_bytes) + al(f_bytes) + al(p_bytes) + al(tgt_bytes) + 2 * al(idx_bytes) +
                                   al(mask_bytes) + 2 * al(nd_bytes) + al(part_bytes));
    if (rc) return rc;
    char* base = (char*)h->scratch;
    auto take = [&](size_t b) {
        char* p = base;
        base += al(b);
        return p;
    };
    float* lv = level_out ? level_out : (float*)take(lv_bytes);
    float* Fj = (float*)take(f_bytes);  // audit slice: exact unstylized F_k (training targets)
    float* Pj = (float*)take(p_bytes);  // audit slice on its estimate: levels 0..k [k+1][3][W*H]
    float* tgt = (float*)take(tgt_bytes);
    uint32_t* lidx = (uint32_t*)take(idx_bytes);  // pixel ids within the audit slice
    uint32_t* gidx = (uint32_t*)take(idx_bytes);  // sample ids s*W*H + pixel
    uint8_t* masks = (uint8_t*)take(mask_bytes);
    double* nd_img = (double*)take(nd_bytes);
    double* nd_fit = (double*)take(nd_bytes);
    double* part = (double*)take(part_bytes);
    float* mean3 = (float*)(part + 3 * 256);
    hipStream_t st = h->stream;

    const int root = h->h_roots.empty() ? -1 : h->h_roots[0];
    const int need_pos = root >= 0 && tree_has(h->h_nodes, root, NLPF_STY_HALFTONE);
    const double log2e = 1.4426950408889634;
    ScreenArgs<float> a0{};
    a0.rec = records;
    a0.anchors = fp16 ? anchors : nullptr;
    a0.out = lv;
    a0.nodes = h->d_nodes;
    a0.root = root;
    a0.chain = make_chain(h->h_nodes.data(), root, 0u);
    a0.need_pos = need_pos;
    a0.W = W;
    a0.H = H;
    a0.S = S;
    a0.L = L;
    a0.out_levels = L;
    a0.plane = plane;
    a0.abx = abx;
    a0.aby = aby;
    a0.kp = (float)sqrt(log2e / (2.0 * (double)h->cfg.sigma_pos * h->cfg.sigma_pos));
    a0.kn = (float)sqrt(log2e / (2.0 * (double)h->cfg.sigma_normal * h->cfg.sigma_normal));
    a0.kl = (float)sqrt(log2e / (2.0 * (double)h->cfg.sigma_lum * h->cfg.sigma_lum));
    a0.rscale = fp16 ? NLPF_FP16_RADIANCE_SCALE : 1.0f;
    a0.tile_mask = nullptr;
    auto level_on = [&](ScreenArgs<float> a, uint32_t d, hipStream_t s) -> hipError_t {
        a.d = d;
        a.out_slot = d;
        a.child_slot = d + 1;
        a.has_child = d + 1 < L;
        // a tile mask covers 64x8 tiles: masked launches take the 8-row tile variant 0
        const int v = a.tile_mask ? 0 : screen_variant_of(h->cfg);
        return fp16 ? screen_level_fp16(a, desc->radius, v, s) : screen_level_fp32(a, desc->radius, v, s);
    };
    auto level = [&](ScreenArgs<float> a, uint32_t d) -> hipError_t { return level_on(a, d, st); };
    auto slice_rec = [&](uint32_t s) {
        return (const void*)((const char*)records + (size_t)s * L * NLPF_SCREEN_PLANES * plane * esz);
    };
    auto slice_anc = [&](uint32_t s) { return fp16 ? anchors + (size_t)s * L * abx * aby * 3 : (const float*)nullptr; };
    auto slot = [&](uint32_t s, uint32_t d) { return lv + ((size_t)s * L + d) * 3 * plane; };
    ScreenArgs<float> aj = a0;  // the audit slice alone
    aj.S = 1;
    aj.rec = slice_rec(sj);
    aj.anchors = slice_anc(sj);
    aj.out = slot(sj, 0);
    // stylize slot k of `n` slices of planar level data io ([n][out_levels][3][W*H]) whose
    // records start at slice s0
    const StyleChain chain_k = make_chain(h->h_nodes.data(), root, 0u);
    if (root >= 0 && !chain_k.valid)
        NLPF_FAIL(h, NLPF_E_UNSUPPORTED, "the learned cache stylizes chains of at most 4 nodes");
    auto style_k = [&](float* io, uint32_t out_levels, uint32_t s0, uint32_t n) -> int {
        if (root < 0 || n == 0) return NLPF_OK;
        const uint64_t cnt = plane * n;
        if (fp16)
            hipLaunchKernelGGL(style_planar_kernel<__half>, dim3((uint32_t)((cnt + 255) / 256)), dim3(256), 0, st, io,
                               n, plane, out_levels, k, chain_k, k, (const __half*)slice_rec(s0), slice_anc(s0), W,
                               H, L, need_pos);
        else
            hipLaunchKernelGGL(style_planar_kernel<float>, dim3((uint32_t)((cnt + 255) / 256)), dim3(256), 0, st, io,
                               n, plane, out_levels, k, chain_k, k, (const float*)slice_rec(s0), nullptr, W, H, L,
                               need_pos);
        return check_launch(h, "style depth k");
    };
    // one level over the slices other than the audit slice: one launch that passes over it
    auto level_others = [&](ScreenArgs<float> a, uint32_t d) -> hipError_t {
        if (S < 2) return hipSuccess;
        a.S = S - 1;
        a.hole1 = sj + 1;
        return level(a, d);
    };

    nlpf_mlp_run_stats rs{};
    rc = record_event(h, 0);
    if (rc) return rc;
    // 1. audit slice, exact: levels D..k+1, unstylized F_k (targets), stylized depth k, levels k-1..0
    for (int d = (int)L - 1; d > (int)k; --d) NLPF_HIP(h, level(aj, (uint32_t)d));
    {
        ScreenArgs<float> af = aj;
        af.root = -1;
        af.chain = StyleChain{};
        NLPF_HIP(h, level(af, k));
    }
    NLPF_HIP(h, hipMemcpyAsync(Fj, slot(sj, k), f_bytes, hipMemcpyDeviceToDevice, st));
    if ((rc = style_k(slot(sj, 0), L, sj, 1))) return rc;
    // levels k-1..0 of the audit slice (its final images, which the tile audit compares against)
    // run on the second stream, overlapping the training and inference below; the handle's
    // stream waits for them before the audit
    if (!h->aux) {
        NLPF_HIP(h, hipStreamCreateWithFlags(&h->aux, hipStreamNonBlocking));
        NLPF_HIP(h, hipEventCreateWithFlags(&h->fork, hipEventDisableTiming));
        NLPF_HIP(h, hipEventCreateWithFlags(&h->join, hipEventDisableTiming));
    }
    NLPF_HIP(h, hipEventRecord(h->fork, st));
    NLPF_HIP(h, hipStreamWaitEvent(h->aux, h->fork, 0));
    for (int d = (int)k - 1; d >= 0; --d) NLPF_HIP(h, level_on(aj, (uint32_t)d, h->aux));
    NLPF_HIP(h, hipEventRecord(h->join, h->aux));
    rc = record_event(h, 1);
    if (rc) return rc;
    // 2. online training on `train_records` audit vertices (a PCG stream of `seed` picks them);
    //    the control estimators train nothing. The picks and the loss live to the end of the call:
    //    their copies complete before the synchronisation of the tile audit below.
    std::vector<uint32_t> pick_l, pick_g;
    float loss = 0.0f;
    if (est == NLPF_CACHE_NETWORK) {
        // a partial Fisher-Yates shuffle of the pixels on the handle's identity permutation, undone
        // once the picks are read (the host work is O(train_records), not O(pixels))
        std::vector<uint32_t>& perm = h->pick_perm;
        if (perm.size() != plane) {
            perm.resize(plane);
            for (uint64_t i = 0; i < plane; ++i) perm[i] = (uint32_t)i;
        }
        Pcg32 g = pcg_seq_offset(0x7EA1ull, splitmix64(run->seed));
        std::vector<uint32_t> js(run->train_records);
        for (uint32_t i = 0; i < run->train_records; ++i) {
            const uint32_t j = i + (uint32_t)(((uint64_t)pcg_next(g) * (plane - i)) >> 32);
            js[i] = j;
            std::swap(perm[i], perm[j]);
        }
        pick_l.assign(perm.begin(), perm.begin() + run->train_records);
        pick_g.resize(run->train_records);
        for (uint32_t i = 0; i < run->train_records; ++i) pick_g[i] = sj * (uint32_t)plane + pick_l[i];
        for (uint32_t i = run->train_records; i-- > 0;) std::swap(perm[i], perm[js[i]]);  // back to the identity
        NLPF_HIP(h, hipMemcpyAsync(lidx, pick_l.data(), idx_bytes, hipMemcpyHostToDevice, st));
        NLPF_HIP(h, hipMemcpyAsync(gidx, pick_g.data(), idx_bytes, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(gather_targets_kernel, dim3((run->train_records + 255) / 256), dim3(256), 0, st, Fj, plane,
                           lidx, run->train_records, tgt);
        if ((rc = check_launch(h, "gather targets"))) return rc;
        for (uint32_t b = 0; b < run->train_records / run->batch; ++b) {
            rc = mlp_train_enqueue(h, desc, records, anchors, k, gidx + (size_t)b * run->batch, nullptr,
                                   tgt + (size_t)b * run->batch * 3, run->batch, run->lr, 1, nullptr);
            if (rc) return rc;
        }
        NLPF_HIP(h, hipMemcpyAsync(&loss, mlp_loss_device(h), sizeof(float), hipMemcpyDeviceToHost, st));
    }
    rc = record_event(h, 2);
    if (rc) return rc;
    // 3. the estimate at depth k: first the audit slice's (slot k of its separate image), its fit
    //    against the targets and its stylization; its audit rerun then runs on the second stream
    //    (4.) while the other slices' estimates are formed and stylized here
    float* Pk = Pj + (size_t)k * 3 * plane;
    const float* c3 = nullptr;  // the controls' constant (zero when NULL)
    if (est == NLPF_CACHE_NETWORK) {
        rc = mlp_infer_planar(h, desc, records, anchors, k, (uint64_t)sj * plane, plane, Pj, k + 1, k, sj);
        if (rc) return rc;
    } else {
        // controls: the audit slice's mean unstylized depth-k value everywhere, or zero (the
        // paths cut at depth k)
        if (est == NLPF_CACHE_MEAN) {
            hipLaunchKernelGGL(plane_partial_kernel, dim3(256, 3), dim3(256), 0, st, Fj, plane, part);
            hipLaunchKernelGGL(plane_mean_kernel, dim3(1), dim3(64), 0, st, part, plane, mean3);
            c3 = mean3;
        }
        hipLaunchKernelGGL(fill_slot_kernel, dim3((uint32_t)((plane + 255) / 256)), dim3(256), 0, st, Pj, 1u, plane,
                           k + 1, k, c3);
        if ((rc = check_launch(h, "cache control estimate"))) return rc;
    }
    hipLaunchKernelGGL(tile_error_kernel, dim3(ntiles), dim3(256), 0, st, Pk, Fj, W, H, tiles_x, nd_fit);
    if ((rc = check_launch(h, "cache fit"))) return rc;
    // a filtered estimate is stylized at depth k; a path cut at depth k carries nothing up
    if (est != NLPF_CACHE_ZERO && (rc = style_k(Pj, k + 1, sj, 1))) return rc;
    // 4. audit, on the second stream after the exact levels: levels k-1..0 of the audit slice on
    //    its estimate
    NLPF_HIP(h, hipEventRecord(h->fork, st));
    NLPF_HIP(h, hipStreamWaitEvent(h->aux, h->fork, 0));
    {
        ScreenArgs<float> ap = aj;
        ap.out = Pj;
        ap.out_levels = k + 1;
        for (int d = (int)k - 1; d >= 0; --d) {
            ScreenArgs<float> b = ap;
            b.d = (uint32_t)d;
            b.out_slot = (uint32_t)d;
            b.child_slot = (uint32_t)d + 1;
            b.has_child = 1;  // slot k holds the estimate
            NLPF_HIP(h, fp16 ? screen_level_fp16(b, desc->radius, screen_variant_of(h->cfg), h->aux)
                             : screen_level_fp32(b, desc->radius, screen_variant_of(h->cfg), h->aux));
        }
    }
    NLPF_HIP(h, hipEventRecord(h->join, h->aux));
    // the other slices' estimates (two ranges around the audit slice) and their stylization
    const uint32_t r0[2] = {0, sj + 1}, rn[2] = {sj, S - 1 - sj};
    for (int r = 0; r < 2; ++r) {
        if (!rn[r]) continue;
        if (est == NLPF_CACHE_NETWORK) {
            rc = mlp_infer_planar(h, desc, records, anchors, k, (uint64_t)r0[r] * plane, (uint64_t)rn[r] * plane, lv,
                                  L, k, 0);
            if (rc) return rc;
        } else {
            hipLaunchKernelGGL(fill_slot_kernel, dim3((uint32_t)((plane * rn[r] + 255) / 256)), dim3(256), 0, st,
                               slot(r0[r], 0), rn[r], plane, L, k, c3);
            if ((rc = check_launch(h, "cache control estimate"))) return rc;
        }
        if (est != NLPF_CACHE_ZERO && (rc = style_k(slot(r0[r], 0), L, r0[r], rn[r]))) return rc;
    }
    rc = record_event(h, 3);
    if (rc) return rc;
    // the per-tile error of the rerun against the audit slice's exact level-0 image (both from
    // the second stream)
    NLPF_HIP(h, hipStreamWaitEvent(st, h->join, 0));  // the rerun and the exact level 0
    hipLaunchKernelGGL(tile_error_kernel, dim3(ntiles), dim3(256), 0, st, Pj, slot(sj, 0), W, H, tiles_x, nd_img);
    if ((rc = check_launch(h, "tile audit"))) return rc;
    std::vector<double> hi(2 * (size_t)ntiles), hf(2 * (size_t)ntiles);
    NLPF_HIP(h, hipMemcpyAsync(hi.data(), nd_img, nd_bytes, hipMemcpyDeviceToHost, st));
    NLPF_HIP(h, hipMemcpyAsync(hf.data(), nd_fit, nd_bytes, hipMemcpyDeviceToHost, st));
    NLPF_HIP(h, hipStreamSynchronize(st));
    rs.train_loss = loss;
    std::vector<uint8_t> m(ntiles, 0);
    double in_ = 0.0, id_ = 0.0, fn_ = 0.0, fd_ = 0.0;
    uint64_t nfall = 0;
    for (uint32_t t = 0; t < ntiles; ++t) {
        in_ += hi[2 * t];
        id_ += hi[2 * t + 1];
        fn_ += hf[2 * t];
        fd_ += hf[2 * t + 1];
        const double e = hi[2 * t + 1] > 0.0 ? std::sqrt(hi[2 * t] / hi[2 * t + 1]) : (hi[2 * t] > 0.0 ? 1e30 : 0.0);
        if (e > (double)run->tile_threshold) {
            m[t] = 1;
            ++nfall;
        }
    }
    rs.audit_rel_l2 = id_ > 0.0 ? (float)std::sqrt(in_ / id_) : 0.0f;
    rs.cache_rel_l2 = fd_ > 0.0 ? (float)std::sqrt(fn_ / fd_) : 0.0f;
    rs.tiles = ntiles;
    rs.fallback_tiles = nfall;
    // 5. exact filtering of the fallback tiles in the other slices (the audit slice is exact)
    if (nfall && S > 1) {
        std::vector<uint8_t> hm((size_t)L * ntiles, 0);
        for (uint32_t d = 1; d < L; ++d) {
            dilate_tiles(m, tiles_x, tiles_y);
            if (d >= k) std::copy(m.begin(), m.end(), hm.begin() + (size_t)d * ntiles);
        }
        for (uint32_t d = k; d < L; ++d)
            for (uint32_t t = 0; t < ntiles; ++t) rs.fallback_tiles_filtered += hm[(size_t)d * ntiles + t];
        NLPF_HIP(h, hipMemcpyAsync(masks, hm.data(), mask_bytes, hipMemcpyHostToDevice, st));
        for (int d = (int)L - 1; d >= (int)k; --d) {
            ScreenArgs<float> am = a0;
            am.tile_mask = masks + (size_t)d * ntiles;
            NLPF_HIP(h, level_others(am, (uint32_t)d));
        }
        NLPF_HIP(h, hipStreamSynchronize(st));  // hm outlives its upload
    }
    rc = record_event(h, 4);
    if (rc) return rc;
    // 6. plain levels k-1..0 on the mixed depth-k output of the other slices, then the film
    for (int d = (int)k - 1; d >= 0; --d) NLPF_HIP(h, level_others(a0, (uint32_t)d));
    NLPF_HIP(h, screen_film_fp32(lv, S, L, 0, plane, film, st));
    rc = record_event(h, 5);
    if (rc) return rc;
    NLPF_HIP(h, hipStreamSynchronize(st));
    rs.audit_ms = event_ms(h, 0, 1);
    rs.train_ms = event_ms(h, 1, 2);
    rs.infer_ms = event_ms(h, 2, 3);
    rs.fallback_ms = event_ms(h, 3, 4);
    rs.filter_ms = event_ms(h, 4, 5);
    rs.total_ms = event_ms(h, 0, 5);
    if (stats) *stats = rs;
    h->stats.total_ms = rs.total_ms;
    return NLPF_OK;
}
