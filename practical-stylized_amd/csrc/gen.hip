// Synthetic screen-mode records, FP16 packing and the standalone stylization kernel.
#include "common.hpp"
#include "scene.hpp"
#include "world_bsdf.hpp"

namespace nlpf {

struct GenArgs {
    float* out;  // fp32 planes for slices [s_begin, s_begin + s_count)
    uint32_t W, H, L, s_begin, s_count;
    uint64_t plane;
    uint64_t seed;
    float mu, sigma, fprob, fscale, alo, ahi;
};

// One thread per (pixel, slice).  Geometry: one camera path per (pixel, slice); the subpixel
// jitter and every bounce's (u1,u2) are shared by all pixels of a (slice, depth), so each
// level image is a coherent picture of the scene (screen-space neighbours stay neighbours in
// world space on smooth surfaces).  Shading fields are i.i.d. per record from a PCG32 stream
// keyed by the record's global index.
__global__ void gen_screen_kernel(GenArgs g) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t npix = g.plane;
    if (i >= npix * g.s_count) return;
    const uint32_t sl = (uint32_t)(i / npix);
    const uint64_t px = i - (uint64_t)sl * npix;
    const uint32_t s = g.s_begin + sl;
    const uint32_t x = (uint32_t)(px % g.W), y = (uint32_t)(px / g.W);
    const uint64_t key = splitmix64(g.seed);

    Pcg32 sg = pcg_seq_offset(0x5EED000000000000ull + s, key);
    const float jx = pcg_uniform(sg), jy = pcg_uniform(sg);
    v3<float> o, d;
    camera_ray(g.W, g.H, (float)x + jx, (float)y + jy, o, d);
    bool escaped = false;
    v3<float> sky_p{0, 0, 0}, sky_n{0, 0, 1};
    for (uint32_t lv = 0; lv < g.L; ++lv) {
        const float bu1 = pcg_uniform(sg), bu2 = pcg_uniform(sg);
        v3<float> p, n;
        if (!escaped) {
            const Hit h = scene_intersect(o, d);
            p = h.p;
            n = h.n;
            if (dot(n, d) > 0.0f) n = n * -1.0f;
            if (h.mat == NLPF_MAT_SKY) {
                escaped = true;
                sky_p = p;
                sky_n = n;
            }
        } else {
            p = sky_p;
            n = sky_n;
        }
        const uint64_t rid = (((uint64_t)s * g.L + lv) * g.H + y) * g.W + x;
        Pcg32 rg = pcg_seq_offset(rid, key);
        const float a0 = g.alo + (g.ahi - g.alo) * pcg_uniform(rg);
        const float a1 = g.alo + (g.ahi - g.alo) * pcg_uniform(rg);
        const float a2 = g.alo + (g.ahi - g.alo) * pcg_uniform(rg);
        // Box-Muller pairs for the three log-normal channels
        const float r0 = m_sqrt(-2.0f * logf(pcg_uniform_open0(rg)));
        const float t0 = 6.2831853f * pcg_uniform(rg);
        const float r1 = m_sqrt(-2.0f * logf(pcg_uniform_open0(rg)));
        const float t1 = 6.2831853f * pcg_uniform(rg);
        const float z0 = r0 * cosf(t0), z1 = r0 * sinf(t0), z2 = r1 * cosf(t1);
        const float ff = pcg_uniform(rg) < g.fprob ? g.fscale : 1.0f;
        float rad[3] = {expf(g.mu + g.sigma * z0) * ff, expf(g.mu + g.sigma * z1) * ff,
                        expf(g.mu + g.sigma * z2) * ff};
        float thr[3] = {pcg_uniform(rg), pcg_uniform(rg), pcg_uniform(rg)};
        if (escaped) {
            rad[0] = rad[1] = rad[2] = 0.0f;
            thr[0] = thr[1] = thr[2] = 0.0f;
        }
        float* base = g.out + ((uint64_t)sl * g.L + lv) * NLPF_SCREEN_PLANES * npix + px;
        const float vals[NLPF_SCREEN_PLANES] = {p.x, p.y, p.z, n.x, n.y, n.z, a0, a1, a2,
                                                 rad[0], rad[1], rad[2], thr[0], thr[1], thr[2]};
#pragma unroll
        for (int c = 0; c < NLPF_SCREEN_PLANES; ++c) base[(uint64_t)c * npix] = vals[c];
        if (!escaped) {
            o = p + n * 1e-4f;
            d = cosine_sample(n, bu1, bu2);
        }
    }
}

// FP32 -> FP16 records.  One wavefront per 8x8 pixel block of a (slice, depth) image: the
// block's anchor is the mean position of its non-sky records, positions are stored as FP16
// offsets from it, radiance is stored divided by NLPF_FP16_RADIANCE_SCALE (a power of two)
// so fireflies stay finite.
__global__ void pack_fp16_kernel(const float* in, __half* out, float* anchors, uint32_t W, uint32_t H,
                                 uint32_t nslices, uint32_t abx, uint32_t aby, uint64_t plane) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t blk = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint64_t nblk = (uint64_t)abx * aby * nslices;
    if (blk >= nblk) return;
    const uint64_t sl = blk / ((uint64_t)abx * aby);
    const uint32_t b = (uint32_t)(blk - sl * abx * aby);
    const uint32_t bx = b % abx, by = b / abx;
    const uint32_t x = bx * 8 + (lane & 7), y = by * 8 + (lane >> 3);
    const bool inside = x < W && y < H;
    const float* src = in + sl * NLPF_SCREEN_PLANES * plane;
    __half* dst = out + sl * NLPF_SCREEN_PLANES * plane;
    const uint64_t o = (uint64_t)y * W + x;
    float p[3] = {0, 0, 0};
    if (inside)
        for (int c = 0; c < 3; ++c) p[c] = src[c * plane + o];
    const bool use = inside && fabsf(p[0]) < 50.0f && fabsf(p[1]) < 50.0f && fabsf(p[2]) < 50.0f;
    float acc[4] = {use ? p[0] : 0.0f, use ? p[1] : 0.0f, use ? p[2] : 0.0f, use ? 1.0f : 0.0f};
    for (int off = 32; off > 0; off >>= 1)
        for (int c = 0; c < 4; ++c) acc[c] += __shfl_xor(acc[c], off);
    float anc[3] = {0, 0, 0};
    if (acc[3] > 0.0f)
        for (int c = 0; c < 3; ++c) anc[c] = acc[c] / acc[3];
    if (lane == 0)
        for (int c = 0; c < 3; ++c) anchors[blk * 3 + c] = anc[c];
    if (!inside) return;
    for (int c = 0; c < 3; ++c) dst[c * plane + o] = __float2half(p[c] - anc[c]);
    for (int c = 3; c < NLPF_SCREEN_PLANES; ++c) {
        float v = src[c * plane + o];
        if (c >= 9 && c < 12) v *= 1.0f / NLPF_FP16_RADIANCE_SCALE;
        dst[c * plane + o] = __float2half(v);
    }
}

__global__ void stylize_kernel(const float* in, float* out, uint64_t n, const uint32_t* depth, uint32_t depth0,
                               const uint8_t* flags, const uint16_t* material, const float* pos,
                               const nlpf_style_node* nodes, const int32_t* roots, uint32_t n_materials) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t m = material ? material[i] : 0u;
    const int root = m < n_materials ? roots[m] : -1;
    v3<float> li{in[3 * i], in[3 * i + 1], in[3 * i + 2]};
    v3<float> p{0, 0, 0};
    if (pos) p = v3<float>{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    if (root >= 0) li = apply_style<float>(nodes, root, depth ? depth[i] : depth0, flags ? flags[i] : 0u, li, p);
    out[3 * i] = li.x;
    out[3 * i + 1] = li.y;
    out[3 * i + 2] = li.z;
}

hipError_t launch_gen_screen(float* out, uint32_t W, uint32_t H, uint32_t L, uint32_t s_begin, uint32_t s_count,
                             const nlpf_gen_desc& gd, hipStream_t st) {
    GenArgs g;
    g.out = out;
    g.W = W;
    g.H = H;
    g.L = L;
    g.s_begin = s_begin;
    g.s_count = s_count;
    g.plane = (uint64_t)W * H;
    g.seed = gd.seed;
    g.mu = gd.mu;
    g.sigma = gd.sigma;
    g.fprob = gd.firefly_prob;
    g.fscale = gd.firefly_scale;
    g.alo = gd.albedo_lo;
    g.ahi = gd.albedo_hi;
    const uint64_t n = g.plane * s_count;
    const uint32_t bs = 256;
    hipLaunchKernelGGL(gen_screen_kernel, dim3((uint32_t)((n + bs - 1) / bs)), dim3(bs), 0, st, g);
    return hipGetLastError();
}

hipError_t launch_pack_fp16(const float* in, __half* out, float* anchors, uint32_t W, uint32_t H, uint32_t nslices,
                            hipStream_t st) {
    const uint32_t abx = (W + 7) / 8, aby = (H + 7) / 8;
    const uint64_t nblk = (uint64_t)abx * aby * nslices;
    const uint32_t wpb = 4;
    hipLaunchKernelGGL(pack_fp16_kernel, dim3((uint32_t)((nblk + wpb - 1) / wpb)), dim3(64 * wpb), 0, st, in, out,
                       anchors, W, H, nslices, abx, aby, (uint64_t)W * H);
    return hipGetLastError();
}

hipError_t launch_stylize(const float* in, float* out, uint64_t n, const uint32_t* depth, uint32_t depth0,
                          const uint8_t* flags, const uint16_t* material, const float* pos,
                          const nlpf_style_node* nodes, const int32_t* roots, uint32_t n_materials, hipStream_t st) {
    const uint32_t bs = 256;
    hipLaunchKernelGGL(stylize_kernel, dim3((uint32_t)((n + bs - 1) / bs)), dim3(bs), 0, st, in, out, n, depth,
                       depth0, flags, material, pos, nodes, roots, n_materials);
    return hipGetLastError();
}

}  // namespace nlpf

// ---------------------------------------------------------------------------------------
// world-mode records: kernel_mk_trace (lm_path_filtering.rs:1062-1218) on the procedural scene
// ---------------------------------------------------------------------------------------
namespace nlpf {

struct WGenArgs {
    float* rec;
    uint8_t* flags;
    uint8_t* new_flags;
    uint16_t* material;
    int32_t* path_depth;
    uint32_t W, H, L;
    uint64_t N, npix;
    uint64_t seed;
    float alo, ahi;
    const nlpf_style_node* nodes;
    const int32_t* roots;
    uint32_t n_materials;
    const WBsdf* bsdf;  // BSDF by material (nlpf_set_bsdfs); diffuse beyond n_bsdf
    uint32_t n_bsdf;
};

__device__ __forceinline__ int mat_root(const WGenArgs& g, int m) {
    return (g.roots && (uint32_t)m < g.n_materials) ? g.roots[m] : -1;
}

#define LIGHT_EMISSION NLPF_LIGHT_EMISSION
#define LIGHT_AREA (4.0f * SceneConst::lh * SceneConst::lh)

// TrowbridgeReitzDistribution::sample_wh with visible normals (microfacet.rs:118-141) in the
// frame (t, b, n): the sampled half vector in world space
__device__ __forceinline__ v3<float> ggx_sample_vndf(v3<float> t, v3<float> b, v3<float> n, v3<float> wo, float alpha,
                                                     float u1, float u2) {
    v3<float> w{alpha * dot(wo, t), alpha * dot(wo, b), dot(wo, n)};
    w = w / m_sqrt(dot(w, w));
    if (w.z < 0.0f) w = w * -1.0f;
    v3<float> t1{1.0f, 0.0f, 0.0f};
    if (w.z < 0.99999f) {
        t1 = v3<float>{-w.y, w.x, 0.0f};  // (0, 0, 1) x w
        t1 = t1 / m_sqrt(dot(t1, t1));
    }
    v3<float> t2{w.y * t1.z - w.z * t1.y, w.z * t1.x - w.x * t1.z, w.x * t1.y - w.y * t1.x};
    t2 = t2 / m_sqrt(dot(t2, t2));
    const float r = m_sqrt(u1), phi = 6.2831853f * u2;  // uniform_sample_disk (sampling.rs:5-9)
    const float px = r * cosf(phi);
    float py = r * sinf(phi);
    const float hh = m_sqrt(1.0f - px * px), s = 0.5f * (1.0f + w.z);
    py = hh + s * (py - hh);
    const float pz = m_sqrt(m_max(0.0f, 1.0f - px * px - py * py));
    const v3<float> nh = t1 * px + t2 * py + w * pz;
    v3<float> hl{alpha * nh.x, alpha * nh.y, m_max(nh.z, 1e-6f)};
    hl = hl / m_sqrt(dot(hl, hl));
    return t * hl.x + b * hl.y + n * hl.z;
}

__global__ void gen_world_kernel(WGenArgs g) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.N) return;
    const uint64_t px = t % g.npix;
    const uint32_t x = (uint32_t)(px % g.W), y = (uint32_t)(px / g.W);
    const uint64_t key = splitmix64(g.seed);
    Pcg32 rng = pcg_seq_offset(t, key);
    v3<float> o, dir;
    const float jx = pcg_uniform(rng), jy = pcg_uniform(rng);
    camera_ray(g.W, g.H, (float)x + jx, (float)y + jy, o, dir);
    uint32_t flags = 0;
    float prev_bsdf_pdf = 0.0f;
    int32_t last = -1;
    const uint64_t N = g.N;
    for (uint32_t d = 0; d < g.L; ++d) {
        const Hit hit = scene_intersect(o, dir);
        if (hit.mat == NLPF_MAT_SKY) break;
        last = (int32_t)d;
        const v3<float> p = hit.p, ng = hit.n, ns = hit.n;
        const v3<float> wo{-dir.x, -dir.y, -dir.z};
        const int mat = hit.mat;
        const v3<float> R = material_albedo(mat, key, g.alo, g.ahi);
        WBsdf bs{};
        bs.kind = NLPF_BSDF_DIFFUSE;
        if ((uint32_t)mat < g.n_bsdf) bs = g.bsdf[mat];
        // AreaLight::le, one-sided (light/area.rs:36-49); MIS against light sampling for d > 0
        float le = 0.0f, mis = 1.0f;
        if (mat == NLPF_MAT_LIGHT) {
            le = dot(ng, dir) < 0.0f ? LIGHT_EMISSION : 0.0f;
            if (d > 0) {
                const float dist2 = hit.t * hit.t;
                const float lpdf = dist2 / (LIGHT_AREA * m_max(fabsf(dot(ng, dir)), 1e-6f));
                mis = prev_bsdf_pdf / (prev_bsdf_pdf + lpdf);
            }
        }
        const uint32_t new_flags = update_flags(g.nodes, mat_root(g, mat), flags);
        float* r = g.rec + (uint64_t)d * NLPF_WORLD_PLANES * N + t;
        const float base[19] = {p.x, p.y, p.z, ns.x, ns.y, ns.z, ng.x, ng.y, ng.z, wo.x, wo.y, wo.z,
                                R.x, R.y, R.z, le, le, le, mis};
        for (int c = 0; c < 19; ++c) r[(uint64_t)c * N] = base[c];
        g.flags[(uint64_t)d * N + t] = (uint8_t)flags;
        g.new_flags[(uint64_t)d * N + t] = (uint8_t)new_flags;
        g.material[(uint64_t)d * N + t] = (uint16_t)mat;
        flags = new_flags;
        float direct[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // wi, irradiance, pdf, mis
        float bsdf_pdf = 0.0f;
        // u0 selects the principled mixture's lobe (BsdfMixture::sample_wi, svm/surface/mod.rs:1595-1611)
        const float u0 = pcg_uniform(rng), u1 = pcg_uniform(rng), u2 = pcg_uniform(rng);
        const float b1 = pcg_uniform(rng), b2 = pcg_uniform(rng);
        bool cont = false;
        // the light and emission-only materials have no lobe to sample: the path ends there
        if (d + 1 < g.L && mat != NLPF_MAT_LIGHT && bs.kind != NLPF_BSDF_EMISSION) {
            // next-event estimation to the quad light (sample_light :1260-1316, area.rs:51-110)
            const v3<float> lp{-SceneConst::lh + 2.0f * SceneConst::lh * u1, SceneConst::ly,
                               -SceneConst::lh + 2.0f * SceneConst::lh * u2};
            v3<float> wi = lp - p;
            const float dist2 = dot(wi, wi);
            const v3<float> ln{0.0f, -1.0f, 0.0f};
            if (dist2 > 0.0f) {
                const float dist = m_sqrt(dist2);
                wi = wi / dist;
                const float cosl = fabsf(dot(ln, wi));
                if (cosl > 0.0f) {
                    const float e = dot(wi, ln) < 0.0f ? LIGHT_EMISSION : 0.0f;
                    v3<float> li{e, e, e};
                    li = apply_style<float>(g.nodes, mat_root(g, NLPF_MAT_LIGHT), d + 1, new_flags, li, lp);
                    const float lpdf = dist2 / (LIGHT_AREA * cosl);
                    const v3<float> ff = dot(ng, wi) >= 0.0f ? ng : ng * -1.0f;
                    const Hit sh = scene_intersect(p + ff * 1e-4f, wi);
                    const bool occluded = sh.mat != NLPF_MAT_SKY && sh.t < dist * (1.0f - 1e-3f);
                    if (!occluded) {
                        direct[0] = wi.x;
                        direct[1] = wi.y;
                        direct[2] = wi.z;
                        direct[3] = li.x;
                        direct[4] = li.y;
                        direct[5] = li.z;
                        direct[6] = lpdf;
                        direct[7] = 1.0f;
                    }
                }
            }
            v3<float> tt, bb;
            onb(ns, tt, bb);
            v3<float> wi2;
            if (bs.kind == NLPF_BSDF_PRINCIPLED && u0 < bs.metallic) {
                // the metal lobe: a visible normal, and wo reflected about it (mod.rs:1912-1924)
                const v3<float> wh = ggx_sample_vndf(tt, bb, ns, wo, bs.alpha, b1, b2);
                wi2 = wh * (2.0f * dot(wo, wh)) - wo;
            } else {
                // cosine BSDF sample on wo's side (diffuse.rs:42-55)
                const float sr = m_sqrt(b1), phi = 6.2831853f * b2;
                const float lx = sr * cosf(phi), ly = sr * sinf(phi);
                float lz = m_sqrt(m_max(0.0f, 1.0f - lx * lx - ly * ly));
                wi2 = tt * lx + bb * ly + ns * lz;
                if (dot(wo, ns) * lz < 0.0f) wi2 = wi2 * -1.0f;
            }
            // the pdf of the sampled direction is the closure's evaluated pdf (SurfaceClosure::sample)
            if (bs.kind == NLPF_BSDF_PRINCIPLED) {
                const GlossCentre gc = gloss_centre(bs, make_float3(R.x, R.y, R.z));
                float3 fc;
                bsdf_pdf = principled_eval(bs, gc, make_float3(R.x, R.y, R.z), make_float3(ns.x, ns.y, ns.z),
                                           make_float3(ng.x, ng.y, ng.z), make_float3(wo.x, wo.y, wo.z),
                                           make_float3(wi2.x, wi2.y, wi2.z), fc);
            } else {
                const float co = dot(wo, ns), ci = dot(wi2, ns);
                auto sgn = [](float v) { return v > 0.0f ? 1.0f : -1.0f; };
                const float fl = sgn(dot(ng, ns));
                const bool valid =
                    (sgn(fl * co) * sgn(dot(wo, ng)) > 0.0f) && (sgn(fl * ci) * sgn(dot(wi2, ng)) > 0.0f);
                if (valid && co * ci >= 0.0f) bsdf_pdf = fabsf(ci) * 0.318309886f;
            }
            if (bsdf_pdf > 0.0f) {
                const v3<float> ff = dot(ng, wi2) >= 0.0f ? ng : ng * -1.0f;
                o = p + ff * 1e-4f;
                dir = wi2;
                prev_bsdf_pdf = bsdf_pdf;
                cont = true;
            }
        }
        for (int c = 0; c < 8; ++c) r[(uint64_t)(19 + c) * N] = direct[c];
        r[(uint64_t)27 * N] = bsdf_pdf;
        if (!cont) break;
    }
    g.path_depth[t] = last;
}

hipError_t launch_gen_world(const WGenArgs& g, hipStream_t st) {
    const uint32_t bs = 256;
    hipLaunchKernelGGL(gen_world_kernel, dim3((uint32_t)((g.N + bs - 1) / bs)), dim3(bs), 0, st, g);
    return hipGetLastError();
}

}  // namespace nlpf

extern "C" int nlpf_generate_world_records(nlpf_handle* h, const nlpf_world_desc* desc, const nlpf_gen_desc* gen,
                                           float* records, uint8_t* flags, uint8_t* new_flags, uint16_t* material,
                                           int32_t* path_depth) {
    using namespace nlpf;
    if (!h) return NLPF_E_INVALID_ARG;
    if (!desc || !gen || !records || !flags || !new_flags || !material || !path_depth)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "nlpf_generate_world_records: NULL argument");
    if (desc->width == 0 || desc->height == 0 || desc->branch == 0 || desc->levels == 0 || desc->levels > 31)
        NLPF_FAIL(h, NLPF_E_INVALID_ARG, "world desc: bad sizes");
    WGenArgs g{};
    g.rec = records;
    g.flags = flags;
    g.new_flags = new_flags;
    g.material = material;
    g.path_depth = path_depth;
    g.W = desc->width;
    g.H = desc->height;
    g.L = desc->levels;
    g.npix = (uint64_t)desc->width * desc->height;
    g.N = g.npix * desc->branch;
    g.seed = gen->seed;
    g.alo = gen->albedo_lo;
    g.ahi = gen->albedo_hi;
    g.nodes = h->d_nodes;
    g.roots = h->d_roots;
    g.n_materials = (uint32_t)h->h_roots.size();
    g.bsdf = (const WBsdf*)h->d_bsdf;
    g.n_bsdf = (uint32_t)h->h_bsdf.size();
    NLPF_HIP(h, hipMemsetAsync(records, 0, sizeof(float) * NLPF_WORLD_PLANES * g.N * g.L, h->stream));
    NLPF_HIP(h, hipMemsetAsync(flags, 0, g.N * g.L, h->stream));
    NLPF_HIP(h, hipMemsetAsync(new_flags, 0, g.N * g.L, h->stream));
    NLPF_HIP(h, hipMemsetAsync(material, 0, 2 * g.N * g.L, h->stream));
    NLPF_HIP(h, launch_gen_world(g, h->stream));
    NLPF_HIP(h, hipStreamSynchronize(h->stream));
    return NLPF_OK;
}
