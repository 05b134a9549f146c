// BSDF evaluation of world-mode neighbour records, shared by the filter kernels: the reuse
// payload (NbrRec), check_wo_wi_valid and the diffuse pdf with the reference's FP32 decisions,
// the principled subset (diffuse / metal mixture) of the material table, and the split of
// pdf_m(wi) into a per-neighbour part (SmisRec, SmisGloss) and a per-direction part (smis_add,
// smis_add_gloss) for the SMIS denominators.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "nlpf.h"

namespace nlpf {

// per-path reuse payload, gathered once per level so neighbour reads are 16-byte loads
struct __align__(16) NbrRec {
    float4 dwi_pdf;   // direct.wi, direct.pdf
    float4 dirr_mis;  // direct.irradiance, direct.mis
    float4 iwi_mis;   // indirect.wi, indirect.mis
    float4 iirr;      // indirect.irradiance, flags-equal key in w (as bits)
    float4 ns_wox;    // ns, wo.x
    float4 ng_woy;    // ng, wo.y
    float4 woz;       // wo.z, material id (as bits), -, -
};

// dot product without fused multiply-adds: the reference decides hemisphere and leak tests
// in plain FP32, so every sign test below rounds exactly like that arithmetic (the pragma is
// what keeps the compiler from contracting; the _rn intrinsics alone do not)
__device__ __forceinline__ float dot_rn(float3 a, float3 b) {
#pragma clang fp contract(off)
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
// SurfaceClosure::check_wo_wi_valid (svm/surface/mod.rs:1701-1716)
__device__ __forceinline__ bool wo_wi_valid(float3 ns, float3 ng, float3 wo, float3 wi) {
    auto sgn = [](float x) { return x > 0.0f ? 1.0f : -1.0f; };
    const float flipped = sgn(dot_rn(ng, ns));
    return (sgn(flipped * dot_rn(wo, ns)) * sgn(dot_rn(wo, ng)) > 0.0f) &&
           (sgn(flipped * dot_rn(wi, ns)) * sgn(dot_rn(wi, ng)) > 0.0f);
}
// DiffuseBsdf::evaluate_impl (svm/surface/diffuse.rs:22-41) through SurfaceClosure::evaluate
// (mod.rs:1725-1744): returns the pdf; fcos = |cos_i| / pi (multiply by R for f*cos)
__device__ __forceinline__ float diffuse_eval(float3 ns, float3 ng, float3 wo, float3 wi, float& fcos) {
    fcos = 0.0f;
    if (!wo_wi_valid(ns, ng, wo, wi)) return 0.0f;
    const float co = dot_rn(wo, ns);
    const float ci = dot_rn(wi, ns);
    if (!(co * ci >= 0.0f)) return 0.0f;
    // |cos| near grazing is a cancellation-prone difference of products: take its magnitude
    // from an FP64 dot (the FP32 value above only decides the hemisphere, as in the reference)
    const double cd = (double)wi.x * ns.x + (double)wi.y * ns.y + (double)wi.z * ns.z;
    const float pdf = (float)(fabs(cd) * 0.318309886183790671538);
    fcos = pdf;
    return pdf;
}

// What pdf_m(wi) of a reused neighbour m needs, apart from wi (diffuse_eval split in two): its
// FP32 normals and cos_o for the reference's FP32 decisions, and three flags. Two 16-byte
// words: the SMIS loops read one record per step for the whole wavefront from LDS, and the
// bytes a record returns to 64 lanes are what bounds them. The FP64 shading normal for
// |cos_i| is the exact widening of the FP32 one, formed in registers.
struct __align__(16) SmisRec {
    float nsx, nsy, nsz, co;
    float ngx, ngy, ngz;
    uint32_t bits;
};
static_assert(sizeof(SmisRec) == 32, "SmisRec layout");
// flags: ns and ng are the same bits (flat shading), and sgn(ng . ns) > 0 (so that
// wi . ng = wi . ns, flipped = +1, and the wi half of check_wo_wi_valid always holds)
constexpr uint32_t SMIS_SAME = 1u;
// the wo half of check_wo_wi_valid holds (otherwise pdf_m(wi) = 0 for every wi)
constexpr uint32_t SMIS_WO_OK = 2u;
// flipped = sgn(ng . ns) is -1
constexpr uint32_t SMIS_FLIP = 4u;
// the neighbour's BSDF is a principled one: its SmisGloss record holds the rest
constexpr uint32_t SMIS_GLOSSY = 8u;

__device__ __forceinline__ SmisRec smis_rec(const NbrRec& o) {
    const float3 ns = make_float3(o.ns_wox.x, o.ns_wox.y, o.ns_wox.z);
    const float3 ng = make_float3(o.ng_woy.x, o.ng_woy.y, o.ng_woy.z);
    const float3 wo = make_float3(o.ns_wox.w, o.ng_woy.w, o.woz.x);
    auto sgn = [](float x) { return x > 0.0f ? 1.0f : -1.0f; };
    const float flipped = sgn(dot_rn(ng, ns));
    const float co = dot_rn(wo, ns);
    const bool wo_ok = sgn(flipped * co) * sgn(dot_rn(wo, ng)) > 0.0f;
    // (flipped > 0 as well: a normal so short that ns . ns underflows keeps the general path)
    const bool same = (__float_as_uint(ns.x) == __float_as_uint(ng.x)) &
                      (__float_as_uint(ns.y) == __float_as_uint(ng.y)) &
                      (__float_as_uint(ns.z) == __float_as_uint(ng.z)) & (flipped > 0.0f);
    SmisRec r;
    r.nsx = ns.x;
    r.nsy = ns.y;
    r.nsz = ns.z;
    r.co = co;
    r.ngx = ng.x;
    r.ngy = ng.y;
    r.ngz = ng.z;
    r.bits = (same ? SMIS_SAME : 0u) | (wo_ok ? SMIS_WO_OK : 0u) | (flipped < 0.0f ? SMIS_FLIP : 0u);
    return r;
}
// sum + |cos_i| of pdf_m(wi) (diffuse_eval without the 1/pi), the term being 0 where pdf_m(wi)
// is 0: the same FP32 sign and hemisphere decisions, the same FP64 magnitude; branch-free.
// `bits` is r.bits (passed apart so a caller can hold it in a scalar register). The term
// enters as fma(|cd|, 1 or 0, sum), which rounds exactly like sum + term (|cd| * 1 and
// |cd| * 0 are exact for the finite cd of unit vectors).
__device__ __forceinline__ double smis_add(double sum, const SmisRec& r, uint32_t bits, float3 wi, double wx,
                                           double wy, double wz) {
    const float ci = dot_rn(wi, make_float3(r.nsx, r.nsy, r.nsz));
    const float cg = dot_rn(wi, make_float3(r.ngx, r.ngy, r.ngz));
    // sgn(flipped * ci) * sgn(cg) > 0: both signs agree (flipped * ci is +-ci exactly); &
    // rather than && so nothing branches
    const bool ok = ((((bits & SMIS_FLIP) ? -ci : ci) > 0.0f) == (cg > 0.0f)) & (r.co * ci >= 0.0f) &
                    ((bits & SMIS_WO_OK) != 0u);
    const double cd = wx * (double)r.nsx + wy * (double)r.nsy + wz * (double)r.nsz;
    return fma(fabs(cd), ok ? 1.0 : 0.0, sum);
}
// smis_add for a record with SMIS_SAME: the sign test reduces to the hemisphere test
__device__ __forceinline__ double smis_add_same(double sum, const SmisRec& r, uint32_t bits, float3 wi, double wx,
                                                double wy, double wz) {
    const float ci = dot_rn(wi, make_float3(r.nsx, r.nsy, r.nsz));
    const bool ok = (r.co * ci >= 0.0f) & ((bits & SMIS_WO_OK) != 0u);
    const double cd = wx * (double)r.nsx + wy * (double)r.nsy + wz * (double)r.nsz;
    return fma(fabs(cd), ok ? 1.0 : 0.0, sum);
}

// ---- the principled subset (svm/surface/principled.rs:120-330 with coat 0 and no specular
// layer, or metallic 1): Mix(diffuse, metal, metallic) ----

// a material's BSDF as the filter uses it (host: nlpf_set_bsdfs)
struct __align__(16) WBsdf {
    uint32_t kind;   // NLPF_BSDF_*
    float metallic;
    float alpha;     // max(roughness^2, 1e-4) in FP32 (TrowbridgeReitzDistribution::from_roughness)
    uint32_t pad;
    float tint[4];   // specular_tint: the edge tint of the conductor parametrisation
};
static_assert(sizeof(WBsdf) == 32, "WBsdf layout");
constexpr float MIX_EPS = 1e-4f;  // BsdfMixture::EPS (svm/surface/mod.rs:1548-1550)
constexpr double PI_D = 3.14159265358979323846;

// Trowbridge-Reitz D at cos(theta_h) = c (microfacet.rs:46-58; isotropic: cos^2 phi + sin^2 phi = 1)
__device__ __forceinline__ double ggx_d(double c, double alpha) {
    const double c2 = c * c;
    const double tan2 = fmax(0.0, 1.0 - c2) / c2;
    const double e = tan2 / (alpha * alpha);
    const double inv_d = PI_D * alpha * alpha * c2 * c2 * (1.0 + e) * (1.0 + e);
    return (!isfinite(tan2) || !isfinite(inv_d) || inv_d == 0.0) ? 0.0 : 1.0 / inv_d;
}
// Smith Lambda at cos(theta) = c (microfacet.rs:60-67)
__device__ __forceinline__ double ggx_lambda(double c, double alpha) {
    const double abs_tan = fabs(sqrt(fmax(0.0, 1.0 - c * c)) / c);
    if (!isfinite(abs_tan)) return 0.0;
    return (-1.0 + sqrt(1.0 + alpha * alpha * abs_tan * abs_tan)) * 0.5;
}
// unpolarised conductor reflectance for index n + ik at cos(theta_i) = c, clamped to
// [0, 0.999] (fr_complex, svm/surface/mod.rs:2144-2157, complex arithmetic util/mod.rs:531-612)
__device__ __forceinline__ double fr_conductor(double c, double n, double k) {
    c = fmin(fmax(c, 0.0), 0.999);
    const double s2 = 1.0 - c * c;
    // sin^2 t = s2 / eta^2
    const double e2r = n * n - k * k, e2i = 2.0 * n * k;
    const double inv = 1.0 / (e2r * e2r + e2i * e2i);
    const double zr = 1.0 - s2 * e2r * inv, zi = s2 * e2i * inv;  // 1 - sin^2 t
    // principal square root, as Complex::sqrt
    const double an = sqrt(zr * zr + zi * zi);
    double tr = 0.0, ti = 0.0;
    if (an != 0.0) {
        const double t1 = sqrt(0.5 * (an + fabs(zr))), t2 = 0.5 * zi / t1;
        if (zr >= 0.0) {
            tr = t1;
            ti = t2;
        } else {
            tr = fabs(t2);
            ti = copysign(t1, zi);
        }
    }
    auto norm_ratio = [](double ar, double ai, double br, double bi) {  // |a / b|^2
        return (ar * ar + ai * ai) / (br * br + bi * bi);
    };
    // r_parl = (eta c - cos t) / (eta c + cos t), r_perp = (c - eta cos t) / (c + eta cos t)
    const double ecr = n * c, eci = k * c;
    const double etr = n * tr - k * ti, eti = n * ti + k * tr;
    return 0.5 * (norm_ratio(ecr - tr, eci - ti, ecr + tr, eci + ti) + norm_ratio(c - etr, -eti, c + etr, eti));
}
// artistic_to_conductor_fresnel for one channel (svm/surface/mod.rs:2128-2142)
__device__ __forceinline__ void conductor_nk(double r, double g, double& n, double& k) {
    r = fmin(fmax(r, 0.0), 0.99);
    const double rs = sqrt(r);
    const double n_min = (1.0 - r) / (1.0 + r), n_max = (1.0 + rs) / (1.0 - rs);
    n = n_max + g * (n_min - n_max);
    k = sqrt(fmax(((n + 1.0) * (n + 1.0) * r - (n - 1.0) * (n - 1.0)) / (1.0 - r), 0.0));
}

// per-vertex part of a principled centre: the conductor's n, k per channel
struct GlossCentre {
    double n[3], k[3];
};
__device__ __forceinline__ GlossCentre gloss_centre(const WBsdf& b, float3 R) {
    GlossCentre g;
    conductor_nk(R.x, b.tint[0], g.n[0], g.k[0]);
    conductor_nk(R.y, b.tint[1], g.n[1], g.k[1]);
    conductor_nk(R.z, b.tint[2], g.n[2], g.k[2]);
    return g;
}

// MicrofacetReflection::evaluate_impl (svm/surface/mod.rs:1882-1911), colour one, in world space
// on the shading normal: the FP32 zero tests of the reference, FP64 magnitudes. Returns the pdf;
// fc = f |cos_i| per channel.
__device__ __forceinline__ double metal_eval(const WBsdf& b, const GlossCentre& gc, float3 ns, float3 wo, float3 wi,
                                             float co, float ci, double3& fc) {
    fc = make_double3(0.0, 0.0, 0.0);
    float3 whf;
    {
#pragma clang fp contract(off)
        whf = make_float3(wo.x + wi.x, wo.y + wi.y, wo.z + wi.z);
    }
    const float t1 = dot_rn(whf, wo), t2 = dot_rn(wi, whf);
    bool zero;
    {
#pragma clang fp contract(off)
        zero = (t1 * t2 < 0.0f) | ((whf.x == 0.0f) & (whf.y == 0.0f) & (whf.z == 0.0f)) | (ci == 0.0f) | (co == 0.0f) |
               !(co * ci >= 0.0f);
    }
    if (zero) return 0.0;
    const double alpha = b.alpha;
    double hx = (double)wo.x + wi.x, hy = (double)wo.y + wi.y, hz = (double)wo.z + wi.z;
    const double hl = 1.0 / sqrt(hx * hx + hy * hy + hz * hz);
    hx *= hl;
    hy *= hl;
    hz *= hl;
    const double cho = (double)wo.x * ns.x + (double)wo.y * ns.y + (double)wo.z * ns.z;
    const double chi = (double)wi.x * ns.x + (double)wi.y * ns.y + (double)wi.z * ns.z;
    const double chh = hx * ns.x + hy * ns.y + hz * ns.z;
    const double D = ggx_d(chh, alpha);
    const double lo = ggx_lambda(cho, alpha), li = ggx_lambda(chi, alpha);
    const double G = 1.0 / (1.0 + lo + li), G1 = 1.0 / (1.0 + lo);
    const double cf = fabs(hx * wi.x + hy * wi.y + hz * wi.z);
    const double s = 0.25 * D * G / fabs(cho);
    fc = make_double3(fr_conductor(cf, gc.n[0], gc.k[0]) * s, fr_conductor(cf, gc.n[1], gc.k[1]) * s,
                      fr_conductor(cf, gc.n[2], gc.k[2]) * s);
    return D * G1 / (4.0 * fabs(cho));
}

// the principled subset behind SurfaceClosure::evaluate: returns the pdf; fc = f |cos_i| per
// channel, base colour R included (the diffuse part is DiffuseBsdf with reflectance R / pi)
__device__ __forceinline__ float principled_eval(const WBsdf& b, const GlossCentre& gc, float3 R, float3 ns, float3 ng,
                                                 float3 wo, float3 wi, float3& fc) {
    fc = make_float3(0.0f, 0.0f, 0.0f);
    if (!wo_wi_valid(ns, ng, wo, wi)) return 0.0f;
    const float co = dot_rn(wo, ns), ci = dot_rn(wi, ns);
    const float m = b.metallic;
    double pd = 0.0, fdx = 0.0, fdy = 0.0, fdz = 0.0;
    if (m < 1.0f - MIX_EPS && co * ci >= 0.0f) {
        const double cd = fabs((double)wi.x * ns.x + (double)wi.y * ns.y + (double)wi.z * ns.z) / PI_D;
        pd = cd;
        fdx = R.x * cd;
        fdy = R.y * cd;
        fdz = R.z * cd;
    }
    double pm = 0.0;
    double3 fm = make_double3(0.0, 0.0, 0.0);
    if (m > MIX_EPS) pm = metal_eval(b, gc, ns, wo, wi, co, ci, fm);
    const double md = m, w = 1.0 - md;
    fc = make_float3((float)(w * fdx + md * fm.x), (float)(w * fdy + md * fm.y), (float)(w * fdz + md * fm.z));
    return (float)(w * pd + md * pm);
}

// What a principled neighbour's pdf_m(wi) needs beyond its SmisRec: its wo and alpha, and the
// two mixture weights in units of pi * pdf (the SMIS sums are sums of |cos| for diffuse terms)
struct __align__(16) SmisGloss {
    float wox, woy, woz, alpha;
    double km;  // metallic * pi * G1(wo) / (4 |cos_o|), 0 when the metal side is skipped or cos_o = 0
    double kd;  // 1 - metallic, 0 when the diffuse side is skipped
};
static_assert(sizeof(SmisGloss) == 32, "SmisGloss layout");
__device__ __forceinline__ SmisGloss smis_gloss(const NbrRec& o, const WBsdf& b, float co) {
    SmisGloss g;
    g.wox = o.ns_wox.w;
    g.woy = o.ng_woy.w;
    g.woz = o.woz.x;
    g.alpha = b.alpha;
    const double cho = (double)g.wox * o.ns_wox.x + (double)g.woy * o.ns_wox.y + (double)g.woz * o.ns_wox.z;
    const float m = b.metallic;
    g.km = (m > MIX_EPS && co != 0.0f)
               ? (double)m * PI_D / ((1.0 + ggx_lambda(cho, (double)b.alpha)) * 4.0 * fabs(cho))
               : 0.0;
    g.kd = m < 1.0f - MIX_EPS ? 1.0 - (double)m : 0.0;
    return g;
}
// smis_add for a principled neighbour: sum + pi * pdf_m(wi), the same FP32 decisions as
// smis_add plus MicrofacetReflection's zero tests, FP64 magnitudes
__device__ __forceinline__ double smis_add_gloss(double sum, const SmisRec& r, const SmisGloss& g, uint32_t bits,
                                                 float3 wi, double wx, double wy, double wz) {
    const float ci = dot_rn(wi, make_float3(r.nsx, r.nsy, r.nsz));
    const float cg = dot_rn(wi, make_float3(r.ngx, r.ngy, r.ngz));
    const bool same = r.co * ci >= 0.0f;
    const bool ok = ((((bits & SMIS_FLIP) ? -ci : ci) > 0.0f) == (cg > 0.0f)) & same & ((bits & SMIS_WO_OK) != 0u);
    if (!ok) return sum;
    const double cd = wx * (double)r.nsx + wy * (double)r.nsy + wz * (double)r.nsz;
    double t = g.kd * fabs(cd);
    if (g.km != 0.0) {
        const float3 wo = make_float3(g.wox, g.woy, g.woz);
        float3 whf;
        {
#pragma clang fp contract(off)
            whf = make_float3(wo.x + wi.x, wo.y + wi.y, wo.z + wi.z);
        }
        const float t1 = dot_rn(whf, wo), t2 = dot_rn(wi, whf);
        bool zero;
        {
#pragma clang fp contract(off)
            zero = (t1 * t2 < 0.0f) | ((whf.x == 0.0f) & (whf.y == 0.0f) & (whf.z == 0.0f)) | (ci == 0.0f);
        }
        if (!zero) {
            const double hx = (double)wo.x + wx, hy = (double)wo.y + wy, hz = (double)wo.z + wz;
            const double chh = (hx * r.nsx + hy * r.nsy + hz * r.nsz) / sqrt(hx * hx + hy * hy + hz * hz);
            t += g.km * ggx_d(chh, (double)g.alpha);
        }
    }
    return sum + t;
}

}  // namespace nlpf
