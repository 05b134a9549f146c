// Lambertian BSDF evaluation of world-mode neighbour records, shared by the filter kernels:
// the reuse payload (NbrRec), check_wo_wi_valid and the diffuse pdf with the reference's FP32
// decisions, and the split of pdf_m(wi) into a per-neighbour part (SmisRec) and a
// per-direction part (smis_add) for the SMIS denominators.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nlpf {

// per-path reuse payload, gathered once per level so neighbour reads are 16-byte loads
struct __align__(16) NbrRec {
    float4 dwi_pdf;   // direct.wi, direct.pdf
    float4 dirr_mis;  // direct.irradiance, direct.mis
    float4 iwi_mis;   // indirect.wi, indirect.mis
    float4 iirr;      // indirect.irradiance, flags-equal key in w (as bits)
    float4 ns_wox;    // ns, wo.x
    float4 ng_woy;    // ng, wo.y
    float4 woz;       // wo.z, -, -, -
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

}  // namespace nlpf
