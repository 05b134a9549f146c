// Lambertian BSDF evaluation of world-mode neighbour records, shared by the filter kernels:
// the reuse payload (NbrRec), check_wo_wi_valid and the diffuse pdf with the reference's FP32
// decisions, and the split of pdf_m(wi) into a per-neighbour part (SmisRec) and a
// per-direction part (smis_term) for the SMIS denominators.
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
// FP32 normals and cos_o for the reference's FP32 decisions, sgn(ng . ns) as +-1, and its FP64
// shading normal for |cos_i|. A neighbour whose wo half of check_wo_wi_valid fails gets a zero
// FP64 normal, so its term is 0 for every wi without a test per pair. Four 16-byte words, so a
// staged record loads with four 128-bit LDS reads.
struct __align__(16) SmisRec {
    double nsx, nsy;
    double nsz;
    float co, fl;
    float nsfx, nsfy, nsfz, ngx;
    float ngy, ngz, pad0, pad1;
};
static_assert(sizeof(SmisRec) == 64, "SmisRec layout");

__device__ __forceinline__ SmisRec smis_rec(const NbrRec& o) {
    const float3 ns = make_float3(o.ns_wox.x, o.ns_wox.y, o.ns_wox.z);
    const float3 ng = make_float3(o.ng_woy.x, o.ng_woy.y, o.ng_woy.z);
    const float3 wo = make_float3(o.ns_wox.w, o.ng_woy.w, o.woz.x);
    auto sgn = [](float x) { return x > 0.0f ? 1.0f : -1.0f; };
    const float flipped = sgn(dot_rn(ng, ns));
    const float co = dot_rn(wo, ns);
    const bool wo_ok = sgn(flipped * co) * sgn(dot_rn(wo, ng)) > 0.0f;
    SmisRec r;
    r.nsx = wo_ok ? (double)ns.x : 0.0;
    r.nsy = wo_ok ? (double)ns.y : 0.0;
    r.nsz = wo_ok ? (double)ns.z : 0.0;
    r.co = co;
    r.fl = flipped;
    r.nsfx = ns.x;
    r.nsfy = ns.y;
    r.nsfz = ns.z;
    r.ngx = ng.x;
    r.ngy = ng.y;
    r.ngz = ng.z;
    r.pad0 = r.pad1 = 0.0f;
    return r;
}
// |cos_i| of pdf_m(wi) (diffuse_eval without the 1/pi), 0 where pdf_m(wi) is 0: the same FP32
// sign and hemisphere decisions, the same FP64 magnitude; branch-free
__device__ __forceinline__ double smis_term(const SmisRec& r, float3 wi, double wx, double wy, double wz) {
    const float ci = dot_rn(wi, make_float3(r.nsfx, r.nsfy, r.nsfz));
    const float cg = dot_rn(wi, make_float3(r.ngx, r.ngy, r.ngz));
    // sgn(flipped * ci) * sgn(cg) > 0: both signs agree; & rather than && so nothing branches
    const bool ok = ((r.fl * ci > 0.0f) == (cg > 0.0f)) & (r.co * ci >= 0.0f);
    const double cd = wx * r.nsx + wy * r.nsy + wz * r.nsz;
    // both arms use cd (0 * cd is a signed zero for the finite cd of unit vectors), so the
    // select needs no branch around the FP64 dot
    return ok ? fabs(cd) : 0.0 * cd;
}

}  // namespace nlpf
