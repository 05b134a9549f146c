/*
 * TEST INFRASTRUCTURE ONLY — world-mode BSDF oracle (CPU, FP64).  Loaded only by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
 *
 * Restated from crates/akari_render/src:
 *   DiffuseBsdf::evaluate_impl              svm/surface/diffuse.rs:22-41 (reflectance = color / pi,
 *                                           :101-121)
 *   SurfaceClosure::check_wo_wi_valid       svm/surface/mod.rs:1701-1716
 *   the principled closure                  svm/surface/principled.rs:120-330: for coat_weight 0
 *                                           and either metallic >= 1 - 1e-4 or
 *                                           specular_ior_level == transmission == 0 it is
 *                                           Mix(diffuse, metal, metallic) (BsdfMixture Mix,
 *                                           mod.rs:1560-1593, EPS 1e-4): the specular layer's
 *                                           weight (f0 * 2 * specular_ior_level) and the coat's
 *                                           (coat_weight) are then 0, so no albedo table enters
 *   metal lobe                              MicrofacetReflection (mod.rs:1882-1911) with
 *                                           FresnelComplex (mod.rs:2267-2277, fr_complex
 *                                           :2144-2157) of artistic_to_conductor_fresnel
 *                                           (mod.rs:2128-2142) and TrowbridgeReitz from_roughness
 *                                           (microfacet.rs:29-45, D :46-58, Lambda :60-67, G1/G
 *                                           :9-15, visible-normal pdf :196-201)
 *   complex arithmetic                      util/mod.rs:531-612
 *
 * The reference works in the shading frame (x, y from the tangent).  The roughness is isotropic
 * here (the principled closure passes (r, r)), so D and Lambda depend on cos(theta) alone:
 * cos^2(phi) + sin^2(phi), which the reference forms from the frame's x and y, is 1.  Decisions
 * (validity, hemispheres, zero tests) are taken in FP32 from FP32 world-space dot products, as
 * the library takes them; magnitudes are FP64.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include "oracle.h"

#define OR_PI 3.14159265358979323846
#define MIX_EPS 1e-4f

static float dot32(const double a[3], const double b[3]) {
    float p0 = (float)a[0] * (float)b[0], p1 = (float)a[1] * (float)b[1], p2 = (float)a[2] * (float)b[2];
    float s = p0 + p1;
    return s + p2;
}
static double dot64(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static int wo_wi_valid(const double ns[3], const double ng[3], const double wo[3], const double wi[3]) {
#define SG(x) ((x) > 0 ? 1.0f : -1.0f)
    float fl = SG(dot32(ng, ns));
    return SG(fl * dot32(wo, ns)) * SG(dot32(wo, ng)) > 0 && SG(fl * dot32(wi, ns)) * SG(dot32(wi, ng)) > 0;
#undef SG
}

/* Trowbridge-Reitz D of a half vector at cos(theta_h) = c (isotropic alpha) */
double or_ggx_d(double c, double alpha) {
    const double c2 = c * c;
    const double sin2 = fmax(0.0, 1.0 - c2);
    const double tan2 = sin2 / c2;
    const double e = tan2 / (alpha * alpha);
    const double inv_d = OR_PI * alpha * alpha * c2 * c2 * (1.0 + e) * (1.0 + e);
    if (!isfinite(tan2) || !isfinite(inv_d) || inv_d == 0.0) return 0.0;
    return 1.0 / inv_d;
}
/* Smith Lambda of a direction at cos(theta) = c */
double or_ggx_lambda(double c, double alpha) {
    const double abs_tan = fabs(sqrt(fmax(0.0, 1.0 - c * c)) / c);
    if (!isfinite(abs_tan)) return 0.0;
    const double a2t2 = alpha * alpha * abs_tan * abs_tan;
    return (-1.0 + sqrt(1.0 + a2t2)) * 0.5;
}

typedef struct {
    double re, im;
} cx;
static cx cmul(cx a, cx b) { return (cx){a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
static cx cdiv(cx a, cx b) {
    const double s = 1.0 / (b.re * b.re + b.im * b.im);
    return (cx){(a.re * b.re + a.im * b.im) * s, (a.im * b.re - a.re * b.im) * s};
}
static cx csqrt_ref(cx z) {
    const double n = sqrt(z.re * z.re + z.im * z.im);
    const double t1 = sqrt(0.5 * (n + fabs(z.re)));
    const double t2 = 0.5 * z.im / t1;
    if (n == 0.0) return (cx){0.0, 0.0};
    if (z.re >= 0.0) return (cx){t1, t2};
    return (cx){fabs(t2), copysign(t1, z.im)};
}
/* unpolarised reflectance of a conductor of index n + ik at cos(theta_i) = c (clamped to
   [0, 0.999] as the reference does) */
double or_fr_complex(double c, double n, double k) {
    c = fmin(fmax(c, 0.0), 0.999);
    const double sin2 = 1.0 - c * c;
    const cx eta = {n, k};
    const cx sin2_t = cdiv((cx){sin2, 0.0}, cmul(eta, eta));
    const cx cos_t = csqrt_ref((cx){1.0 - sin2_t.re, -sin2_t.im});
    const cx ec = {eta.re * c, eta.im * c};
    const cx r_parl = cdiv((cx){ec.re - cos_t.re, ec.im - cos_t.im}, (cx){ec.re + cos_t.re, ec.im + cos_t.im});
    const cx ect = cmul(eta, cos_t);
    const cx r_perp = cdiv((cx){c - ect.re, -ect.im}, (cx){c + ect.re, ect.im});
    return (r_parl.re * r_parl.re + r_parl.im * r_parl.im + r_perp.re * r_perp.re + r_perp.im * r_perp.im) * 0.5;
}
/* Gulbrandsen's artistic colour (reflectivity r, edge tint g) to n, k for one channel */
void or_conductor_nk(double r, double g, double* n_out, double* k_out) {
    r = fmin(fmax(r, 0.0), 0.99);
    const double rs = sqrt(r);
    const double n_min = (1.0 - r) / (1.0 + r);
    const double n_max = (1.0 + rs) / (1.0 - rs);
    const double n = n_max + g * (n_min - n_max);
    const double k2 = ((n + 1.0) * (n + 1.0) * r - (n - 1.0) * (n - 1.0)) / (1.0 - r);
    *n_out = n;
    *k_out = sqrt(fmax(k2, 0.0));
}

/* Metal lobe of the principled BSDF (color one): returns its pdf and fills f * |cos_i| per
   channel when fc != NULL; zero where MicrofacetReflection returns zero. */
static double metal_eval(const or_bsdf* b, const double R[3], const double ns[3], const double wo[3],
                         const double wi[3], float co, float ci, double fc[3]) {
    if (fc) fc[0] = fc[1] = fc[2] = 0.0;
    const double whf[3] = {(double)((float)wo[0] + (float)wi[0]), (double)((float)wo[1] + (float)wi[1]),
                           (double)((float)wo[2] + (float)wi[2])};
    const float t1 = dot32(whf, wo), t2 = dot32(wi, whf);
    if (t1 * t2 < 0.0f || (whf[0] == 0.0 && whf[1] == 0.0 && whf[2] == 0.0) || ci == 0.0f || co == 0.0f ||
        !(co * ci >= 0.0f))
        return 0.0;
    const float af = fmaxf(b->roughness * b->roughness, 1e-4f);
    const double alpha = af;
    double h[3] = {wo[0] + wi[0], wo[1] + wi[1], wo[2] + wi[2]};
    const double hl = sqrt(dot64(h, h));
    for (int i = 0; i < 3; ++i) h[i] /= hl;
    const double cho = dot64(wo, ns), chi = dot64(wi, ns), chh = dot64(h, ns);
    const double D = or_ggx_d(chh, alpha);
    const double lo = or_ggx_lambda(cho, alpha), li = or_ggx_lambda(chi, alpha);
    const double G = 1.0 / (1.0 + lo + li), G1 = 1.0 / (1.0 + lo);
    const double woh = fabs(dot64(wo, h));
    if (fc) {
        /* F at wi . face_forward(h, n), taken in absolute value */
        const double cf = fabs(dot64(wi, h));
        const double s = fabs(0.25 * D * G / (chi * cho)) * fabs(chi);
        for (int ch = 0; ch < 3; ++ch) {
            double n, k;
            or_conductor_nk((double)(float)R[ch], (double)b->specular_tint[ch], &n, &k);
            fc[ch] = or_fr_complex(cf, n, k) * s;
        }
    }
    return D * G1 * woh / fabs(cho) / (4.0 * woh);
}

double or_bsdf_eval(const or_bsdf* b, const double R[3], const double ns[3], const double ng[3], const double wo[3],
                    const double wi[3], double fc[3]) {
    if (fc) fc[0] = fc[1] = fc[2] = 0.0;
    const uint32_t kind = b ? b->kind : OR_BSDF_DIFFUSE;
    if (kind == OR_BSDF_EMISSION) return 0.0;
    if (!wo_wi_valid(ns, ng, wo, wi)) return 0.0;
    const float co = dot32(wo, ns), ci = dot32(wi, ns);
    const int same = co * ci >= 0.0f;
    const double cd = fabs(dot64(wi, ns));
    if (kind == OR_BSDF_DIFFUSE) {
        if (!same) return 0.0;
        if (fc)
            for (int ch = 0; ch < 3; ++ch) fc[ch] = R[ch] * cd / OR_PI;
        return cd / OR_PI;
    }
    /* principled: Mix(diffuse, metal, metallic) with the EPS gates of BsdfMixture */
    const float m = b->metallic;
    double fd[3] = {0, 0, 0}, fm[3] = {0, 0, 0}, pd = 0.0, pm = 0.0;
    if (m < 1.0f - MIX_EPS && same) {
        pd = cd / OR_PI;
        for (int ch = 0; ch < 3; ++ch) fd[ch] = R[ch] * cd / OR_PI;
    }
    if (m > MIX_EPS) pm = metal_eval(b, R, ns, wo, wi, co, ci, fc ? fm : NULL);
    if (fc)
        for (int ch = 0; ch < 3; ++ch) fc[ch] = (1.0 - m) * fd[ch] + m * fm[ch];
    return (1.0 - m) * pd + m * pm;
}
