// Stylization nonlinearities evaluated on the GPU (templated on float / double).
//
// Each function restates the arithmetic of one reference wrapper surface; the
// reference applies the wrapped (inner) surface first and then its own transform,
// and the tree walk in apply_style() keeps that order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/nlpf.h"

#define NLPF_HD __host__ __device__ __forceinline__

namespace nlpf {

template <typename T>
struct v3 {
    T x, y, z;
};
template <typename T> NLPF_HD v3<T> mk(T a, T b, T c) { return v3<T>{a, b, c}; }
template <typename T> NLPF_HD v3<T> operator+(v3<T> a, v3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> NLPF_HD v3<T> operator-(v3<T> a, v3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> NLPF_HD v3<T> operator*(v3<T> a, v3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <typename T> NLPF_HD v3<T> operator*(v3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> NLPF_HD v3<T> operator/(v3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T> NLPF_HD T dot(v3<T> a, v3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

NLPF_HD float m_pow(float a, float b) { return powf(a, b); }
NLPF_HD double m_pow(double a, double b) { return pow(a, b); }
NLPF_HD float m_floor(float a) { return floorf(a); }
NLPF_HD double m_floor(double a) { return floor(a); }
NLPF_HD float m_round(float a) { return roundf(a); }
NLPF_HD double m_round(double a) { return round(a); }
NLPF_HD float m_fmod(float a, float b) { return fmodf(a, b); }
NLPF_HD double m_fmod(double a, double b) { return fmod(a, b); }
NLPF_HD float m_sqrt(float a) { return sqrtf(a); }
NLPF_HD double m_sqrt(double a) { return sqrt(a); }
// Quotients inside the stylization nodes.  On the GPU an FP32 quotient is a * rcp(b) (v_rcp_f32,
// within 1 ulp; the product within 2 ulp of the correctly rounded quotient, against an FP32
// tolerance of 1e-5): a correctly rounded division takes about ten instructions.  Every
// divisor here is a luminance, a colour channel difference, a ramp segment or a constant, far
// from the denormal range.  FP64 (the error oracle) and host code divide exactly.
template <typename T> NLPF_HD T m_div(T a, T b) { return a / b; }
NLPF_HD float m_div(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return a * __builtin_amdgcn_rcpf(b);
#else
    return a / b;
#endif
}
template <typename T> NLPF_HD T m_min(T a, T b) { return a < b ? a : b; }
template <typename T> NLPF_HD T m_max(T a, T b) { return a > b ? a : b; }
template <typename T> NLPF_HD T m_clamp(T v, T lo, T hi) { return m_min(m_max(v, lo), hi); }
template <typename T> NLPF_HD v3<T> clamp3(v3<T> v, T lo, T hi) {
    return {m_clamp(v.x, lo, hi), m_clamp(v.y, lo, hi), m_clamp(v.z, lo, hi)};
}
template <typename T> NLPF_HD T hmax(v3<T> v) { return m_max(m_max(v.x, v.y), v.z); }
template <typename T> NLPF_HD T hmin(v3<T> v) { return m_min(m_min(v.x, v.y), v.z); }
// f64::rem_euclid / f32::rem_euclid for m > 0.  fmod(a, m) is exact, and for a in (-m, 2m) it
// is a or a - m (exact by Sterbenz), so that range skips the general remainder sequence with
// the same result bit for bit; the HSV conversions only ever pass values in it.
template <typename T> NLPF_HD T rem_euclid(T a, T m) {
    T r;
    if (a > -m && a < m + m)
        r = a >= m ? a - m : a;
    else
        r = m_fmod(a, m);
    return r < T(0) ? r + m : r;
}

// Color::lum, Rec.709 weights (crates/akari_render/src/color.rs:356-361)
template <typename T> NLPF_HD T lum709(v3<T> c) {
    return c.x * T(0.2126) + c.y * T(0.7152) + c.z * T(0.0722);
}
// IEC 61966-2-1 transfer functions (color.rs:531-549)
template <typename T> NLPF_HD T srgb_to_linear1(T v) {
    return v <= T(0.04045) ? v / T(12.92) : m_pow((v + T(0.055)) / T(1.055), T(2.4));
}
template <typename T> NLPF_HD T linear_to_srgb1(T v) {
    return v <= T(0.0031308) ? v * T(12.92) : m_pow(v, T(1) / T(2.4)) * T(1.055) - T(0.055);
}
template <typename T> NLPF_HD v3<T> srgb_to_linear(v3<T> c) {
    return {srgb_to_linear1(c.x), srgb_to_linear1(c.y), srgb_to_linear1(c.z)};
}
template <typename T> NLPF_HD v3<T> linear_to_srgb(v3<T> c) {
    return {linear_to_srgb1(c.x), linear_to_srgb1(c.y), linear_to_srgb1(c.z)};
}
// HSV conversions (color.rs:551-589)
// The three hue sectors differ only in the numerator and the offset: they are selected before
// the one division, so every lane runs the same instructions (same arithmetic as branching).
template <typename T> NLPF_HD v3<T> rgb_to_hsv(v3<T> c) {
    const T mx = hmax(c), mn = hmin(c), diff = mx - mn;
    const bool rx = mx == c.x, ry = !rx && mx == c.y;
    const T num = rx ? c.y - c.z : (ry ? c.z - c.x : c.x - c.y);
    const T q = m_div(num, diff);
    T h = rx ? rem_euclid(q, T(6)) : q + (ry ? T(2) : T(4));
    if (mx == mn) h = T(0);
    h = m_div(h, T(6));
    const T s = mx == T(0) ? T(0) : m_div(diff, mx);
    return {h, s, mx};
}
template <typename T> NLPF_HD T hsv_channel(T n, v3<T> hsv) {
    T k = rem_euclid(n + hsv.x * T(6), T(6));
    T t = m_max(m_min(m_min(k, T(4) - k), T(1)), T(0));
    return hsv.z - hsv.z * hsv.y * t;
}
template <typename T> NLPF_HD v3<T> hsv_to_rgb(v3<T> hsv) {
    return {hsv_channel(T(5), hsv), hsv_channel(T(3), hsv), hsv_channel(T(1), hsv)};
}

// StylizationFlags::get_style_index (svm/surface/mod.rs:80-93)
NLPF_HD uint32_t style_index(uint32_t flags) {
    uint32_t i = 0;
    if (flags & NLPF_FLAG_STYLE2) i = 1;
    if (flags & NLPF_FLAG_STYLE3) i = 2;
    if (flags & NLPF_FLAG_STYLE4) i = 3;
    return i;
}

// ToonSurface::apply_stylization (svm/surface/mod.rs:572-610), own transform only
template <typename T> NLPF_HD v3<T> toon(v3<T> li, T w) {
    li = clamp3(li, T(0), T(100000));
    T l = lum709(li);
    if (l <= T(0)) return li;
    v3<T> hue = li / l;
    T nl;
    if (l > T(1)) {
        nl = m_floor(l / w) * w;
    } else {
        T a = linear_to_srgb1(l);
        nl = srgb_to_linear1(m_floor(a / w) * w);
    }
    return hue * nl;
}

// colour ramp (no reference node): hue-preserving piecewise-linear luminance remap.
// Segment k is the first stop with !(pos[k] <= x), n when every stop is <= x.  The stops
// ascend (nlpf_set_style rejects other tables), so k is the number of stops <= x, found by a
// branch-free binary search in power-of-two steps (the same k as the linear search, NaN
// included: every comparison is false and k = 0).
#define NLPF_RAMP_STOPS 16
template <typename T, typename P> NLPF_HD uint32_t ramp_segment(P pos, uint32_t n, T x) {
    uint32_t k = 0;
#pragma unroll
    for (uint32_t step = NLPF_RAMP_STOPS; step > 0; step >>= 1) {
        const uint32_t j = k + step;
        if (j <= n && T(pos[j - 1]) <= x) k = j;
    }
    return k;
}
template <typename T, typename P> NLPF_HD T ramp_eval(P pos, P val, uint32_t n, T x) {
    if (n == 0) return x;
    if (n > NLPF_RAMP_STOPS) n = NLPF_RAMP_STOPS;
    const uint32_t k = ramp_segment(pos, n, x);
    if (k == 0) return T(val[0]);
    if (k == n) return T(val[n - 1]);
    const T x0 = T(pos[k - 1]), x1 = T(pos[k]);
    const T t = m_div(x - x0, x1 - x0);
    return T(val[k - 1]) + t * (T(val[k]) - T(val[k - 1]));
}
template <typename T, typename P> NLPF_HD v3<T> ramp(v3<T> li, P pos, P val, uint32_t n) {
    li = clamp3(li, T(0), T(100000));
    T l = lum709(li);
    if (l <= T(0)) return li;
    return li * m_div(ramp_eval(pos, val, n, l), l);
}

// sty_hue_shift: hue rotation in the reference's HSV (no clamping, so HDR-safe)
template <typename T> NLPF_HD v3<T> hue_shift(v3<T> li, T shift) {
    v3<T> hsv = rgb_to_hsv(li);
    hsv.x = rem_euclid(hsv.x + shift, T(1));
    return hsv_to_rgb(hsv);
}

// SaturationAdjustSurface::apply_stylization (svm/surface/mod.rs:820-851)
template <typename T> NLPF_HD v3<T> saturation(v3<T> li, uint32_t depth, T power, T iscale) {
    v3<T> rgb = li;
    if (hmax(rgb) == T(0)) return li;
    if (depth != 0) rgb = rgb / iscale;
    rgb = clamp3(rgb, T(0), T(1));
    rgb = linear_to_srgb(rgb);
    v3<T> hsv = rgb_to_hsv(rgb);
    hsv.y = m_pow(hsv.y, power);
    rgb = srgb_to_linear(hsv_to_rgb(hsv));
    if (depth != 0) rgb = rgb * iscale;
    return rgb;
}

// SaturationAdjustRexSurface::apply_stylization (svm/surface/mod.rs:1059-1089)
template <typename T> NLPF_HD v3<T> rex(v3<T> li, T blend) {
    T l = (li.x + li.y + li.z) / T(3);
    if (l <= T(0) || hmin(li) < T(0)) return li;
    T cmin = hmin(li), cmax = hmax(li);
    v3<T> imax = (li - mk(cmin, cmin, cmin)) / cmax;
    T u = (imax.x + imax.y + imax.z) / T(3);
    v3<T> base = mk(blend * l, blend * l, blend * l);
    if (u != T(0)) return base + imax * (T(1) - blend) * l / u;
    T o = (T(1) - blend) * l;
    return base + mk(o, o, o);
}

// ColorPaletteSurface::apply_stylization (svm/surface/mod.rs:696-735)
template <typename T> NLPF_HD v3<T> palette(v3<T> li, uint32_t depth, T iscale, T gamma,
                                           const float* pal, uint32_t n) {
    v3<T> rgb = li;
    T l = (rgb.x + rgb.y + rgb.z) / T(3);
    if (depth != 0) rgb = rgb / iscale;
    rgb = clamp3(rgb, T(0), T(1));
    rgb = mk(m_pow(rgb.x, gamma), m_pow(rgb.y, gamma), m_pow(rgb.z, gamma));
    T best = T(1e16);
    v3<T> bc = mk(T(0), T(0), T(0));
    for (uint32_t i = 0; i < n; ++i) {
        v3<T> c = mk(T(pal[3 * i]), T(pal[3 * i + 1]), T(pal[3 * i + 2]));
        v3<T> d = c - rgb;
        T dist = dot(d, d);
        if (dist < best) {
            best = dist;
            bc = c;
        }
    }
    T ig = T(1) / gamma;
    rgb = mk(m_pow(bc.x, ig), m_pow(bc.y, ig), m_pow(bc.z, ig));
    if (depth != 0) {
        T nl = (rgb.x + rgb.y + rgb.z) / T(3);
        if (nl != T(0)) rgb = rgb * (l / nl);
    }
    return rgb;
}

// EightBitSurface::apply_stylization (svm/surface/mod.rs:1315-1345)
template <typename T> NLPF_HD v3<T> map_8bit(v3<T> c) {
    v3<T> s = linear_to_srgb(c);
    return srgb_to_linear(mk(m_round(s.x * T(8)) / T(8), m_round(s.y * T(8)) / T(8),
                             m_round(s.z * T(4)) / T(4)));
}
template <typename T> NLPF_HD v3<T> eight_bit(v3<T> li, uint32_t depth, T iscale) {
    T l = hmax(li);
    if (l == T(0)) return li;
    if (depth > 0) return map_8bit(li / iscale) * iscale;
    T nl = m_min(l, T(1));
    return map_8bit(li / l * nl);
}

// HalfToneSurface::apply_stylization (svm/surface/mod.rs:1180-1233)
template <typename T> NLPF_HD v3<T> halftone(v3<T> li, uint32_t depth, const nlpf_style_node& nd, v3<T> p) {
    if (depth > nd.max_depth) return li;
    T l = (li.x + li.y + li.z) / T(3);
    if (l == T(0)) return li;
    const float* q = nd.p;
    T iscale = q[0], rscale = q[1], radius = q[2], d = q[3], blend = q[4];
    T r = depth == 0 ? l / rscale : m_round(l / iscale) / rscale;
    r = T(1) - m_clamp(r, T(0), T(1));
    r = r * d * T(0.5) * radius;
    const float* m = q + 11;
    v3<T> tp = mk(T(m[0]) * p.x + T(m[1]) * p.y + T(m[2]) * p.z + T(m[3]),
                  T(m[4]) * p.x + T(m[5]) * p.y + T(m[6]) * p.z + T(m[7]),
                  T(m[8]) * p.x + T(m[9]) * p.y + T(m[10]) * p.z + T(m[11]));
    v3<T> ip = mk(m_floor(tp.x / d) * d, m_floor(tp.y / d) * d, m_floor(tp.z / d) * d);
    T best = T(1e16);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                v3<T> g = ip + mk(T(i), T(j), T(k)) * d;
                v3<T> dd = g - p;
                T dist = m_sqrt(dot(dd, dd));
                if (dist < best) best = dist;
            }
    v3<T> tone = best < r ? mk(T(q[5]), T(q[6]), T(q[7])) : mk(T(q[8]), T(q[9]), T(q[10]));
    v3<T> c = tone * li;
    return li + (c - li) * blend;
}

NLPF_HD bool depth_active(uint32_t mask, uint32_t depth) {
    return (mask >> (depth < 31u ? depth : 31u)) & 1u;
}

// ramp_tab (optional): a copy of the node's 32 ramp parameters to read instead of nd.p (a
// kernel keeps it in LDS, where the stop lookups of the binary search are cheap)
template <typename T>
NLPF_HD v3<T> apply_node(const nlpf_style_node& nd, uint32_t depth, v3<T> li, v3<T> p,
                         const float* ramp_tab = nullptr) {
    if (!depth_active(nd.depth_mask, depth)) return li;
    switch (nd.type) {
        case NLPF_STY_TOON: return toon(li, T(nd.p[0]));
        case NLPF_STY_PALETTE: return palette(li, depth, T(nd.p[0]), T(nd.p[1]), nd.p + 2, nd.count);
        case NLPF_STY_SATURATION: return saturation(li, depth, T(nd.p[0]), T(nd.p[1]));
        case NLPF_STY_REX: return rex(li, T(nd.p[0]));
        case NLPF_STY_EIGHT_BIT: return eight_bit(li, depth, T(nd.p[0]));
        case NLPF_STY_HALFTONE: return halftone(li, depth, nd, p);
        case NLPF_STY_HUE_SHIFT: return hue_shift(li, T(nd.p[0]));
        case NLPF_STY_RAMP: {
#if defined(__HIP_DEVICE_COMPILE__)
            // the table lives in LDS: reading it through an LDS pointer keeps the lookups
            // ds_read (a pointer that may also point at the parameters would be a flat access)
            if (ramp_tab) {
                const auto q = (const __attribute__((address_space(3))) float*)ramp_tab;
                return ramp(li, q, q + 16, nd.count);
            }
#endif
            return ramp(li, nd.p, nd.p + 16, nd.count);
        }
        default: return li;
    }
}

#define NLPF_STYLE_STACK 16
// The wrapper nodes SurfaceClosure::apply_stylization (svm/surface/mod.rs:1824-1842) runs for
// one material and flags: descend from the root picking EnableIf branches by the flags' style
// index; seq[0..n) receives the wrapper transforms outermost first (they apply innermost first).
NLPF_HD int resolve_style(const nlpf_style_node* nodes, int root, uint32_t flags, int* seq) {
    if (flags & NLPF_FLAG_DISABLED) return 0;
    int sp = 0;
    int cur = root;
    const uint32_t sidx = style_index(flags);
    while (cur >= 0 && sp < NLPF_STYLE_STACK) {
        const nlpf_style_node& nd = nodes[cur];
        if (nd.type == NLPF_STY_ENABLE_IF) {
            cur = (sidx == nd.idx) ? nd.inner : nd.else_;
        } else if (nd.type == NLPF_STY_SET_IDX || nd.type == NLPF_STY_BASE) {
            cur = nd.inner;
        } else {
            seq[sp++] = cur;
            cur = nd.inner;
        }
    }
    return sp;
}
template <typename T>
NLPF_HD v3<T> apply_style(const nlpf_style_node* nodes, int root, uint32_t depth, uint32_t flags,
                          v3<T> li, v3<T> p) {
    int stack[NLPF_STYLE_STACK];
    int sp = resolve_style(nodes, root, flags, stack);
    while (sp > 0) li = apply_node<T>(nodes[stack[--sp]], depth, li, p);
    return li;
}

// A resolved chain held by value (kernel arguments): node[0] applies first.  Kernels whose
// vertices all share one material and flags read every parameter as a uniform scalar load.
#define NLPF_CHAIN_MAX 4
struct StyleChain {
    int32_t valid;  // 0 (zero-initialised): not resolved, walk the node table from the root
    int32_t n;      // nodes in use
    nlpf_style_node node[NLPF_CHAIN_MAX];
};
inline StyleChain make_chain(const nlpf_style_node* nodes, int root, uint32_t flags) {
    StyleChain c{};
    int seq[NLPF_STYLE_STACK];
    const int n = root < 0 ? 0 : resolve_style(nodes, root, flags, seq);
    if (n > NLPF_CHAIN_MAX) return c;
    c.valid = 1;
    c.n = n;
    for (int i = 0; i < n; ++i) c.node[i] = nodes[seq[n - 1 - i]];
    return c;
}
// tabs (optional): [NLPF_CHAIN_MAX][32] copies of the ramp nodes' parameters
template <typename T>
NLPF_HD v3<T> apply_chain(const StyleChain& c, uint32_t depth, v3<T> li, v3<T> p, const float* tabs = nullptr) {
    for (int i = 0; i < NLPF_CHAIN_MAX; ++i)
        if (i < c.n) li = apply_node<T>(c.node[i], depth, li, p, tabs ? tabs + 32 * i : nullptr);
    return li;
}

// update_stylization_flags: SurfaceClosure forwards to its root surface only
// (svm/surface/mod.rs:1806-1815); wrapper surfaces keep the trait default (return the flags
// unchanged, :266-274) and only a SetStyleIdx root rewrites the one-hot style bits (:941-953,
// set_style_index :94-101, which leaves stylization_disabled alone).
NLPF_HD uint32_t update_flags(const nlpf_style_node* nodes, int root, uint32_t flags) {
    if (root < 0 || nodes[root].type != NLPF_STY_SET_IDX) return flags;
    const uint32_t idx = nodes[root].idx;
    uint32_t f = flags & NLPF_FLAG_DISABLED;
    if (idx == 1) f |= NLPF_FLAG_STYLE2;
    if (idx == 2) f |= NLPF_FLAG_STYLE3;
    if (idx == 3) f |= NLPF_FLAG_STYLE4;
    return f;
}

}  // namespace nlpf
