/*
 * TEST INFRASTRUCTURE ONLY — CPU oracle for the stylization nodes (FP64).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * as the checker; the library never calls it.
 *
 * Restated from the reference sources:
 *   color helpers       crates/akari_render/src/color.rs:328-337 (clamp), 356-361 (lum),
 *                       531-549 (sRGB), 551-589 (HSV)
 *   ToonSurface         crates/akari_render/src/svm/surface/mod.rs:572-610
 *   ColorPaletteSurface mod.rs:696-735
 *   SaturationAdjust    mod.rs:820-851
 *   SaturationAdjustRex mod.rs:1059-1089
 *   HalfToneSurface     mod.rs:1180-1233
 *   EightBitSurface     mod.rs:1315-1345
 *   EnableOrSet         mod.rs:941-976, SurfaceClosure::apply_stylization mod.rs:1824-1842
 *   StylizationFlags    mod.rs:58-101
 * Hue shift and colour ramp have no reference implementation; they follow the definitions
 * written in include/nlpf.h.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include "oracle.h"

static double lum(const double c[3]) { return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]; }
static double s2l(double v) { return v <= 0.04045 ? v / 12.92 : pow((v + 0.055) / 1.055, 2.4); }
static double l2s(double v) { return v <= 0.0031308 ? v * 12.92 : pow(v, 1.0 / 2.4) * 1.055 - 0.055; }
static double dmin(double a, double b) { return a < b ? a : b; }
static double dmax(double a, double b) { return a > b ? a : b; }
static double clampd(double v, double lo, double hi) { return dmin(dmax(v, lo), hi); }
static double remeu(double a, double m) {
    double r = fmod(a, m);
    return r < 0 ? r + m : r;
}

void or_rgb_to_hsv(const double c[3], double hsv[3]) {
    double mx = dmax(dmax(c[0], c[1]), c[2]);
    double mn = dmin(dmin(c[0], c[1]), c[2]);
    double diff = mx - mn, h;
    if (mx == mn)
        h = 0;
    else if (mx == c[0])
        h = remeu((c[1] - c[2]) / diff, 6.0);
    else if (mx == c[1])
        h = (c[2] - c[0]) / diff + 2.0;
    else
        h = (c[0] - c[1]) / diff + 4.0;
    hsv[0] = h / 6.0;
    hsv[1] = mx == 0 ? 0 : diff / mx;
    hsv[2] = mx;
}

void or_hsv_to_rgb(const double hsv[3], double c[3]) {
    static const double ns[3] = {5.0, 3.0, 1.0};
    for (int i = 0; i < 3; ++i) {
        double k = remeu(ns[i] + hsv[0] * 6.0, 6.0);
        double t = dmax(dmin(dmin(k, 4.0 - k), 1.0), 0.0);
        c[i] = hsv[2] - hsv[2] * hsv[1] * t;
    }
}

static void toon(double c[3], double w) {
    for (int i = 0; i < 3; ++i) c[i] = clampd(c[i], 0.0, 100000.0);
    double l = lum(c);
    if (l <= 0) return;
    double nl;
    if (l > 1.0)
        nl = floor(l / w) * w;
    else
        nl = s2l(floor(l2s(l) / w) * w);
    for (int i = 0; i < 3; ++i) c[i] = c[i] / l * nl;
}

static double ramp_at(const float* pos, const float* val, uint32_t n, double x) {
    if (n == 0) return x;
    uint32_t k = 0;
    while (k < n && (double)pos[k] <= x) ++k;
    if (k == 0) return val[0];
    if (k == n) return val[n - 1];
    double t = (x - pos[k - 1]) / ((double)pos[k] - pos[k - 1]);
    return val[k - 1] + t * ((double)val[k] - val[k - 1]);
}

static void ramp(double c[3], const float* pos, const float* val, uint32_t n) {
    for (int i = 0; i < 3; ++i) c[i] = clampd(c[i], 0.0, 100000.0);
    double l = lum(c);
    if (l <= 0) return;
    double r = ramp_at(pos, val, n, l);
    for (int i = 0; i < 3; ++i) c[i] = c[i] / l * r;
}

static void hue_shift(double c[3], double shift) {
    double hsv[3];
    or_rgb_to_hsv(c, hsv);
    hsv[0] = remeu(hsv[0] + shift, 1.0);
    or_hsv_to_rgb(hsv, c);
}

static void saturation(double c[3], uint32_t depth, double power, double iscale) {
    if (dmax(dmax(c[0], c[1]), c[2]) == 0) return;
    double r[3], hsv[3];
    for (int i = 0; i < 3; ++i) {
        r[i] = depth != 0 ? c[i] / iscale : c[i];
        r[i] = l2s(clampd(r[i], 0, 1));
    }
    or_rgb_to_hsv(r, hsv);
    hsv[1] = pow(hsv[1], power);
    or_hsv_to_rgb(hsv, r);
    for (int i = 0; i < 3; ++i) c[i] = s2l(r[i]) * (depth != 0 ? iscale : 1.0);
}

static void rex(double c[3], double blend) {
    double l = (c[0] + c[1] + c[2]) / 3.0;
    double mn = dmin(dmin(c[0], c[1]), c[2]), mx = dmax(dmax(c[0], c[1]), c[2]);
    if (l <= 0 || mn < 0) return;
    double im[3], u = 0;
    for (int i = 0; i < 3; ++i) {
        im[i] = (c[i] - mn) / mx;
        u += im[i];
    }
    u /= 3.0;
    for (int i = 0; i < 3; ++i) c[i] = blend * l + (1.0 - blend) * (u != 0 ? im[i] * l / u : l);
}

static void palette(double c[3], uint32_t depth, double iscale, double gamma, const float* pal, uint32_t n) {
    double l = (c[0] + c[1] + c[2]) / 3.0;
    double r[3];
    for (int i = 0; i < 3; ++i) r[i] = pow(clampd(depth != 0 ? c[i] / iscale : c[i], 0, 1), gamma);
    double best = 1e16, bc[3] = {0, 0, 0};
    for (uint32_t k = 0; k < n; ++k) {
        double d = 0;
        for (int i = 0; i < 3; ++i) d += ((double)pal[3 * k + i] - r[i]) * ((double)pal[3 * k + i] - r[i]);
        if (d < best) {
            best = d;
            for (int i = 0; i < 3; ++i) bc[i] = pal[3 * k + i];
        }
    }
    for (int i = 0; i < 3; ++i) c[i] = pow(bc[i], 1.0 / gamma);
    if (depth != 0) {
        double nl = (c[0] + c[1] + c[2]) / 3.0;
        if (nl != 0)
            for (int i = 0; i < 3; ++i) c[i] *= l / nl;
    }
}

static void map8(double c[3]) {
    static const double q[3] = {8.0, 8.0, 4.0};
    for (int i = 0; i < 3; ++i) c[i] = s2l(round(l2s(c[i]) * q[i]) / q[i]);
}

static void eight_bit(double c[3], uint32_t depth, double iscale) {
    double l = dmax(dmax(c[0], c[1]), c[2]);
    if (l == 0) return;
    if (depth > 0) {
        for (int i = 0; i < 3; ++i) c[i] /= iscale;
        map8(c);
        for (int i = 0; i < 3; ++i) c[i] *= iscale;
    } else {
        double nl = dmin(l, 1.0);
        for (int i = 0; i < 3; ++i) c[i] = c[i] / l * nl;
        map8(c);
    }
}

static void halftone(double c[3], uint32_t depth, const or_node* nd, const double p[3]) {
    if (depth > nd->max_depth) return;
    double l = (c[0] + c[1] + c[2]) / 3.0;
    if (l == 0) return;
    const float* q = nd->p;
    double r = depth == 0 ? l / q[1] : round(l / q[0]) / q[1];
    r = (1.0 - clampd(r, 0, 1)) * q[3] * 0.5 * q[2];
    const float* m = q + 11;
    double ip[3];
    for (int i = 0; i < 3; ++i) {
        double t = m[4 * i] * p[0] + m[4 * i + 1] * p[1] + m[4 * i + 2] * p[2] + m[4 * i + 3];
        ip[i] = floor(t / q[3]) * q[3];
    }
    double best = 1e16;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                double g[3] = {ip[0] + i * q[3], ip[1] + j * q[3], ip[2] + k * q[3]};
                double dist = sqrt((g[0] - p[0]) * (g[0] - p[0]) + (g[1] - p[1]) * (g[1] - p[1]) +
                                   (g[2] - p[2]) * (g[2] - p[2]));
                if (dist < best) best = dist;
            }
    const float* tone = best < r ? q + 5 : q + 8;
    for (int i = 0; i < 3; ++i) {
        double cc = tone[i] * c[i];
        c[i] = c[i] + (cc - c[i]) * q[4];
    }
}

static uint32_t style_idx(uint32_t flags) {
    uint32_t i = 0;
    if (flags & 2u) i = 1;
    if (flags & 4u) i = 2;
    if (flags & 8u) i = 3;
    return i;
}

/* recursive form of the reference's wrapper recursion: inner first, then own transform */
static void apply_rec(const or_node* nodes, int cur, uint32_t depth, uint32_t flags, double c[3], const double p[3],
                      int guard) {
    if (cur < 0 || guard > 32) return;
    const or_node* nd = &nodes[cur];
    if (nd->type == OR_ENABLE_IF) {
        apply_rec(nodes, style_idx(flags) == nd->idx ? nd->inner : nd->else_, depth, flags, c, p, guard + 1);
        return;
    }
    apply_rec(nodes, nd->inner, depth, flags, c, p, guard + 1);
    if (nd->type == OR_SET_IDX || nd->type == OR_BASE) return;
    uint32_t bit = depth < 31 ? depth : 31;
    if (!((nd->depth_mask >> bit) & 1u)) return;
    switch (nd->type) {
        case OR_TOON: toon(c, nd->p[0]); break;
        case OR_PALETTE: palette(c, depth, nd->p[0], nd->p[1], nd->p + 2, nd->count); break;
        case OR_SATURATION: saturation(c, depth, nd->p[0], nd->p[1]); break;
        case OR_REX: rex(c, nd->p[0]); break;
        case OR_EIGHT_BIT: eight_bit(c, depth, nd->p[0]); break;
        case OR_HALFTONE: halftone(c, depth, nd, p); break;
        case OR_HUE_SHIFT: hue_shift(c, nd->p[0]); break;
        case OR_RAMP: ramp(c, nd->p, nd->p + 16, nd->count); break;
        default: break;
    }
}

void or_apply_style(const or_node* nodes, int root, uint32_t depth, uint32_t flags, const double in[3],
                    const double p[3], double out[3]) {
    double c[3] = {in[0], in[1], in[2]};
    static const double zero[3] = {0, 0, 0};
    if (!(flags & 1u)) apply_rec(nodes, root, depth, flags, c, p ? p : zero, 0);
    memcpy(out, c, sizeof(c));
}

uint32_t or_update_flags(const or_node* nodes, int root, uint32_t flags) {
    if (root < 0 || nodes[root].type != OR_SET_IDX) return flags;
    uint32_t f = flags & 1u;
    uint32_t idx = nodes[root].idx;
    if (idx == 1) f |= 2u;
    if (idx == 2) f |= 4u;
    if (idx == 3) f |= 8u;
    return f;
}

int or_stylize_batch(const or_node* nodes, const int32_t* roots, uint32_t n_materials, const double* in,
                     double* out, uint64_t n, const uint32_t* depth, uint32_t depth0, const uint8_t* flags,
                     const uint16_t* material, const double* p) {
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t m = material ? material[i] : 0u;
        int root = m < n_materials ? roots[m] : -1;
        or_apply_style(nodes, root, depth ? depth[i] : depth0, flags ? flags[i] : 0u, in + 3 * i,
                       p ? p + 3 * i : NULL, out + 3 * i);
    }
    return 0;
}
