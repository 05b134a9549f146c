// Procedural scene used to make synthetic path-vertex records: an open-front box
// [-0.5,0.5]^3 of 5 diffuse walls, two diffuse spheres (r = 0.2, 0.3) resting on the floor
// and a downward-facing quad area light under the ceiling.  Rays leaving through the open
// front reach a black "sky" sphere of radius 100.
#pragma once
#include "style.hpp"

namespace nlpf {

// PCG-XSH-RR 64/32, the generator of crates/akari_render/src/sampler/mod.rs:79-137
struct Pcg32 {
    uint64_t state, inc;
};
NLPF_HD uint32_t pcg_next(Pcg32& g) {
    const uint64_t old = g.state;
    g.state = old * 0x5851f42d4c957f2dull + g.inc;
    const uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
    const uint32_t rot = (uint32_t)(old >> 59);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
}
// Pcg32Var::new_seq_offset (sampler/mod.rs:95-101)
NLPF_HD Pcg32 pcg_seq_offset(uint64_t seq, uint64_t offset) {
    Pcg32 g{0ull, (seq << 1) | 1ull};
    pcg_next(g);
    g.state += offset;
    pcg_next(g);
    return g;
}
// ChaCha block function (Bernstein; RFC 7539 2.3 layout): `rounds` rounds over 16 words, then
// the input added back
NLPF_HD uint32_t rotl32(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }
NLPF_HD void chacha_qr(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
    a += b; d = rotl32(d ^ a, 16);
    c += d; b = rotl32(b ^ c, 12);
    a += b; d = rotl32(d ^ a, 8);
    c += d; b = rotl32(b ^ c, 7);
}
NLPF_HD void chacha_block(const uint32_t in[16], int rounds, uint32_t out[16]) {
    uint32_t x[16];
    for (int i = 0; i < 16; ++i) x[i] = in[i];
    for (int r = 0; r < rounds; r += 2) {
        chacha_qr(x[0], x[4], x[8], x[12]);
        chacha_qr(x[1], x[5], x[9], x[13]);
        chacha_qr(x[2], x[6], x[10], x[14]);
        chacha_qr(x[3], x[7], x[11], x[15]);
        chacha_qr(x[0], x[5], x[10], x[15]);
        chacha_qr(x[1], x[6], x[11], x[12]);
        chacha_qr(x[2], x[7], x[8], x[13]);
        chacha_qr(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}
// 32-byte key of rand 0.8.5's StdRng::seed_from_u64 (rand_core 0.6.4: eight outputs of a
// PCG32 that advances before each output, little-endian)
struct StdKey {
    uint32_t w[8];
};
NLPF_HD StdKey std_seed_from_u64(uint64_t state) {
    StdKey k;
    for (int i = 0; i < 8; ++i) {
        state = state * 6364136223846793005ull + 11634580027462260723ull;
        const uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
        const uint32_t rot = (uint32_t)(state >> 59);
        k.w[i] = (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    return k;
}
// t-th next_u64 of StdRng (rand_chacha 0.3.1 ChaCha12Rng, stream 0): keystream words 2t and
// 2t+1 of block t/8, the block counter in state words 12-13
NLPF_HD uint64_t std_rng_u64(const StdKey& k, uint64_t t) {
    uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
    for (int i = 0; i < 8; ++i) in[4 + i] = k.w[i];
    const uint64_t b = t >> 3;
    in[12] = (uint32_t)b;
    in[13] = (uint32_t)(b >> 32);
    in[14] = in[15] = 0u;
    uint32_t out[16];
    chacha_block(in, 12, out);
    const int w = (int)(t & 7) * 2;
    return (uint64_t)out[w] | ((uint64_t)out[w + 1] << 32);
}
// reservoir stream t as the reference seeds it (init_pcg32_buffer_with_seed,
// sampler/mod.rs:157-174): new_seq_offset(t, t-th u64 of StdRng::seed_from_u64(seed))
NLPF_HD Pcg32 reservoir_stream(const StdKey& k, uint64_t t) { return pcg_seq_offset(t, std_rng_u64(k, t)); }
NLPF_HD uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
// uniform in [0,1): 24 high bits
NLPF_HD float pcg_uniform(Pcg32& g) { return (float)(pcg_next(g) >> 8) * (1.0f / 16777216.0f); }
// uniform in (0,1]
NLPF_HD float pcg_uniform_open0(Pcg32& g) { return (float)((pcg_next(g) >> 8) + 1u) * (1.0f / 16777216.0f); }

#define NLPF_SCENE_SKY 100.0f
#define NLPF_MAT_LIGHT 7
#define NLPF_MAT_SKY 8
#define NLPF_SCENE_MATERIALS 8
// radiance of the quad light's lower face (emission colour 1, strength 15)
#define NLPF_LIGHT_EMISSION 15.0f

// diffuse reflectance of material m: three U[lo, hi) draws from the stream of the material,
// keyed by the scene seed; the light is a pure emitter (reflectance 0)
NLPF_HD v3<float> material_albedo(int m, uint64_t key, float lo, float hi) {
    if (m == NLPF_MAT_LIGHT) return v3<float>{0.0f, 0.0f, 0.0f};
    Pcg32 r = pcg_seq_offset(0xA1BED0ull + (uint64_t)m, key);
    const float a = lo + (hi - lo) * pcg_uniform(r);
    const float b = lo + (hi - lo) * pcg_uniform(r);
    const float c = lo + (hi - lo) * pcg_uniform(r);
    return v3<float>{a, b, c};
}

struct Hit {
    float t;
    v3<float> p, n;  // n: geometric normal, outward for spheres, inward for walls
    int mat;         // 0 floor 1 ceiling 2 back 3 left 4 right 5 sphere0 6 sphere1 7 light 8 sky
};

struct SceneConst {
    static constexpr float h = 0.5f;
    static constexpr float s0x = -0.22f, s0y = -0.3f, s0z = -0.18f, s0r = 0.2f;
    static constexpr float s1x = 0.17f, s1y = -0.2f, s1z = 0.08f, s1r = 0.3f;
    static constexpr float ly = 0.499f, lh = 0.12f;  // light plane y, half extent in x and z
    static constexpr float cam_z = 1.6f, tan_half_fov = 0.41421356f;  // 45 degree vertical fov
};

NLPF_HD void hit_plane(v3<float> o, v3<float> d, int axis, float c, float nsign, int mat, float lo0, float hi0,
                       float lo1, float hi1, Hit& best) {
    const float oc = axis == 0 ? o.x : (axis == 1 ? o.y : o.z);
    const float dc = axis == 0 ? d.x : (axis == 1 ? d.y : d.z);
    if (dc == 0.0f) return;
    const float t = (c - oc) / dc;
    if (!(t > 1e-5f) || !(t < best.t)) return;
    const v3<float> p = o + d * t;
    const float a = axis == 0 ? p.y : p.x;
    const float b = axis == 2 ? p.y : p.z;
    if (a < lo0 || a > hi0 || b < lo1 || b > hi1) return;
    best.t = t;
    best.p = p;
    best.n = v3<float>{axis == 0 ? nsign : 0.0f, axis == 1 ? nsign : 0.0f, axis == 2 ? nsign : 0.0f};
    best.mat = mat;
}

NLPF_HD void hit_sphere(v3<float> o, v3<float> d, v3<float> c, float r, int mat, Hit& best) {
    const v3<float> oc = o - c;
    const float b = dot(oc, d);
    const float cc = dot(oc, oc) - r * r;
    const float disc = b * b - cc;
    if (disc < 0.0f) return;
    const float sq = m_sqrt(disc);
    float t = -b - sq;
    if (!(t > 1e-5f)) t = -b + sq;
    if (!(t > 1e-5f) || !(t < best.t)) return;
    best.t = t;
    best.p = o + d * t;
    best.n = (best.p - c) / r;
    best.mat = mat;
}

// nearest hit along unit direction d; misses report the sky sphere
NLPF_HD Hit scene_intersect(v3<float> o, v3<float> d) {
    typedef SceneConst C;
    Hit h;
    h.t = 1e30f;
    h.mat = -1;
    hit_plane(o, d, 1, -C::h, 1.0f, 0, -C::h, C::h, -C::h, C::h, h);    // floor
    hit_plane(o, d, 1, C::h, -1.0f, 1, -C::h, C::h, -C::h, C::h, h);    // ceiling
    hit_plane(o, d, 2, -C::h, 1.0f, 2, -C::h, C::h, -C::h, C::h, h);    // back
    hit_plane(o, d, 0, -C::h, 1.0f, 3, -C::h, C::h, -C::h, C::h, h);    // left
    hit_plane(o, d, 0, C::h, -1.0f, 4, -C::h, C::h, -C::h, C::h, h);    // right
    hit_plane(o, d, 1, C::ly, -1.0f, NLPF_MAT_LIGHT, -C::lh, C::lh, -C::lh, C::lh, h);  // light
    hit_sphere(o, d, v3<float>{C::s0x, C::s0y, C::s0z}, C::s0r, 5, h);
    hit_sphere(o, d, v3<float>{C::s1x, C::s1y, C::s1z}, C::s1r, 6, h);
    if (h.mat < 0) {
        h.t = NLPF_SCENE_SKY;
        h.p = o + d * NLPF_SCENE_SKY;
        h.n = v3<float>{-d.x, -d.y, -d.z};
        h.mat = NLPF_MAT_SKY;
    }
    return h;
}

NLPF_HD v3<float> normalize3(v3<float> v) { return v / m_sqrt(dot(v, v)); }

// camera ray through pixel (x + jx, y + jy), y = 0 at the top of the image
NLPF_HD void camera_ray(uint32_t W, uint32_t H, float x, float y, v3<float>& o, v3<float>& d) {
    typedef SceneConst C;
    const float aspect = (float)W / (float)H;
    const float u = (2.0f * x / (float)W - 1.0f) * C::tan_half_fov * aspect;
    const float v = (1.0f - 2.0f * y / (float)H) * C::tan_half_fov;
    o = v3<float>{0.0f, 0.0f, C::cam_z};
    d = normalize3(v3<float>{u, v, -1.0f});
}

// orthonormal basis around n (Duff et al. 2017, "Building an Orthonormal Basis, Revisited")
NLPF_HD void onb(v3<float> n, v3<float>& t, v3<float>& b) {
    const float sgn = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sgn + n.z);
    const float bb = n.x * n.y * a;
    t = v3<float>{1.0f + sgn * n.x * n.x * a, sgn * bb, -sgn * n.x};
    b = v3<float>{bb, sgn + n.y * n.y * a, -n.y};
}

// cosine-weighted hemisphere sample around n (disk mapping of sampling.rs:5-21)
NLPF_HD v3<float> cosine_sample(v3<float> n, float u1, float u2) {
    const float r = m_sqrt(u1);
    const float phi = u2 * 6.2831853f;
    const float lx = r * cosf(phi), ly = r * sinf(phi);
    const float lz = m_sqrt(m_max(0.0f, 1.0f - lx * lx - ly * ly));
    v3<float> t, b;
    onb(n, t, b);
    return t * lx + b * ly + n * lz;
}

}  // namespace nlpf
