/*
 * TEST INFRASTRUCTURE ONLY — world-mode NL-PF oracle (CPU, FP64).  Loaded only by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
 *
 * Restated from crates/akari_integrator/src/lm_path_filtering.rs:
 *   hash / get_cell / for_each_neighbor      :20-123 (27 cells, dz-dy-dx order, every bucket
 *                                                     walked, bucket collisions revisit entries)
 *   mis_weight (balance heuristic)            :305-316
 *   evaluate_filtered_radiance                :757-972 (reservoir :813-821, direct :839-863,
 *                                                     indirect + SMIS :864-945, stylize :948-971)
 *   kernel_filter / kernel_propagate          :999-1060, :973-998
 *   depth loop, radius / resolution           :1563-1598, :1673-1774
 *   kernel_splat / kernel_update_film         :479-506, :428-477 (+ film.rs:196-229)
 * with each material's BSDF behind SurfaceClosure::evaluate (svm/surface/mod.rs:1701-1744;
 * DiffuseBsdf, or the principled subset of oracle_bsdf.c) and Pcg32 (sampler/mod.rs:79-137).
 * The SMIS denominators evaluate each reused neighbour's own BSDF, normals and wo (:886-935).
 *
 * Reservoir streams as the reference seeds them (init_pcg32_buffer_with_seed,
 * sampler/mod.rs:157-174): the t-th point of a batch draws from Pcg32 new_seq_offset(t, s_t),
 * where s_t is the t-th u64 of rand 0.8.5's StdRng::seed_from_u64(seed) (rand_chacha 0.3.1
 * ChaCha12 keyed by rand_core 0.6.4's PCG32 seed expansion; those crates are not vendored, so
 * their published algorithms are restated below), carried across depths.
 *
 * Determinism choice shared with the library (the reference's order is scheduling-dependent):
 * grid entries of one bucket are visited in ascending path id.
 *
 * Clustering (cluster = true) is the library's deterministic replacement of the reference's
 * racy CAS seeding / joining (:532-623, 1675-1710), restated sequentially:
 *   repeat: every bucket's smallest unclustered path id becomes a seed;
 *     repeat: every unclustered vertex proposes to the first seed (27-cell visit order) within
 *       the radius whose cluster has fewer than K joiners; each seed accepts its proposals in
 *       ascending vertex order until it has K joiners;  until no vertex proposes;
 *   until every vertex is clustered.
 * The filter then gathers around the seed's position and keeps the members of the vertex's
 * own cluster (:782-790, 807-811).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define NP 28

typedef struct {
    uint64_t state, inc;
} pcg_t;
static uint32_t pcg_next(pcg_t* g) {
    uint64_t old = g->state;
    g->state = old * 0x5851f42d4c957f2dull + g->inc;
    uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
    uint32_t rot = (uint32_t)(old >> 59);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
}
static pcg_t pcg_seq(uint64_t seq, uint64_t off) {
    pcg_t g = {0, (seq << 1) | 1};
    pcg_next(&g);
    g.state += off;
    pcg_next(&g);
    return g;
}

/* ChaCha block function (Bernstein; RFC 7539 2.3 for the layout of constants, key, counter and
 * nonce): `rounds` rounds (rand_chacha's ChaCha12 uses 12) over the 16 input words, then the
 * input added back. */
#define OR_ROTL(v, c) (((v) << (c)) | ((v) >> (32 - (c))))
#define OR_QR(a, b, c, d)                                                                          \
    do {                                                                                           \
        a += b; d ^= a; d = OR_ROTL(d, 16);                                                        \
        c += d; b ^= c; b = OR_ROTL(b, 12);                                                        \
        a += b; d ^= a; d = OR_ROTL(d, 8);                                                         \
        c += d; b ^= c; b = OR_ROTL(b, 7);                                                         \
    } while (0)
void or_chacha_block(const uint32_t in[16], int rounds, uint32_t out[16]) {
    uint32_t x[16];
    memcpy(x, in, sizeof(x));
    for (int r = 0; r < rounds; r += 2) {
        OR_QR(x[0], x[4], x[8], x[12]);
        OR_QR(x[1], x[5], x[9], x[13]);
        OR_QR(x[2], x[6], x[10], x[14]);
        OR_QR(x[3], x[7], x[11], x[15]);
        OR_QR(x[0], x[5], x[10], x[15]);
        OR_QR(x[1], x[6], x[11], x[12]);
        OR_QR(x[2], x[7], x[8], x[13]);
        OR_QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}

/* rand_core 0.6.4 SeedableRng::seed_from_u64: the 32-byte seed is eight outputs of a PCG32
 * (XSH-RR) that advances its state before each output, stored little-endian. */
void or_std_seed_from_u64(uint64_t state, uint32_t key[8]) {
    for (int i = 0; i < 8; ++i) {
        state = state * 6364136223846793005ull + 11634580027462260723ull;
        const uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
        const uint32_t rot = (uint32_t)(state >> 59);
        key[i] = (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
}

/* rand_chacha 0.3.1 ChaCha12Rng from a 32-byte key (stream 0): out[t] = the t-th next_u64, i.e.
 * keystream words 2t (low half) and 2t+1; block b holds words 16b..16b+15 at 64-bit counter b
 * (state words 12-13), stream words 14-15 zero. */
void or_chacha12_u64(const uint32_t key[8], uint64_t first, uint64_t n, uint64_t* out) {
    uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
    uint32_t blk[16];
    uint64_t cur = ~0ull;
    memcpy(in + 4, key, 32);
    for (uint64_t t = first; t < first + n; ++t) {
        const uint64_t b = t >> 3;
        if (b != cur) {
            in[12] = (uint32_t)b;
            in[13] = (uint32_t)(b >> 32);
            in[14] = in[15] = 0;
            or_chacha_block(in, 12, blk);
            cur = b;
        }
        const int w = (int)(t & 7) * 2;
        out[t - first] = (uint64_t)blk[w] | ((uint64_t)blk[w + 1] << 32);
    }
}

/* the reference's reservoir streams, init_pcg32_buffer_with_seed(count, seed): stream t is
 * Pcg32 new_seq_offset(t, t-th u64 of StdRng::seed_from_u64(seed)) */
static int std_pcg_streams(uint64_t seed, uint64_t count, pcg_t* out) {
    uint32_t key[8];
    uint64_t* s = (uint64_t*)malloc(sizeof(uint64_t) * (count ? count : 1));
    if (!s) return -1;
    or_std_seed_from_u64(seed, key);
    or_chacha12_u64(key, 0, count, s);
    for (uint64_t t = 0; t < count; ++t) out[t] = pcg_seq(t, s[t]);
    free(s);
    return 0;
}

/* n draws of Pcg32 new_seq_offset(seq, offset).gen_u32() (sampler/mod.rs:91-117) */
int or_pcg32_fill(uint64_t seq, uint64_t offset, uint32_t n, uint32_t* out) {
    pcg_t g = pcg_seq(seq, offset);
    for (uint32_t i = 0; i < n; ++i) out[i] = pcg_next(&g);
    return 0;
}

/* the first gen_u32 draws of the reference's reservoir streams first..first+count-1 */
int or_reservoir_streams(uint64_t seed, uint64_t first, uint64_t count, uint32_t draws, uint32_t* out) {
    pcg_t* g = (pcg_t*)malloc(sizeof(pcg_t) * (first + count ? first + count : 1));
    if (!g) return -1;
    if (std_pcg_streams(seed, first + count, g)) {
        free(g);
        return -1;
    }
    for (uint64_t t = 0; t < count; ++t)
        for (uint32_t k = 0; k < draws; ++k) out[t * draws + k] = pcg_next(&g[first + t]);
    free(g);
    return 0;
}

typedef struct {
    const float* rec;
    uint64_t N;
} lvl_t;
static double F(lvl_t l, int c, uint64_t i) { return l.rec[(uint64_t)c * l.N + i]; }

/* Hemisphere and leak tests are decided in FP32, as the reference computes them (its
 * closures run in f32), inside or_bsdf_eval; the estimator arithmetic itself stays FP64. */
/* BSDF of material m: its table entry, diffuse without one */
static const or_bsdf* bsdf_of(const or_bsdf* bsdfs, uint32_t n_bsdfs, uint16_t m) {
    return bsdfs && m < n_bsdfs ? &bsdfs[m] : NULL;
}

static uint32_t bucket_of(const float pmin[3], const float pmax[3], const uint32_t res[3], uint64_t table,
                          const int32_t cell[3]) {
    (void)pmin;
    (void)pmax;
    (void)res;
    uint64_t h = ((uint64_t)(uint32_t)cell[0] * 73856093ull) ^ ((uint64_t)(uint32_t)cell[1] * 19349663ull) ^
                 ((uint64_t)(uint32_t)cell[2] * 83492791ull);
    return (uint32_t)(h % table);
}
static void cell_of(const float pmin[3], const float pmax[3], const uint32_t res[3], float px, float py, float pz,
                    int32_t c[3]) {
    float p[3] = {px, py, pz};
    for (int i = 0; i < 3; ++i) {
        float t = (p[i] - pmin[i]) / (pmax[i] - pmin[i]);
        t = (t < 0 ? 0 : (t > 1 ? 1 : t)) * (float)res[i];
        uint32_t u = (uint32_t)t;
        c[i] = (int32_t)(u < res[i] - 1 ? u : res[i] - 1);
    }
}

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y;
}

/* Neighbour multiset of every point of one level: for tests of the grid itself.
 * out_count[t] = number of entries visited within the radius (duplicates included). */
int or_world_neighbor_counts(const float* pos, uint32_t n, const float pmin[3], const float pmax[3], float radius,
                             uint64_t table, uint32_t* out_count) {
    uint32_t res[3];
    for (int i = 0; i < 3; ++i) res[i] = (uint32_t)floorf((pmax[i] - pmin[i]) / radius);
    uint64_t* kv = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
    if (!kv) return -1;
    for (uint32_t t = 0; t < n; ++t) {
        int32_t c[3];
        cell_of(pmin, pmax, res, pos[3 * t], pos[3 * t + 1], pos[3 * t + 2], c);
        kv[t] = ((uint64_t)bucket_of(pmin, pmax, res, table, c) << 32) | t;
    }
    qsort(kv, n, sizeof(uint64_t), cmp_u64);
    for (uint32_t t = 0; t < n; ++t) {
        int32_t c[3];
        cell_of(pmin, pmax, res, pos[3 * t], pos[3 * t + 1], pos[3 * t + 2], c);
        uint32_t cnt = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int32_t nc[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
                    if (nc[0] < 0 || nc[1] < 0 || nc[2] < 0 || nc[0] >= (int)res[0] || nc[1] >= (int)res[1] ||
                        nc[2] >= (int)res[2])
                        continue;
                    uint64_t b = bucket_of(pmin, pmax, res, table, nc);
                    /* lower bound of bucket b */
                    uint32_t lo = 0, hi = n;
                    while (lo < hi) {
                        uint32_t mid = (lo + hi) / 2;
                        if ((kv[mid] >> 32) < b) lo = mid + 1; else hi = mid;
                    }
                    for (uint32_t e = lo; e < n && (kv[e] >> 32) == b; ++e) {
                        uint32_t j = (uint32_t)kv[e];
                        float ddx = pos[3 * j] - pos[3 * t], ddy = pos[3 * j + 1] - pos[3 * t + 1],
                              ddz = pos[3 * j + 2] - pos[3 * t + 2];
                        if (sqrtf(ddx * ddx + ddy * ddy + ddz * ddz) <= radius) ++cnt;
                    }
                }
        out_count[t] = cnt;
    }
    free(kv);
    return 0;
}

/* first entry e of bucket b in the sorted (bucket << 32 | path) keys */
static uint32_t bucket_lo(const uint64_t* kv, uint32_t n, uint64_t b) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if ((kv[mid] >> 32) < b) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* deterministic clustering of one level (see the header comment); cid[i] = seed of i */
static void cluster_level(lvl_t lv, const float pmin[3], const float pmax[3], const uint32_t res[3], uint64_t table,
                          float radius, uint32_t K, const uint32_t* pts, uint32_t n, const uint64_t* kv,
                          uint32_t* cid, uint32_t* csize, uint64_t* prop) {
    const uint32_t UNSET = 0xFFFFFFFFu;
    for (uint32_t e = 0; e < n; ++e) {
        cid[pts[e]] = UNSET;
        csize[pts[e]] = 0;
    }
    for (;;) {
        /* seeds: the smallest unclustered path id of each bucket (its entries ascend by path id;
           a bucket can hold colliding cells, as on the GPU) */
        uint32_t nseed = 0;
        for (uint32_t e = 0; e < n; ++e) {
            const uint32_t i = pts[e];
            if (cid[i] != UNSET) continue;
            int32_t c[3];
            cell_of(pmin, pmax, res, (float)F(lv, 0, i), (float)F(lv, 1, i), (float)F(lv, 2, i), c);
            const uint64_t b = bucket_of(pmin, pmax, res, table, c);
            uint32_t smallest = UNSET;
            for (uint32_t f = bucket_lo(kv, n, b); f < n && (kv[f] >> 32) == b; ++f) {
                const uint32_t j = (uint32_t)kv[f];
                if (cid[j] == UNSET) {
                    smallest = j;
                    break;
                }
            }
            if (smallest == i) prop[nseed++] = i;
        }
        for (uint32_t s = 0; s < nseed; ++s) cid[prop[s]] = (uint32_t)prop[s];
        uint32_t n_uncl = 0;
        for (;;) {
            uint32_t np = 0;
            n_uncl = 0;
            for (uint32_t e = 0; e < n; ++e) {
                const uint32_t i = pts[e];
                if (cid[i] != UNSET) continue;
                ++n_uncl;
                const float px = (float)F(lv, 0, i), py = (float)F(lv, 1, i), pz = (float)F(lv, 2, i);
                int32_t c[3];
                cell_of(pmin, pmax, res, px, py, pz, c);
                int done = 0;
                for (int dz = -1; dz <= 1 && !done; ++dz)
                    for (int dy = -1; dy <= 1 && !done; ++dy)
                        for (int dx = -1; dx <= 1 && !done; ++dx) {
                            int32_t nc[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
                            if (nc[0] < 0 || nc[1] < 0 || nc[2] < 0 || nc[0] >= (int)res[0] ||
                                nc[1] >= (int)res[1] || nc[2] >= (int)res[2])
                                continue;
                            const uint64_t b = bucket_of(pmin, pmax, res, table, nc);
                            for (uint32_t f = bucket_lo(kv, n, b); f < n && (kv[f] >> 32) == b; ++f) {
                                const uint32_t j = (uint32_t)kv[f];
                                const float ddx = (float)F(lv, 0, j) - px, ddy = (float)F(lv, 1, j) - py,
                                            ddz = (float)F(lv, 2, j) - pz;
                                if (!(sqrtf(ddx * ddx + ddy * ddy + ddz * ddz) <= radius)) continue;
                                if (cid[j] == j && csize[j] < K) {
                                    prop[np++] = ((uint64_t)j << 32) | i;
                                    done = 1;
                                    break;
                                }
                            }
                        }
            }
            if (np == 0) break;
            qsort(prop, np, sizeof(uint64_t), cmp_u64);
            for (uint32_t a = 0; a < np;) {
                const uint32_t seed = (uint32_t)(prop[a] >> 32);
                uint32_t z = a;
                while (z < np && (uint32_t)(prop[z] >> 32) == seed) ++z;
                const uint32_t size0 = csize[seed], room = K > size0 ? K - size0 : 0u;
                for (uint32_t r = 0; r < z - a && r < room; ++r) cid[(uint32_t)prop[a + r]] = seed;
                csize[seed] = size0 + ((z - a) < room ? (z - a) : room);
                a = z;
            }
        }
        if (n_uncl == 0) break;
    }
}

/*
 * Whole world-mode pass.  Arguments mirror nlpf_world_run; cfg_* carry the "pf" config.
 * film [H][W][3] (double), radiance_out optional [L][N][3] (double).
 */
int or_world_run(uint32_t W, uint32_t H, uint32_t branch, uint32_t L, const float aabb_min[3],
                 const float aabb_max[3], const float* rec, const uint8_t* flags, const uint8_t* new_flags,
                 const uint16_t* material, const int32_t* path_depth, const float* radius_scale, uint32_t n_rs,
                 const uint32_t* reuse, uint32_t n_reuse, uint64_t vram_mb, int smis, int no_first_hit_stylization,
                 int disable_flag_check, int cluster, uint64_t seed, uint64_t table, const or_node* nodes,
                 const int32_t* roots, uint32_t n_materials, double* film, double* radiance_out, uint64_t* nbr_out,
                 uint32_t* cluster_out, const or_bsdf* bsdfs, uint32_t n_bsdfs) {
    const uint64_t npix = (uint64_t)W * H, N = npix * branch;
    double* ind = (double*)calloc(8 * N, sizeof(double));
    double* rad = (double*)calloc(3 * N * L, sizeof(double));
    uint32_t* pts = (uint32_t*)malloc(sizeof(uint32_t) * N);
    uint64_t* kv = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint32_t* cid = (uint32_t*)malloc(sizeof(uint32_t) * N);
    uint32_t* csize = (uint32_t*)malloc(sizeof(uint32_t) * N);
    uint64_t* prop = (uint64_t*)malloc(sizeof(uint64_t) * N);
    if (!ind || !rad || !pts || !kv || !cid || !csize || !prop) return -1;
    if (cluster_out)
        for (uint64_t i = 0; i < (uint64_t)L * N; ++i) cluster_out[i] = 0xFFFFFFFFu;
    uint32_t kmax = 1;
    for (uint32_t d = 0; d < L; ++d) {
        uint32_t K = d < n_reuse ? reuse[d] : reuse[n_reuse - 1];
        if (K > kmax) kmax = K;
    }
    uint64_t batch_max = 1;
    for (uint32_t d = 0; d < L; ++d) {
        uint64_t K = d < n_reuse ? reuse[d] : reuse[n_reuse - 1];
        uint64_t b = vram_mb * 1024 * 1024 / (K * 4);
        if (b < 1) b = 1;
        if (b > batch_max) batch_max = b;
    }
    if (batch_max > N) batch_max = N;
    pcg_t* rng = (pcg_t*)malloc(sizeof(pcg_t) * batch_max);
    if (!rng) return -1;
    if (std_pcg_streams(seed, batch_max, rng)) {
        free(rng);
        return -1;
    }
    float ext[3];
    for (int i = 0; i < 3; ++i) ext[i] = aabb_max[i] - aabb_min[i];
    const float avg = (ext[0] + ext[1] + ext[2]) / 3.0f;
    for (int d = (int)L - 1; d >= 0; --d) {
        lvl_t lv = {rec + (uint64_t)d * NP * N, N};
        uint32_t n = 0;
        for (uint64_t t = 0; t < N; ++t)
            if (path_depth[t] >= d) pts[n++] = (uint32_t)t;
        const float radius = avg * 0.01f * (d < (int)n_rs ? radius_scale[d] : radius_scale[n_rs - 1]);
        uint32_t res[3];
        for (int i = 0; i < 3; ++i) res[i] = (uint32_t)floorf(ext[i] / radius);
        for (uint32_t e = 0; e < n; ++e) {
            int32_t c[3];
            uint32_t i = pts[e];
            cell_of(aabb_min, aabb_max, res, (float)F(lv, 0, i), (float)F(lv, 1, i), (float)F(lv, 2, i), c);
            kv[e] = ((uint64_t)bucket_of(aabb_min, aabb_max, res, table, c) << 32) | i;
        }
        qsort(kv, n, sizeof(uint64_t), cmp_u64);
        const uint32_t K = d < (int)n_reuse ? reuse[d] : reuse[n_reuse - 1];
        if (cluster) cluster_level(lv, aabb_min, aabb_max, res, table, radius, K, pts, n, kv, cid, csize, prop);
        uint64_t batch = vram_mb * 1024 * 1024 / ((uint64_t)K * 4);
        if (batch < 1) batch = 1;
        uint64_t nbr_total = 0;
        double* radd = rad + (uint64_t)d * 3 * N;
        /* Points are independent except for the reservoir stream of batch slot tl, which the
           points e = tl, tl + batch, ... draw from in that order: slots run in parallel, each
           slot's points in order, so the results do not depend on the thread count. */
        const uint64_t nslots = batch < n ? batch : n;
        int alloc_fail = 0;
#pragma omp parallel reduction(+ : nbr_total) reduction(| : alloc_fail)
        {
            uint32_t* buf = (uint32_t*)malloc(sizeof(uint32_t) * kmax);
            if (!buf) alloc_fail = 1;
#pragma omp for schedule(dynamic, 64)
            for (uint64_t tl = 0; tl < nslots; ++tl) {
                for (uint64_t e = tl; e < n && buf; e += batch) {
                    const uint32_t c = pts[e];
                    const uint32_t q = cluster ? cid[c] : c;  /* the seed's position when clustered */
                    const float qx = (float)F(lv, 0, q), qy = (float)F(lv, 1, q), qz = (float)F(lv, 2, q);
                    const double ns[3] = {F(lv, 3, c), F(lv, 4, c), F(lv, 5, c)};
                    const double ng[3] = {F(lv, 6, c), F(lv, 7, c), F(lv, 8, c)};
                    const double wo[3] = {F(lv, 9, c), F(lv, 10, c), F(lv, 11, c)};
                    const double R[3] = {F(lv, 12, c), F(lv, 13, c), F(lv, 14, c)};
                    const or_bsdf* bc = bsdf_of(bsdfs, n_bsdfs, material[(uint64_t)d * N + c]);
                    const uint8_t myf = new_flags[(uint64_t)d * N + c];
                    int32_t cell[3];
                    cell_of(aabb_min, aabb_max, res, qx, qy, qz, cell);
                    uint32_t cnt = 0;
                    for (int dz = -1; dz <= 1; ++dz)
                        for (int dy = -1; dy <= 1; ++dy)
                            for (int dx = -1; dx <= 1; ++dx) {
                                int32_t nc[3] = {cell[0] + dx, cell[1] + dy, cell[2] + dz};
                                if (nc[0] < 0 || nc[1] < 0 || nc[2] < 0 || nc[0] >= (int)res[0] ||
                                    nc[1] >= (int)res[1] || nc[2] >= (int)res[2])
                                    continue;
                                uint64_t b = bucket_of(aabb_min, aabb_max, res, table, nc);
                                uint32_t lo = 0, hi = n;
                                while (lo < hi) {
                                    uint32_t mid = (lo + hi) / 2;
                                    if ((kv[mid] >> 32) < b) lo = mid + 1; else hi = mid;
                                }
                                for (uint32_t k = lo; k < n && (kv[k] >> 32) == b; ++k) {
                                    uint32_t j = (uint32_t)kv[k];
                                    float ddx = (float)F(lv, 0, j) - qx, ddy = (float)F(lv, 1, j) - qy,
                                          ddz = (float)F(lv, 2, j) - qz;
                                    if (!(sqrtf(ddx * ddx + ddy * ddy + ddz * ddz) <= radius)) continue;
                                    if (cluster && cid[j] != q) continue;
                                    if (!disable_flag_check && new_flags[(uint64_t)d * N + j] != myf) continue;
                                    if (cnt < K) {
                                        buf[cnt] = j;
                                    } else {
                                        uint32_t r = pcg_next(&rng[tl]) % (cnt + 1);
                                        if (r < K) buf[r] = j;
                                    }
                                    ++cnt;
                                }
                            }
                    nbr_total += cnt;
                    const uint32_t used = cnt < K ? cnt : K;
                    double direct[3] = {0, 0, 0}, indirect[3] = {0, 0, 0};
                    for (uint32_t k = 0; k < used; ++k) {
                        const uint32_t j = buf[k];
                        const double dpdf = F(lv, 25, j);
                        if (dpdf > 0) {
                            double wi[3] = {F(lv, 19, j), F(lv, 20, j), F(lv, 21, j)}, fc[3];
                            double bp = or_bsdf_eval(bc, R, ns, ng, wo, wi, fc);
                            double w = dpdf / (dpdf + bp);
                            for (int ch = 0; ch < 3; ++ch)
                                direct[ch] += fc[ch] * F(lv, 22 + ch, j) * F(lv, 26, j) * w / dpdf / (double)used;
                        }
                        double wi[3] = {ind[j], ind[N + j], ind[2 * N + j]};
                        if (sqrt(wi[0] * wi[0] + wi[1] * wi[1] + wi[2] * wi[2]) == 0) continue;
                        double fc[3];
                        double bp = or_bsdf_eval(bc, R, ns, ng, wo, wi, fc);
                        double reused[3];
                        for (int ch = 0; ch < 3; ++ch) reused[ch] = fc[ch] * ind[(3 + ch) * N + j] * ind[6 * N + j];
                        if (smis) {
                            double sum = 0;
                            for (uint32_t m = 0; m < used; ++m) {
                                uint32_t o = buf[m];
                                double ons[3] = {F(lv, 3, o), F(lv, 4, o), F(lv, 5, o)};
                                double ong[3] = {F(lv, 6, o), F(lv, 7, o), F(lv, 8, o)};
                                double owo[3] = {F(lv, 9, o), F(lv, 10, o), F(lv, 11, o)};
                                double oR[3] = {F(lv, 12, o), F(lv, 13, o), F(lv, 14, o)};
                                sum += or_bsdf_eval(bsdf_of(bsdfs, n_bsdfs, material[(uint64_t)d * N + o]), oR, ons, ong,
                                                    owo, wi, NULL);
                            }
                            for (int ch = 0; ch < 3; ++ch) indirect[ch] += reused[ch] / sum;
                        } else if (bp > 0) {
                            for (int ch = 0; ch < 3; ++ch) indirect[ch] += reused[ch] / bp / (double)used;
                        }
                    }
                    double Lr[3];
                    for (int ch = 0; ch < 3; ++ch) Lr[ch] = F(lv, 15 + ch, c) + direct[ch] + indirect[ch];
                    if (d > 0 || !no_first_hit_stylization) {
                        uint16_t m = material[(uint64_t)d * N + c];
                        int root = m < n_materials ? roots[m] : -1;
                        double p[3] = {F(lv, 0, c), F(lv, 1, c), F(lv, 2, c)};
                        if (root >= 0) or_apply_style(nodes, root, (uint32_t)d, myf, Lr, p, Lr);
                    }
                    for (int ch = 0; ch < 3; ++ch) radd[3 * (uint64_t)c + ch] = Lr[ch];
                }
            }
            free(buf);
        }
        if (alloc_fail) return -1;
        if (nbr_out) nbr_out[d] = nbr_total;
        if (cluster && cluster_out)
            for (uint32_t e = 0; e < n; ++e) cluster_out[(uint64_t)d * N + pts[e]] = cid[pts[e]];
        /* propagate to depth d-1 */
        memset(ind, 0, sizeof(double) * 8 * N);
        if (d > 0) {
            lvl_t pv = {rec + (uint64_t)(d - 1) * NP * N, N};
            for (uint32_t e = 0; e < n; ++e) {
                uint32_t i = pts[e];
                for (int ch = 0; ch < 3; ++ch) {
                    ind[ch * N + i] = -F(lv, 9 + ch, i);
                    ind[(3 + ch) * N + i] = radd[3 * (uint64_t)i + ch];
                }
                ind[6 * N + i] = F(lv, 18, i);
                ind[7 * N + i] = F(pv, 27, i);
            }
        }
    }
    /* splat + update_film */
    lvl_t l0 = {rec, N};
    for (uint64_t px = 0; px < npix; ++px) {
        double s[3] = {0, 0, 0}, acc[3] = {0, 0, 0}, wsum = 0;
        for (uint32_t b = 0; b < branch; ++b) {
            uint64_t t = (uint64_t)b * npix + px;
            if (path_depth[t] < 0) continue;
            for (int ch = 0; ch < 3; ++ch) s[ch] += rad[3 * t + ch];
        }
        for (uint32_t b = 0; b < branch; ++b) {
            uint64_t t = (uint64_t)b * npix + px;
            if (path_depth[t] < 0) continue;
            double Lr[3] = {s[0] / branch, s[1] / branch, s[2] / branch};
            if (!no_first_hit_stylization) {
                uint16_t m = material[t];
                int root = m < n_materials ? roots[m] : -1;
                double p[3] = {F(l0, 0, t), F(l0, 1, t), F(l0, 2, t)};
                if (root >= 0) or_apply_style(nodes, root, 0u, flags[t], Lr, p, Lr);
            }
            if (isnan(Lr[0]) || isnan(Lr[1]) || isnan(Lr[2])) Lr[0] = Lr[1] = Lr[2] = 0;
            for (int ch = 0; ch < 3; ++ch) acc[ch] += Lr[ch];
            wsum += 1;
        }
        for (int ch = 0; ch < 3; ++ch) film[3 * px + ch] = wsum > 0 ? acc[ch] / wsum : 0;
    }
    if (radiance_out) memcpy(radiance_out, rad, sizeof(double) * 3 * N * L);
    free(ind);
    free(rad);
    free(pts);
    free(kv);
    free(rng);
    free(cid);
    free(csize);
    free(prop);
    return 0;
}
