/*
 * TEST INFRASTRUCTURE ONLY — CPU oracle of the NL-PF hot path.  Loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker; never linked into
 * or called by libnlpf.
 *
 * The node table uses the same 256-byte layout as the library's stylization nodes (the
 * exchange format of the C ABI), so one table feeds both sides.
 */
#ifndef NLPF_ORACLE_H
#define NLPF_ORACLE_H
#include <stdint.h>

enum {
    OR_BASE = 0,
    OR_TOON = 1,
    OR_PALETTE = 2,
    OR_SATURATION = 3,
    OR_REX = 4,
    OR_EIGHT_BIT = 5,
    OR_HALFTONE = 6,
    OR_HUE_SHIFT = 7,
    OR_RAMP = 8,
    OR_ENABLE_IF = 9,
    OR_SET_IDX = 10
};

typedef struct or_node {
    int32_t type, inner, else_;
    uint32_t idx, depth_mask, max_depth, count, reserved;
    float p[56];
} or_node;

void or_rgb_to_hsv(const double c[3], double hsv[3]);
void or_hsv_to_rgb(const double hsv[3], double c[3]);
void or_apply_style(const or_node* nodes, int root, uint32_t depth, uint32_t flags, const double in[3],
                    const double p[3], double out[3]);
uint32_t or_update_flags(const or_node* nodes, int root, uint32_t flags);
int or_stylize_batch(const or_node* nodes, const int32_t* roots, uint32_t n_materials, const double* in,
                     double* out, uint64_t n, const uint32_t* depth, uint32_t depth0, const uint8_t* flags,
                     const uint16_t* material, const double* p);

/* Screen-mode NL-PF over FP32 record planes [S][L][15][H][W] (positions already absolute).
 * film [H][W][3], level_out optional [S][L][3][H][W]. */
int or_screen_run_f64(uint32_t W, uint32_t H, uint32_t S, uint32_t L, int radius, double sigma_pos,
                      double sigma_normal, double sigma_lum, const float* rec, const or_node* nodes, int root,
                      double* film, double* level_out, int threads);
int or_screen_run_f32(uint32_t W, uint32_t H, uint32_t S, uint32_t L, int radius, double sigma_pos,
                      double sigma_normal, double sigma_lum, const float* rec, const or_node* nodes, int root,
                      float* film, float* level_out, int threads);
/* The same with per-vertex material (uint16) and flags (uint8, NULL: 0) planes [S][L][H][W]:
 * vertex i takes roots[material[i]]'s tree (identity past n_materials or for root -1) with
 * flags[i]; material NULL: root's tree with flags 0 for every vertex. */
int or_screen_run_shaded_f64(uint32_t W, uint32_t H, uint32_t S, uint32_t L, int radius, double sigma_pos,
                             double sigma_normal, double sigma_lum, const float* rec, const or_node* nodes, int root,
                             const int32_t* roots, uint32_t n_materials, const uint16_t* material,
                             const uint8_t* flags, double* film, double* level_out, int threads);
int or_screen_run_shaded_f32(uint32_t W, uint32_t H, uint32_t S, uint32_t L, int radius, double sigma_pos,
                             double sigma_normal, double sigma_lum, const float* rec, const or_node* nodes, int root,
                             const int32_t* roots, uint32_t n_materials, const uint16_t* material,
                             const uint8_t* flags, float* film, float* level_out, int threads);

/* World-mode BSDFs per material (oracle_bsdf.c); the same 40-byte layout as the library's
 * nlpf_bsdf. */
enum { OR_BSDF_DIFFUSE = 0, OR_BSDF_PRINCIPLED = 1, OR_BSDF_EMISSION = 2 };
typedef struct or_bsdf {
    uint32_t kind;
    float metallic, roughness;
    float specular_tint[3];
    float specular_ior_level, transmission, coat_weight;
    uint32_t reserved;
} or_bsdf;
double or_ggx_d(double cos_h, double alpha);
double or_ggx_lambda(double cos_w, double alpha);
double or_fr_complex(double cos_i, double n, double k);
void or_conductor_nk(double r, double g, double* n, double* k);
/* pdf of wi; fc (optional) = f(wo, wi) |cos_i| per channel, R = the record's albedo (base colour) */
double or_bsdf_eval(const or_bsdf* b, const double R[3], const double ns[3], const double ng[3], const double wo[3],
                    const double wi[3], double fc[3]);

/* World-mode NL-PF, FP64; see oracle_world.c.  cluster != 0 runs the deterministic clustering
 * (cluster_out, optional, receives every level's cluster ids [L][N], 0xFFFFFFFF off-level). */
int or_world_neighbor_counts(const float* pos, uint32_t n, const float pmin[3], const float pmax[3], float radius,
                             uint64_t table, uint32_t* out_count);
int or_world_run(uint32_t W, uint32_t H, uint32_t branch, uint32_t L, const float aabb_min[3],
                 const float aabb_max[3], const float* rec, const uint8_t* flags, const uint8_t* new_flags,
                 const uint16_t* material, const int32_t* path_depth, const float* radius_scale, uint32_t n_rs,
                 const uint32_t* reuse, uint32_t n_reuse, uint64_t vram_mb, int smis, int no_first_hit_stylization,
                 int disable_flag_check, int cluster, uint64_t seed, uint64_t table, const or_node* nodes,
                 const int32_t* roots, uint32_t n_materials, double* film, double* radiance_out, uint64_t* nbr_out,
                 uint32_t* cluster_out, const or_bsdf* bsdfs, uint32_t n_bsdfs);

/* Random streams (oracle_world.c): Pcg32 new_seq_offset(seq, offset) draws; the ChaCha block
 * function; rand_core's seed_from_u64 key expansion; rand_chacha's ChaCha12 u64 stream; the
 * first `draws` outputs of reservoir streams first..first+count-1 as the reference seeds them. */
int or_pcg32_fill(uint64_t seq, uint64_t offset, uint32_t n, uint32_t* out);
void or_chacha_block(const uint32_t in[16], int rounds, uint32_t out[16]);
void or_std_seed_from_u64(uint64_t state, uint32_t key[8]);
void or_chacha12_u64(const uint32_t key[8], uint64_t first, uint64_t n, uint64_t* out);
int or_reservoir_streams(uint64_t seed, uint64_t first, uint64_t count, uint32_t draws, uint32_t* out);

/* FP64 reference gradient of the screen-mode NL-PF; returns the number of parameters
 * (3 + 16 * ramp nodes) written to grad, or -1. */
int or_screen_grad_f64(uint32_t W, uint32_t H, uint32_t S, uint32_t L, int radius, const double sig[3],
                       const float* rec, const or_node* nodes, uint32_t n_nodes, int root, const double* dfilm,
                       double* grad, double* loss_unused);

#endif
