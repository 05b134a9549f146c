/*
 * nlpf.h — C ABI of the MI355X-native nonlinear path-filtering (NL-PF) library.
 *
 * The library replaces the `"type":"pf"` integrator of the reference
 * (crates/akari_integrator/src/lm_path_filtering.rs:1802-1813, dispatched from
 * crates/akari_integrator/src/lib.rs:169-177) and the stylization hooks it calls
 * (crates/akari_render/src/svm/surface/mod.rs:262-286, 1824-1842).
 *
 * Conventions
 *  - every entry point returns an int status (NLPF_OK == 0); on failure
 *    nlpf_last_error(h) (or nlpf_last_error(NULL) for the calls without a handle:
 *    nlpf_create, nlpf_style_from_scene, nlpf_bsdfs_from_scene, nlpf_write_world_scene,
 *    nlpf_write_image_host)
 *    holds a message;
 *  - buffers are device pointers owned by the caller, or allocated through
 *    nlpf_alloc (hipMalloc on the handle's device);
 *  - one HIP stream per handle; calls on one handle are ordered on that stream;
 *    a handle is thread-compatible but not thread-safe;
 *  - no torch types cross this boundary: plain pointers and sizes only.
 */
#ifndef NLPF_H
#define NLPF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

/* ---- status codes (the reference panics instead: lm_path_filtering.rs:829, 927, 938) ---- */
#define NLPF_OK 0
#define NLPF_E_INVALID_CONFIG 1
#define NLPF_E_INVALID_ARG 2
#define NLPF_E_OOM 3
#define NLPF_E_HIP 4
#define NLPF_E_UNSUPPORTED 5
#define NLPF_E_NUMERIC 6
#define NLPF_E_IO 7

/* ---- storage precision of record planes ---- */
#define NLPF_FP32 0
#define NLPF_FP16 1
#define NLPF_FP64 2

/* ---- stylization node types ----
 * Each node wraps `inner` (its input surface) exactly as the reference's wrapper
 * surfaces do: apply_stylization first runs inner, then its own transform.
 *   BASE       plain BSDF, identity        (svm/surface/mod.rs:275-285 default)
 *   TOON       ToonSurface                 (svm/surface/mod.rs:572-610)
 *   PALETTE    ColorPaletteSurface         (svm/surface/mod.rs:696-735)
 *   SATURATION SaturationAdjustSurface     (svm/surface/mod.rs:820-851)
 *   REX        SaturationAdjustRexSurface  (svm/surface/mod.rs:1059-1089)
 *   EIGHT_BIT  EightBitSurface             (svm/surface/mod.rs:1315-1345)
 *   HALFTONE   HalfToneSurface             (svm/surface/mod.rs:1180-1233)
 *   HUE_SHIFT  sty_hue_shift: parsed by the reference (akari_scenegraph/src/shader.rs:218-223)
 *              but never compiled (svm/compiler.rs:478 todo!()); defined here as an HSV hue
 *              rotation h <- fract(h + shift) using color.rs:551-589.
 *   RAMP       colour ramp: no reference node; piecewise-linear luminance remap, hue preserved.
 *   ENABLE_IF  EnableOrSetStylizationSurface, mode_set=false (svm/surface/mod.rs:959-976)
 *   SET_IDX    EnableOrSetStylizationSurface, mode_set=true  (svm/surface/mod.rs:941-953)
 *
 * Parameter layout p[] per type:
 *   TOON       p[0]=color_band_width
 *   PALETTE    p[0]=indirect_scale p[1]=gamma p[2+3k..4+3k]=entry k rgb (count<=18)
 *   SATURATION p[0]=power p[1]=indirect_scale; max_depth field
 *   REX        p[0]=blend
 *   EIGHT_BIT  p[0]=indirect_scale
 *   HALFTONE   p[0]=indirect_scale p[1]=radius_scale p[2]=radius p[3]=grid_distance p[4]=blend
 *              p[5..7]=dark_tone p[8..10]=light_tone p[11..22]=affine 3x4 row-major; max_depth field
 *   HUE_SHIFT  p[0]=shift p[1]=indirect_scale (hue rotation is scale-equivariant, so unused)
 *   RAMP       p[0..15]=stop positions (ascending) p[16..31]=stop values (count<=16)
 * depth_mask: bit d set => the node's own transform runs at path depth d (bits >= 31 map to
 * bit 31). 0xFFFFFFFF reproduces the reference (no depth gating on toon/palette/...). */
#define NLPF_STY_BASE 0
#define NLPF_STY_TOON 1
#define NLPF_STY_PALETTE 2
#define NLPF_STY_SATURATION 3
#define NLPF_STY_REX 4
#define NLPF_STY_EIGHT_BIT 5
#define NLPF_STY_HALFTONE 6
#define NLPF_STY_HUE_SHIFT 7
#define NLPF_STY_RAMP 8
#define NLPF_STY_ENABLE_IF 9
#define NLPF_STY_SET_IDX 10

#define NLPF_STY_MAX_PARAMS 56
#define NLPF_STY_MAX_NODES 1024
#define NLPF_STY_MAX_MATERIALS 65536

typedef struct nlpf_style_node {
    int32_t type;
    int32_t inner;       /* wrapped node index, -1 = base BSDF (identity) */
    int32_t else_;       /* ENABLE_IF else branch, -1 = null surface (identity) */
    uint32_t idx;        /* ENABLE_IF / SET_IDX style index 0..3 */
    uint32_t depth_mask; /* see above */
    uint32_t max_depth;  /* SATURATION / HALFTONE max_depth */
    uint32_t count;      /* PALETTE entries / RAMP stops */
    uint32_t reserved;
    float p[NLPF_STY_MAX_PARAMS];
} nlpf_style_node; /* 256 bytes */

/* StylizationFlags (svm/surface/mod.rs:58-101) packed into one byte:
 * bit0 stylization_disabled, bit1 use_style_2, bit2 use_style_3, bit3 use_style_4. */
#define NLPF_FLAG_DISABLED 1u
#define NLPF_FLAG_STYLE2 2u
#define NLPF_FLAG_STYLE3 4u
#define NLPF_FLAG_STYLE4 8u

typedef struct nlpf_handle nlpf_handle;

/* ---- lifecycle ----
 * config_json: the reference's "pf" method schema (lm_path_filtering.rs:264-298, all fields
 * optional with the reference defaults: max_depth 7, repeats 1, branch 1, radius_scale
 * [0.1,0.05], reuse [128,64], vram 2048, smis true, no_first_hit_stylization false,
 * disable_flag_check false, cluster true, scene_bounds null) plus the screen-mode fields
 * mode ("world"|"screen", default "world"), pixel_radius (7), sigma_pos (0.05),
 * sigma_normal (0.9), sigma_lum (0.5), precision ("fp32"|"fp16"|"fp64"), seed (0),
 * hash_table_size (64 Mi buckets), world_kernel ("auto"|"thread"|"wave": world-mode filter
 * with one thread or one wavefront per vertex, same results; "auto" picks by the expected
 * neighbour count), world_grid ("auto"|"sort": the world-mode grid's counting build, or the
 * radix sort at every depth; bit-identical results) and screen_kernel ("auto"|"tiles"|"strips":
 * the screen-mode level kernel; results agree within FP32 summation order).
 * Unknown keys are ignored, as serde does for the reference (no deny_unknown_fields). */
/* Replaces PathFiltering::new / Config deserialisation (lm_path_filtering.rs:264-298, 1317)
 * and the device setup of lm_path_filtering::render (:1802-1813). */
int nlpf_create(const char* config_json, nlpf_handle** out);
int nlpf_destroy(nlpf_handle* h);
/* Message of the last failing call on h; h == NULL gives the last failure of a call without a
 * handle (nlpf_create, nlpf_style_from_scene, nlpf_bsdfs_from_scene, nlpf_write_world_scene,
 * nlpf_write_image_host). */
const char* nlpf_last_error(const nlpf_handle* h);
/* Parsed configuration echoed back as JSON (for host-side mirrors and tests). */
int nlpf_get_config_json(const nlpf_handle* h, char* buf, size_t cap);
/* Overwrite the screen-mode bandwidths (sigma_pos, sigma_normal, sigma_lum); no reference
 * counterpart (screen mode is an addition). */
int nlpf_set_sigmas(nlpf_handle* h, float sigma_pos, float sigma_normal, float sigma_lum);

/* ---- device memory (hipMalloc / hipMemcpy on the handle's device and stream) ----
 * Stand in for the LuisaCompute device buffers the reference allocates per render
 * (lm_path_filtering.rs:1319-1560: Buffer<Vertex>, per-depth vertex buffers, film). */
int nlpf_alloc(nlpf_handle* h, size_t bytes, void** dptr);
int nlpf_free(nlpf_handle* h, void* dptr);
int nlpf_upload(nlpf_handle* h, void* dst_device, const void* src_host, size_t bytes);
int nlpf_download(nlpf_handle* h, void* dst_host, const void* src_device, size_t bytes);
int nlpf_memset(nlpf_handle* h, void* dst_device, int value, size_t bytes);
int nlpf_copy(nlpf_handle* h, void* dst_device, const void* src_device, size_t bytes);
int nlpf_sync(nlpf_handle* h);

/* ---- stylization ----
 * Replaces the SVM compile of sty_* nodes (svm/compiler.rs:337-477) with a flat node table.
 * material_roots[m] = root node of material m's stylization tree (-1 = unstylized). */
int nlpf_set_style(nlpf_handle* h, const nlpf_style_node* nodes, uint32_t n_nodes,
                   const int32_t* material_roots, uint32_t n_materials);

/* Build a style table from a reference scene file (akari_scenegraph JSON, the scene format the
 * reference loads; akari_scenegraph/src/scene.rs:121-129, shader.rs:24-28): material m's root is
 * the lowering of the shader graph of the material named materials[m]. Lowering follows the SVM
 * compiler (svm/compiler.rs:337-477), starting from the graph's output node:
 *  - sty_* nodes become nodes of this table, and their inputs must be constant
 *    (float / rgb / float3 nodes);
 *  - any other surface node ends the chain as the identity;
 *  - identical subtrees share one node.
 * On entry *n_nodes is the capacity of nodes[]; on return it is the number written. Host only
 * (no GPU, no handle); errors via nlpf_last_error(NULL). Pass the result to nlpf_set_style. */
int nlpf_style_from_scene(const char* scene_json, const char* const* materials, uint32_t n_materials,
                          nlpf_style_node* nodes, uint32_t* n_nodes, int32_t* material_roots);

/* Apply SurfaceClosure::apply_stylization (svm/surface/mod.rs:1824-1842) of material
 * `material[i]` (or material 0 when material == NULL) at path depth depth[i] (or `depth0`
 * when depth == NULL) with flags[i] (0 when NULL) and world position p[i] (halftone only;
 * may be NULL otherwise) to rgb_in[i]; all arrays are device pointers, n elements,
 * rgb as 3 interleaved floats. */
int nlpf_stylize(nlpf_handle* h, const float* rgb_in, float* rgb_out, uint64_t n,
                 const uint32_t* depth, uint32_t depth0, const uint8_t* flags,
                 const uint16_t* material, const float* p);

/* FP16 records store radiance divided by this power of two so fireflies stay finite. */
#define NLPF_FP16_RADIANCE_SCALE 64.0f

/* ---- screen mode: per-(spp slice, depth) cross-bilateral NL-PF ----
 * Record planes, element type by precision (float / half / double), layout
 *   records[((s*levels + d)*NLPF_SCREEN_PLANES + c)*H*W + y*W + x]
 * planes c: 0-2 position, 3-5 normal, 6-8 albedo, 9-11 radiance, 12-14 throughput.
 * FP16 positions are stored as offsets from a per-8x8-block FP32 anchor
 *   anchors[((s*levels + d)*ceil(H/8) + by)*ceil(W/8) + bx][3]  (NULL for FP32/FP64).
 * Filter per level (depth D..0):
 *   in_i  = radiance_i + throughput_i * out_{d+1,i}        (out_{D+1} = 0)
 *   w_ij  = exp(-|p_i-p_j|^2/2sp^2 - (1-n_i.n_j)^2/2sn^2 - (l_i-l_j)^2/2sl^2), l = lum(albedo)
 *   out_i = S_d( sum_j w_ij in_j / sum_j w_ij ),  j over the (2r+1)^2 window clipped to the image
 * S_d = material 0's stylization tree at depth d (flags 0; nlpf_screen_run_materials takes each
 * vertex's material and flags). film[y][x][c] = mean_s out_{0}.
 * Outputs are double for NLPF_FP64 and float otherwise:
 *   film_rgb  [H][W][3];
 *   level_out (optional, NULL to skip) every level's output [spp][levels][3][H][W]. */
#define NLPF_SCREEN_PLANES 15
typedef struct nlpf_screen_desc {
    uint32_t width, height, spp, levels;
    int32_t precision;
    int32_t radius;
} nlpf_screen_desc;

/* Screen-mode pass: no reference function (the screen-space filtering workload defines it); it takes the
 * place of the depth loop PathFiltering::render runs (lm_path_filtering.rs:1673-1774:
 * kernel_filter :999-1060 per depth, kernel_propagate :973-998 to the parent, kernel_splat +
 * kernel_update_film :428-506) with an image-space window instead of the hash grid. */
int nlpf_screen_run(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records,
                    const float* anchors, void* film_rgb, void* level_out);

/* nlpf_screen_run with each vertex's own material and stylization flags in place of material 0
 * and flags 0: vertex i of level d is stylized by S = material[i]'s tree (identity when
 * material[i] >= the style table's material count or its root is -1) with flags[i], as
 * kernel_filter stylizes a vertex through its own surface closure with its own flags
 * (lm_path_filtering.rs:947-969, apply_stylization svm/surface/mod.rs:1824-1842).  material
 * (uint16) and flags (uint8, NULL: all 0) are device planes laid out like one record plane,
 *   material[(s*levels + d)*H*W + y*W + x].
 * Halftone positions are read whenever any material's tree holds a halftone node. */
int nlpf_screen_run_materials(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records,
                              const float* anchors, const uint16_t* material, const uint8_t* flags,
                              void* film_rgb, void* level_out);

/* One chunk of a frame too large for HBM at once: desc->spp slices that are spp slices
 * [slice0, slice0 + desc->spp) of a total_spp-slice frame (records, anchors and level_out
 * hold just this chunk).  Chunks are passed in slice order, slice0 = 0 first; film_rgb holds
 * the running per-pixel sums of out_0 between calls and, after the chunk that ends at
 * total_spp, the frame's film, equal bit for bit to one nlpf_screen_run over all total_spp
 * slices.  No reference function: it stands for the memory-limited batch loop of
 * PathFiltering::render (lm_path_filtering.rs:1711-1755) in screen mode. */
int nlpf_screen_run_chunk(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records,
                          const float* anchors, void* film_rgb, void* level_out, uint32_t slice0,
                          uint32_t total_spp);

/* nlpf_screen_run without level_out, allocating nothing: no hipMalloc / hipFree, no event, no
 * synchronisation (film_rgb is ready once h's stream is; nlpf_sync waits for it).  No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  The level outputs live in a static device array of
 * NLPF_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the process ends;
 * slices run in groups of as many as fit (two 3 x H x W images of float, double for FP64, per
 * slice; one for a single level), so one slice must fit: FP32/FP16 up to 11.1 M pixels (4K),
 * FP64 up to 5.5 M.  Larger images return NLPF_E_INVALID_ARG.  film_rgb equals
 * nlpf_screen_run's bit for bit; it holds partial sums until the pass has finished.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (several handles or
 * streams) race.  Timings in nlpf_get_stats stay 0. */
#define NLPF_NOALLOC_SCRATCH_BYTES (256ull << 20)
int nlpf_screen_run_noalloc(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records,
                            const float* anchors, void* film_rgb);

/* Adjoint of nlpf_screen_run (FP32 records) with respect to the bandwidths and every colour-ramp
 * node's stop values, for a loss whose gradient with respect to the film is dL_dfilm
 * ([H][W][3], device).  level_out must hold the forward pass's level outputs (same records,
 * sigmas and style table).  Weights are recomputed per tile, never stored.  grad_params
 * (device, float[3 + 16 * ramps]) receives dL/dsigma_pos, dL/dsigma_normal, dL/dsigma_lum,
 * then dL/dvalue[k] of the r-th RAMP node of the style table at 3 + 16 r + k.  At most 2 ramp
 * nodes and a chain of at most 8 stylization nodes.  No reference counterpart (the reference
 * has no differentiable NL-PF). */
int nlpf_screen_backward(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records,
                         const void* level_out, const float* dL_dfilm, float* grad_params);

/* Kernel time (ms, HIP events on the handle's stream) of the last nlpf_screen_run, total
 * and per level (levels entries, deepest level first in execution order, indexed by depth).
 * The reference only logs timings and the average neighbour count
 * (lm_path_filtering.rs:1767-1773, 1798). */
typedef struct nlpf_stats {
    double total_ms;
    double level_ms[32];
    double film_ms;
    uint64_t vertices;
    uint64_t bytes_algorithmic;  /* records read once, level outputs written once, read once */
    double flops_algorithmic;    /* 27 per neighbour pair (window^2 pairs per vertex) */
    /* world mode: points gathered and neighbours visited per depth (the reference logs the
     * average, lm_path_filtering.rs:1767-1773) */
    uint64_t points[32];
    uint64_t neighbors[32];
    /* world mode with SMIS: sum over the depth's points of (reused neighbours)^2, the pdf
     * evaluations of the SMIS denominators */
    uint64_t smis_pairs[32];
} nlpf_stats;
int nlpf_get_stats(const nlpf_handle* h, nlpf_stats* out);

/* Parameters of the synthetic record generators below. */
typedef struct nlpf_gen_desc {
    uint64_t seed;
    float mu, sigma;
    float firefly_prob, firefly_scale;
    float albedo_lo, albedo_hi;
} nlpf_gen_desc;

/* ---- world mode: the reference's hash-grid NL-PF ----
 * Replaces PathFiltering::render's filter loop (lm_path_filtering.rs:1673-1788) for records
 * traced elsewhere (kernel_mk_trace :1062-1218 or nlpf_generate_world_records).
 * N = width*height*branch paths, path t = b*width*height + pixel.  Float planes per level d,
 * layout records[(d*NLPF_WORLD_PLANES + c)*N + t]:
 *   0-2 p, 3-5 ns (shading normal), 6-8 ng (geometric normal), 9-11 wo, 12-14 albedo R
 *   (base colour: a diffuse material has f = R/pi |cos|; see nlpf_set_bsdfs), 15-17 le,
 *   18 mis (emission MIS weight),
 *   19-21 direct.wi, 22-24 direct.irradiance, 25 direct.pdf, 26 direct.mis, 27 bsdf_pdf.
 * flags / new_flags: StylizationFlags bytes [L][N]; material: stylization material [L][N];
 * path_depth[t]: last depth with a vertex, -1 for a path whose camera ray missed.
 * The indirect record of depth d is produced by the filter of depth d+1 (kernel_propagate,
 * :973-998).  film: with desc->film_accumulate == 0, [H][W][3] receives the pass's film after
 * kernel_splat + kernel_update_film (:428-506, depth-0 re-stylization of the spp mean included),
 * i.e. the mean of the pass's samples per pixel (0 where no path hit).  With film_accumulate != 0,
 * film is a film buffer (NLPF_FILM_FLOATS, below) and the pass adds its samples to it as
 * Film::add_sample does (one sample of weight 1 per path that hit, film.rs:187-222), so repeated
 * passes accumulate as the reference's `repeats` loop does (:1629-1783); nlpf_film_resolve then gives
 * the image.  Per pixel the pass's samples are added in branch order.  radiance_out (optional)
 * [L][N][3] = every vertex's filtered, stylized radiance (v.radiance). Config fields used:
 * max_depth, branch, radius_scale, reuse, vram (batching), smis, cluster,
 * no_first_hit_stylization, disable_flag_check, scene_bounds, seed, hash_table_size.
 * Reservoir draws (more candidates than reuse) of the t-th point of a batch come from stream t,
 * Pcg32 new_seq_offset(t, t-th u64 of StdRng::seed_from_u64(seed)), as
 * init_pcg32_buffer_with_seed (sampler/mod.rs:157-174) seeds the reference's rng_buf (with seed
 * 0, :1349-1353); streams advance across depths.  The reference's tracer draws from the same
 * buffer before filtering, so its filter starts from streams its trace has advanced. */
#define NLPF_WORLD_PLANES 28
typedef struct nlpf_world_desc {
    uint32_t width, height, branch, levels;
    float aabb_min[3], aabb_max[3]; /* scene AABB, used unless the config sets scene_bounds */
    uint32_t film_accumulate;       /* 0: film is [H][W][3]; else a film buffer (see above) */
} nlpf_world_desc;
int nlpf_world_run(nlpf_handle* h, const nlpf_world_desc* desc, const float* records,
                   const uint8_t* flags, const uint8_t* new_flags, const uint16_t* material,
                   const int32_t* path_depth, float* film, float* radiance_out);

/* ---- world mode: each material's BSDF ----
 * The world filter evaluates every centre vertex's BSDF for its neighbours' directions and, for
 * SMIS, every reused neighbour's own BSDF (Svm::dispatch_surface, svm/eval.rs:589-624, through
 * SurfaceClosure::evaluate, svm/surface/mod.rs:1701-1744).  Entry m of the table is material
 * m's BSDF; a material without an entry is diffuse.  The base colour is the record's albedo
 * plane in every kind.
 *   NLPF_BSDF_DIFFUSE     DiffuseBsdf, reflectance albedo / pi (svm/surface/diffuse.rs:22-41,
 *                         101-121).
 *   NLPF_BSDF_PRINCIPLED  the principled BSDF (svm/surface/principled.rs:120-330) where it needs
 *                         no precomputed albedo table: coat_weight == 0 and either
 *                         metallic >= 1 - 1e-4 (the metal lobe; specular_ior_level and
 *                         transmission are then never evaluated) or specular_ior_level == 0 and
 *                         transmission == 0 (the specular layer's weight is then 0).  It is then
 *                         Mix(diffuse, metal, metallic) (BsdfMixture, mod.rs:1560-1593), the metal
 *                         lobe a Trowbridge-Reitz reflection (roughness^2, at least 1e-4, visible
 *                         normal pdf; microfacet.rs) with the conductor Fresnel of
 *                         artistic_to_conductor_fresnel(albedo, specular_tint) (mod.rs:2128-2157).
 *                         Other principled parameter sets are rejected with NLPF_E_UNSUPPORTED.
 *   NLPF_BSDF_EMISSION    an emission-only surface: no reflection (EmissiveSurface without an
 *                         inner BSDF, mod.rs:1347-1372).
 * No reference function takes this table: it stands for the SVM material data the reference's
 * filter reads per hit. */
#define NLPF_BSDF_DIFFUSE 0u
#define NLPF_BSDF_PRINCIPLED 1u
#define NLPF_BSDF_EMISSION 2u
typedef struct nlpf_bsdf {
    uint32_t kind;
    float metallic, roughness;
    float specular_tint[3];
    float specular_ior_level, transmission, coat_weight;
    uint32_t reserved; /* 0 */
} nlpf_bsdf;            /* 40 bytes */
int nlpf_set_bsdfs(nlpf_handle* h, const nlpf_bsdf* table, uint32_t n_materials);
/* Read each named material's BSDF from a reference scene file (the graph below its output node,
 * through the sty_* nodes' "surface" inputs, which is what their closures evaluate): "diffuse" ->
 * NLPF_BSDF_DIFFUSE, "emission" -> NLPF_BSDF_EMISSION, "principled" -> NLPF_BSDF_PRINCIPLED with
 * its constant metallic, roughness, specular_tint, specular_ior_level, transmission_weight and
 * coat_weight inputs (the base colour may be anything: it comes from the records).  Any other
 * surface, a non-constant input of those, a normal map, or a principled parameter set outside
 * the subset above fails with a message naming the material. */
int nlpf_bsdfs_from_scene(const char* scene_json, const char* const* materials, uint32_t n_materials,
                          nlpf_bsdf* out);

/* Trace the procedural scene (below) into world-mode records, one path per (pixel, branch),
 * as kernel_mk_trace does (lm_path_filtering.rs:1062-1218, area-light NEE light/area.rs:50-131): camera ray, NEE to the quad light with its radiance stylized at
 * depth+1 by material `light material`, cosine BSDF sampling, StylizationFlags updated by
 * each hit's material root. Materials: 0 floor 1 ceiling 2 back 3 left 4 right 5 sphere0
 * 6 sphere1 7 light; wall/sphere albedos U[albedo_lo, albedo_hi] per material from `seed`. */
int nlpf_generate_world_records(nlpf_handle* h, const nlpf_world_desc* desc,
                                const nlpf_gen_desc* gen, float* records, uint8_t* flags,
                                uint8_t* new_flags, uint16_t* material, int32_t* path_depth);

/* ---- film (crates/akari_render/src/film.rs) ----
 * A film buffer is one device buffer of NLPF_FILM_FLOATS(width, height) floats laid out as the
 * reference's Film::data for the sRGB representation (film.rs:66-67, 88-93, 108-110):
 *   [0, 3P)   pixels:  rgb of pixel i at 3i .. 3i+2  (P = width * height)
 *   [3P, 6P)  splat:   rgb of pixel i at 3P + 3i ..
 *   [6P, 7P)  weights: weight of pixel i at 6P + i
 * Samples are (x, y) positions in pixels (xy[2n]), rgb[3n] and an optional weight[n] (NULL
 * means 1); a sample lands in pixel (floor x, floor y) (Film::linear_index, film.rs:150-160).
 * The reference adds with float atomics (film.rs:180-184, 210-219), so its sums depend on
 * scheduling; here a pixel's samples are added in input order, onto the film's current values,
 * so results are the sequential sums and repeat bit for bit.  Samples outside the film (or at a
 * NaN position), which the reference only asserts against in debug builds, are skipped and
 * counted in *dropped (optional).  width * height < 2^31, n < 2^31. */
#define NLPF_FILM_FLOATS(w, h) (7ull * (uint64_t)(w) * (uint64_t)(h))
/* Film::clear (film.rs:224-227): every value to 0. */
int nlpf_film_clear(nlpf_handle* h, float* film, uint32_t width, uint32_t height);
/* Film::add_sample (film.rs:187-222): c = remove_nan(rgb) * w; pixels += c, weights += w
 * (remove_nan zeroes a colour with any NaN channel, color.rs:319-327). */
int nlpf_film_add_samples(nlpf_handle* h, float* film, uint32_t width, uint32_t height,
                          const float* xy, const float* rgb, const float* weight, uint64_t n,
                          uint64_t* dropped);
/* Film::add_splat (film.rs:162-186): c = remove_nan(rgb) * w, a NaN channel of c set to 0;
 * splat += c. */
int nlpf_film_add_splats(nlpf_handle* h, float* film, uint32_t width, uint32_t height,
                         const float* xy, const float* rgb, const float* weight, uint64_t n,
                         uint64_t* dropped);
/* Film::copy_to_rgba_image (film.rs:120-147, 229-246): rgba[4P] (device) gets
 * pixels / (weight == 0 ? 1 : weight) + splat * splat_scale, sRGB-encoded (color.rs:536-539)
 * unless hdr, alpha 1. */
int nlpf_film_resolve(nlpf_handle* h, const float* film, uint32_t width, uint32_t height,
                      float splat_scale, int hdr, float* rgba);

/* ---- image output (crates/akari_render/src/util/mod.rs:57-121) ----
 * util::write_image: a path ending in ".exr" gets the R, G, B channels as 32-bit floats in an
 * uncompressed scanline OpenEXR file; any other path gets an 8-bit RGB PNG of
 * (linear_to_srgb(c) * 255).clamp(0, 255) truncated (color.rs:590-596, util/mod.rs:85-93).
 * The parent directory is created.  rgba[4 * width * height] is row-major, top row first.
 * nlpf_write_image reads rgba from the device (after the handle's earlier work); the _host
 * variant reads host memory and needs no GPU. */
int nlpf_write_image(nlpf_handle* h, const float* rgba, uint32_t width, uint32_t height,
                     const char* path);
int nlpf_write_image_host(const float* rgba, uint32_t width, uint32_t height, const char* path);

/* ---- the procedural scene as a reference scene file ----
 * Writes the scene that nlpf_generate_world_records traces as an akari_scenegraph scene
 * (akari_scenegraph/src/scene.rs:121-129, 368-383): `path` gets the JSON, and "<stem>.bin"
 * next to it gets the one buffer it names (resolved relative to the JSON, scene.rs:641-690).
 *  - Geometry is triangle meshes, one instance per material: exact quads for the walls and the
 *    light, and UV spheres of sphere_segments segments (even, 4..4096) with analytic vertex
 *    normals.
 *  - Materials are named "mat<m>_<name>", so that they sort in material-id order. Each holds a
 *    diffuse BSDF of the generator's reflectance for `gen` (the light instead holds an emission
 *    node of colour 1 and strength 15), or, where the BSDF table (bsdfs, n_bsdfs; may be empty)
 *    gives one, a principled node with that entry's parameters and the reflectance as base
 *    colour, or an emission node of strength 0 for NLPF_BSDF_EMISSION; it is wrapped in the
 *    sty_* nodes lifted from the style table (nodes, roots).  nlpf_style_from_scene and
 *    nlpf_bsdfs_from_scene read the same tables back. Hue shift, ramp and
 *    depth-gated nodes have no reference counterpart: NLPF_E_UNSUPPORTED.
 *  - The camera is the generator's pinhole, in the reference's PerspectiveCamera conventions
 *    (camera/mod.rs:130-152). Rendering the file with a box pixel filter of radius 1 matches
 *    the generator's jitter.
 * Host only; errors via nlpf_last_error(NULL). No reference counterpart (the reference's scenes
 * come from its Blender exporter). */
int nlpf_write_world_scene(const char* path, uint32_t width, uint32_t height, const nlpf_gen_desc* gen,
                           const nlpf_style_node* nodes, uint32_t n_nodes, const int32_t* material_roots,
                           uint32_t n_materials, uint32_t sphere_segments, const nlpf_bsdf* bsdfs,
                           uint32_t n_bsdfs);

/* ---- random streams (host side, the same code the kernels run) ----
 * nlpf_pcg32_fill: n draws of Pcg32 new_seq_offset(seq, offset).gen_u32() (sampler/mod.rs:91-117).
 * nlpf_reservoir_streams: the first `draws` gen_u32 outputs of world-mode reservoir streams
 * first..first+count-1 for config seed `seed` ([count][draws] into host_out).
 * No reference entry point: the reference's sampler is internal to its kernels; these let a
 * caller reproduce the library's draws. */
int nlpf_pcg32_fill(uint64_t seq, uint64_t offset, uint32_t n, uint32_t* host_out);
int nlpf_reservoir_streams(uint64_t seed, uint64_t first, uint64_t count, uint32_t draws, uint32_t* host_out);

/* ---- learned inner-expectation cache (radiance-cache MLP) ----
 * 84 inputs (clamped position, sin/cos of 2^k pi p for k = 0..11 per coordinate, normal,
 * albedo, view direction towards the parent vertex) padded to 96, five hidden layers of 64
 * with ReLU, 3 linear outputs (padded to 16); BF16 operands on MFMA with FP32 accumulation,
 * FP32 master weights and Adam.  Parameters: W_l [out][in] then b_l [out] for l = 0..5
 * (nlpf_mlp_param_count() floats).  Samples address screen records as slice*W*H + pixel
 * at level `depth`; `feats` ([n][96] floats, device) replaces the record encoder when given.
 * No reference counterpart: the reference's learned cache is tiny-cuda-nn's NRC
 * (crates/akari_integrator/src/nrc_pt.rs:1159-1353), a different network. */
uint64_t nlpf_mlp_param_count(void);
int nlpf_mlp_init(nlpf_handle* h, uint64_t seed);
int nlpf_mlp_get_params(nlpf_handle* h, float* host_params, uint64_t n);
int nlpf_mlp_set_params(nlpf_handle* h, const float* host_params, uint64_t n);
int nlpf_mlp_encode(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records,
                    const float* anchors, uint32_t depth, const uint32_t* idx, uint64_t n,
                    float* feats);
int nlpf_mlp_infer(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records,
                   const float* anchors, uint32_t depth, const uint32_t* idx, const float* feats,
                   uint64_t n, float* out_rgb);
/* One training step on `batch` samples (a multiple of 128) against target [batch][3] (device):
 * L2 loss mean_i |y_i - t_i|^2, Adam (beta 0.9/0.999, eps 1e-8) with learning rate lr when
 * apply_update != 0.  loss (host) receives the pre-update loss, grad (device, optional,
 * nlpf_mlp_param_count() floats) the gradient. */
int nlpf_mlp_train_step(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records,
                        const float* anchors, uint32_t depth, const uint32_t* idx,
                        const float* feats, const float* target, uint32_t batch, float lr,
                        int apply_update, float* loss, float* grad);
/* Inference weights as an exponential moving average of the trained parameters:
 * ema_t = decay ema_{t-1} + (1 - decay) theta_t after every applied Adam step, used
 * bias-corrected (ema_t / (1 - decay^t)); decay 0 (the default) uses the parameters themselves.
 * Setting it (and nlpf_mlp_init / nlpf_mlp_set_params) restarts the average at the current
 * parameters.  nlpf_mlp_infer, nlpf_mlp_reference and the cache pass use these weights,
 * training uses the parameters. */
int nlpf_mlp_set_ema(nlpf_handle* h, float decay);
/* FP64 VALU evaluation of the same network (bf16 weights and activations as on MFMA, FP64
 * sums): the error oracle of the learned path. */
int nlpf_mlp_reference(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records,
                       const float* anchors, uint32_t depth, const uint32_t* idx, uint64_t n,
                       double* out_rgb);

/* Screen-mode NL-PF with the learned cache standing in for the filtered estimate at depth
 * `depth` (records deeper than `depth` are read only for the audit slice and fallback tiles):
 *   audit slice  spp slice `audit_slice` filtered exactly at every level: training targets
 *                (its unstylized depth-`depth` values at `train_records` vertices, Adam steps
 *                of `batch`, rate `lr`) and its exact level-0 image;
 *   inference    the network's estimate at depth `depth` for every slice, stylized there;
 *   audit        levels depth-1..0 of the audit slice rerun on the estimate; 64x8 tiles whose
 *                relative L2 error against the exact level-0 image exceeds `tile_threshold`
 *                are filtered exactly at depths levels-1..depth in every slice (with the
 *                dependency halo: the level-0 tile mask dilated by one tile per level);
 *   levels depth-1..0 and the film as nlpf_screen_run.
 * The audit slice's exact levels depth-1..0 run on a second stream of the handle, beside the
 * training and inference; the call returns after both streams are done.
 * The network must be initialised (nlpf_mlp_init); training continues across calls.
 * film: float [H][W][3]; level_out (optional): float [spp][levels][3][H][W], where the levels
 * deeper than `depth` of the non-audit slices are defined only inside fallback tiles. */
#define NLPF_CACHE_NETWORK 0 /* the trained network (the learned cache) */
#define NLPF_CACHE_ZERO 1    /* control: depth-`depth` output 0, the paths cut there (no training) */
#define NLPF_CACHE_MEAN 2    /* control: the audit slice's mean unstylized depth-`depth` value at every
                                vertex, stylized there (no training) */
typedef struct nlpf_mlp_run_desc {
    uint32_t depth;
    uint32_t audit_slice;
    uint32_t train_records;
    uint32_t batch;
    float lr;
    float tile_threshold;
    uint64_t seed;
    uint32_t estimator; /* NLPF_CACHE_*: what stands in for the filtered depth-`depth` values */
    uint32_t reserved;
} nlpf_mlp_run_desc;
typedef struct nlpf_mlp_run_stats {
    double total_ms, audit_ms, train_ms, infer_ms, fallback_ms, filter_ms;
    uint64_t tiles;                   /* 64x8 tiles per image */
    uint64_t fallback_tiles;          /* level-0 tiles over the threshold */
    uint64_t fallback_tiles_filtered; /* tile-levels refiltered exactly per slice (halo included) */
    float train_loss;                 /* last training batch's loss */
    float cache_rel_l2;               /* estimate vs exact unstylized depth-`depth` values, audit slice */
    float audit_rel_l2;               /* level-0 image from the estimate vs exact, audit slice */
} nlpf_mlp_run_stats;
int nlpf_screen_run_mlp(nlpf_handle* h, const nlpf_screen_desc* desc, const void* records,
                        const float* anchors, const nlpf_mlp_run_desc* run, float* film,
                        float* level_out, nlpf_mlp_run_stats* stats);

/* ---- synthetic records (procedural unit box + 2 spheres + quad light) ----
 * Traces one path per (pixel, spp slice) through the analytic scene and writes screen-mode
 * records: p, n from the hits; albedo U[0.1,0.9]; radiance lognormal(mu, sigma) with a
 * `firefly_prob` fraction multiplied by `firefly_scale`; throughput U[0,1]; all from PCG32
 * streams keyed by (seed, record index). FP16 output also writes the anchors. */
int nlpf_generate_screen_records(nlpf_handle* h, const nlpf_screen_desc* desc,
                                 const nlpf_gen_desc* gen, void* records, float* anchors);
/* Convert FP32 records to FP16 records + anchors (same layout rules as above).  The synthetic
 * generator and the conversion have no reference counterpart: the reference's records come
 * from kernel_mk_trace (lm_path_filtering.rs:1062-1218) and are FP32 only. */
int nlpf_convert_records_fp16(nlpf_handle* h, const nlpf_screen_desc* desc_fp32,
                              const float* records_fp32, void* records_fp16, float* anchors);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif
#ifdef __cplusplus
}
#endif
#endif /* NLPF_H */
