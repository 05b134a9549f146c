/*
 * TEST INFRASTRUCTURE ONLY — screen-mode NL-PF oracle (CPU, OpenMP).  The FP64 instance is
 * the error oracle; the FP32 instance is the CPU baseline that bench.py times.  Loaded only
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Screen mode has no reference counterpart (the reference NL-PF is the world-space gather of
 * crates/akari_integrator/src/lm_path_filtering.rs:757-972); its per-level recursion follows
 * the reference's depth loop D..0 (:1673-1774) with the child's filtered output propagated
 * to the parent as in kernel_propagate (:973-998).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define REAL double
#define FN or_screen_run_shaded_f64
#define EXPF exp
#include "oracle_screen_core.inc"
#undef REAL
#undef FN
#undef EXPF

#define REAL float
#define FN or_screen_run_shaded_f32
#define EXPF expf
#include "oracle_screen_core.inc"
#undef REAL
#undef FN
#undef EXPF

/* one tree (root) with flags 0 for every vertex */
int or_screen_run_f64(uint32_t W, uint32_t H, uint32_t S, uint32_t L, int radius, double sigma_pos,
                      double sigma_normal, double sigma_lum, const float* rec, const or_node* nodes, int root,
                      double* film, double* level_out, int threads) {
    return or_screen_run_shaded_f64(W, H, S, L, radius, sigma_pos, sigma_normal, sigma_lum, rec, nodes, root, NULL,
                                    0u, NULL, NULL, film, level_out, threads);
}
int or_screen_run_f32(uint32_t W, uint32_t H, uint32_t S, uint32_t L, int radius, double sigma_pos,
                      double sigma_normal, double sigma_lum, const float* rec, const or_node* nodes, int root,
                      float* film, float* level_out, int threads) {
    return or_screen_run_shaded_f32(W, H, S, L, radius, sigma_pos, sigma_normal, sigma_lum, rec, nodes, root, NULL,
                                    0u, NULL, NULL, film, level_out, threads);
}
