/*
 * crush_avx2.h — TEST INFRASTRUCTURE AND CPU BASELINE ONLY.
 *
 * A C restatement of the reference solver's default element path
 * (Settings.simd = true): honeycomb elements packed 8 to a block, every
 * element quantity held as eight f32 lanes and computed with AVX2
 * (src/kernel_simd.rs:180-596 of osullivryan/crushrs).  It is used as the
 * timed CPU baseline next to the GPU path and is checked against the f64
 * oracle; nothing in the product links, loads or calls it.
 *
 * Pinning: the reference's own lane-vs-scalar test (a crushing element
 * stepped 200 times must track the f64 update to single precision,
 * src/kernel_simd.rs:605-680) and whole solves against the f64 oracle.
 */
#ifndef CRUSH_AVX2_H
#define CRUSH_AVX2_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lane_blocks lane_blocks;

/* Room for n elements (Blocks::build, src/kernel_simd.rs:523-546). */
lane_blocks* lane_blocks_create(int n);
void lane_blocks_free(lane_blocks* b);

/* Block::set_lane (src/kernel_simd.rs:237-283) for the k-th honeycomb
 * element, k = 0, 1, ... in order: conn[8] node ids, grad[24] ([3*i+c]),
 * volume, gamma[32] ([4*i+a]), H[144] (symmetric, 12x12), youngs, knots[24]
 * ([c, sigma, slope] x 8, unused at +inf), cap[3], state[25] (F_prev^-1,
 * R row-major, stress[6], compaction) and the erosion flag. */
void lane_blocks_set(lane_blocks* b, int k, const int32_t* conn, const double* grad, double volume,
                     const double* gamma, const double* H, double youngs, const double* knots,
                     const double* cap, const double* state, int eroded);

/* Blocks::compute_forces + Blocks::scatter (src/kernel_simd.rs:549-583):
 * forces of every block at displacement u (positions X + u), added into
 * global[3N] in block and lane order.  threads > 1 splits the blocks over
 * OpenMP threads (chunks of 4 blocks); the scatter stays sequential, so the
 * sum does not depend on the thread count.  Returns the number of elements
 * that inverted in this call. */
int lane_blocks_forces(lane_blocks* b, const double* positions, const double* u, double* global,
                       int threads);

/* Erosion flag and f64 copy of the state of the k-th element
 * (Blocks::mirror_erosion / Block::lane_state). */
int lane_blocks_eroded(const lane_blocks* b, int k);
void lane_blocks_state(const lane_blocks* b, int k, double* state25);

/* Last forces of the k-th element (f32 out[24], as computed). */
void lane_blocks_out(const lane_blocks* b, int k, float* out24);

#ifdef __cplusplus
}
#endif
#endif
