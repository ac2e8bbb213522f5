/*
 * crush_oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * An f64 CPU restatement of the reference solver's scalar path
 * (Settings.simd = false: src/solver.rs, src/kernel.rs, src/element.rs,
 * src/material.rs, src/contact.rs of osullivryan/crushrs).  It is the
 * checker the GPU path is compared against; nothing in the product links,
 * loads or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it.
 *
 * Pinning (what fixes each function independently of this restatement):
 *   oracle_hex_geometry      cube closed form (src/element.rs:203-212),
 *                            rank-12 == dense 8-Gauss K.u and zero on uniform
 *                            gradients (src/element.rs:215-248)
 *   oracle_honeycomb_update  uniaxial closed form (src/kernel.rs:369-386),
 *                            knot values (src/kernel_simd.rs:674-680),
 *                            Hencky-foam equivalence, objectivity,
 *                            idempotence (src/material.rs:961-1023)
 *   oracle_eigen3            numpy.linalg.eigh to 1e-12 (src/material.rs:886-925)
 *   oracle_hencky            pressure K ln J (src/material.rs:878-884),
 *                            objectivity, small-strain limit
 *   oracle_contact_*         brute-force restatement itself; pinned by the
 *                            Newton's-third-law sum and hand-built cases
 *   oracle_run               momentum conservation (tests/impact.rs:38) and
 *                            the crash 30 30 numbers of README.md:33-38
 */
#ifndef CRUSH_ORACLE_H
#define CRUSH_ORACLE_H

#include <stdint.h>
#include "../include/crushrs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* HexGeometry::new (src/element.rs:141-161) + Hourglass::from_dense
 * (src/element.rs:91-114).  x[8*3] node positions; outputs grad[8*3]
 * (dN_i/dX_c at the centre, [3*i+c]), volume, lumped[8], gamma[8*4]
 * ([4*i+a]), H[12*12] row-major in q[3*mode+dir] coordinates.  Returns 0,
 * or -1 for a degenerate hexahedron. */
int oracle_hex_geometry(const double* x, double youngs, double poisson, double density,
                        double hourglass, double* grad, double* volume, double* lumped,
                        double* gamma, double* H);

/* Dense hourglass stiffness K (24x24 row-major) of src/element.rs:146-154. */
int oracle_dense_hourglass_k(const double* x, double youngs, double poisson,
                             double hourglass, double* K);

/* HoneycombParams::from_material knot table (src/material.rs:430-449):
 * knots[8*3] = [c, sigma, slope]; returns 0 or -1 if the curve is invalid. */
int oracle_honeycomb_params(const crush_material* m, double* knots, double* cap_factor);

/* HoneycombParams::yield_at (src/material.rs:459-467). */
void oracle_knot_yield_at(const double* knots, double c, double* sy, double* h);

/* Plasticity::yield_at (src/material.rs:83-102). */
void oracle_plasticity_yield_at(const crush_material* m, double c, double* sy, double* h);

/* honeycomb_rate_update (src/material.rs:472-535).
 * state[25] = f_prev_inv[9] (row-major), rotation[9], stress[6], compaction;
 * updated in place.  tau[9] Kirchhoff stress, row-major. */
void oracle_honeycomb_update(double youngs, const double* knots, const double* cap_factor,
                             const double* F, double* state, double* tau);

/* symmetric_eigen3 (src/material.rs:570-659): vals ascending, vecs[9]
 * row-major with eigenvectors as columns. */
void oracle_eigen3(const double* A, double* vals, double* vecs);

/* finite_strain_return, elastic (src/material.rs:708-771 with
 * plasticity = None and C_p^-1 = I). */
void oracle_hencky_elastic(double youngs, double poisson, const double* F, double* tau);

/* characteristic_length_and_volume (src/element.rs:176-186). */
void oracle_char_length(const double* p, double* length, double* volume);

/* Elements::element_force (src/kernel.rs:116-156) for one element given the
 * 8 current positions x[24] and displacements u[24].  kind: CRUSH_MAT_*.
 * Returns 1 (ok) or 0 (inverted: out zeroed). */
int oracle_element_force(const double* x, const double* u, const double* grad, double volume,
                         const double* gamma, const double* H, int kind, double youngs,
                         double poisson, const double* knots, const double* cap_factor,
                         double* state, double* out);

/* The whole solve (src/solver.rs:190-381, simd = false). */
int oracle_run(const crush_model* model, crush_results* out);

/* The first n_steps steps with per-step taps; same layout as
 * crush_debug_run_steps in include/crushrs.h. */
int oracle_run_steps(const crush_model* model, int32_t n_steps, double* f_int, double* f_ext,
                     int32_t* hit, uint8_t* eroded, double* dt, int32_t* steps_done);

/* The whole solve with the reference's default element path
 * (Settings.simd = true): honeycomb elements through the f32 AVX2 lane
 * blocks of crush_avx2.h, the rest scalar f64.  threads > 1 spreads the
 * element pass over OpenMP threads (results do not depend on it).  This is
 * the timed CPU baseline. */
int oracle_run_lanes(const crush_model* model, int32_t threads, crush_results* out);

/* oracle_run_steps on the lane path. */
int oracle_run_steps_lanes(const crush_model* model, int32_t n_steps, int32_t threads, double* f_int,
                           double* f_ext, int32_t* hit, uint8_t* eroded, double* dt,
                           int32_t* steps_done);

/* signal.rs restated (src/signal.rs:9-76): the CFC filter (out may not
 * alias x), linear resampling (returns the sample count, writes at most
 * cap) and the cumulative trapezoid. */
void oracle_cfc_filter(const double* x, int32_t n, double dt, double cfc, double* out);
int32_t oracle_resample(const double* t, const double* y, int32_t n, double t0, double dt, double t_end, double* out,
                        int32_t cap);
void oracle_integrate(const double* t, const double* y, int32_t n, double y0, double* out);

/* Test hook: when set, oracle_run / oracle_run_steps (the scalar path)
 * take the force of every element whose material kind k has bit (1 << k)
 * in kinds_mask from fn(ctx, element, kind, u[24] of its nodes, out[24]),
 * which returns 1, 0 when the element inverts (out zeroed), or -1 to
 * leave that element to the oracle's own arithmetic.  Lets an
 * alternative element arithmetic run inside the oracle's solve loop.
 * Process-global; NULL clears it. */
typedef int (*oracle_element_hook)(void* ctx, int32_t element, int32_t kind, const double* u24, double* out24);
void oracle_set_element_hook(oracle_element_hook fn, void* ctx, int32_t kinds_mask);

/* Displacement after each of the first n_steps steps, u[s*3N..]; simd
 * selects the lane path. */
int oracle_run_displacements(const crush_model* model, int32_t n_steps, int32_t simd, double* u,
                             int32_t* steps_done);

/* One contact evaluation (ContactRuntime::apply, src/contact.rs:92-160)
 * over the contacts of `model` at displacement u, adding into force[3N];
 * hit[] as in oracle_run_steps.  Returns the number of penetrating nodes. */
int oracle_contact_apply(const crush_model* model, const double* u, const double* masses,
                         double dt, double* force, int32_t* hit);

#ifdef __cplusplus
}
#endif
#endif
