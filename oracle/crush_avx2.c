/*
 * crush_avx2.c — TEST INFRASTRUCTURE AND CPU BASELINE ONLY.
 *
 * The reference's f32 lane kernel for honeycomb hexahedra
 * (src/kernel_simd.rs:300-491, block_forces_impl) restated with AVX2
 * intrinsics: eight elements per block, one 256-bit register per element
 * quantity.  Global positions, displacements and forces stay f64; the
 * element arithmetic and the material state are f32.
 *
 * Rounding follows the reference: its `fma(a, b, c)` helper is a multiply
 * followed by an add (Rust never contracts them), so this file is built with
 * -ffp-contract=off and writes every product-sum as two roundings.
 */
#include "crush_avx2.h"

#include <immintrin.h>
#include <stdlib.h>
#include <string.h>

#define LANES 8
#define NKNOT 8

typedef __m256 V;

typedef struct {
    float grad[24][LANES];
    float volume[LANES];
    float gamma[32][LANES];
    float h[144][LANES]; /* h[12*col + row] */
    float youngs[LANES];
    float trans[LANES];
    float knot_c[NKNOT][LANES];
    float knot_s[NKNOT][LANES];
    float knot_h[NKNOT][LANES];
    float f_prev_inv[9][LANES];
    float rotation[9][LANES];
    float stress[6][LANES];
    float compaction[LANES];
    float out[24][LANES]; /* last forces, lane-major per component */
    int32_t conn[LANES][8];
    uint8_t eroded[LANES];
    int count;
} __attribute__((aligned(32))) block_t;

struct lane_blocks {
    int n, nblocks;
    block_t* blocks;
};

static inline V ld(const float* p) { return _mm256_load_ps(p); }
static inline void st(float* p, V v) { _mm256_store_ps(p, v); }
static inline V splat(float x) { return _mm256_set1_ps(x); }
static inline V add(V a, V b) { return _mm256_add_ps(a, b); }
static inline V sub(V a, V b) { return _mm256_sub_ps(a, b); }
static inline V mul(V a, V b) { return _mm256_mul_ps(a, b); }
static inline V dv(V a, V b) { return _mm256_div_ps(a, b); }
/* a*b + c with two roundings (the reference's `fma` helper) */
static inline V mad(V a, V b, V c) { return _mm256_add_ps(_mm256_mul_ps(a, b), c); }
/* lanes of `a` where mask is set, else `b` */
static inline V sel(V mask, V a, V b) { return _mm256_blendv_ps(b, a, mask); }
static inline V vmax(V a, V b) { return _mm256_max_ps(a, b); }
static inline V clampv(V a, V lo, V hi) { return _mm256_min_ps(_mm256_max_ps(a, lo), hi); }

static void mat_mul(const V* a, const V* b, V* r) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            V acc = mul(a[3 * i], b[j]);
            acc = mad(a[3 * i + 1], b[3 + j], acc);
            r[3 * i + j] = mad(a[3 * i + 2], b[6 + j], acc);
        }
}

/* a * b^T */
static void mat_mul_t(const V* a, const V* b, V* r) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            V acc = mul(a[3 * i], b[3 * j]);
            acc = mad(a[3 * i + 1], b[3 * j + 1], acc);
            r[3 * i + j] = mad(a[3 * i + 2], b[3 * j + 2], acc);
        }
}

/* a^T * b */
static void mat_t_mul(const V* a, const V* b, V* r) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            V acc = mul(a[i], b[j]);
            acc = mad(a[3 + i], b[3 + j], acc);
            r[3 * i + j] = mad(a[6 + i], b[6 + j], acc);
        }
}

static V det3(const V* m) {
    V c0 = sub(mul(m[4], m[8]), mul(m[5], m[7]));
    V c1 = sub(mul(m[5], m[6]), mul(m[3], m[8]));
    V c2 = sub(mul(m[3], m[7]), mul(m[4], m[6]));
    return mad(m[0], c0, mad(m[1], c1, mul(m[2], c2)));
}

/* cofactor inverse; det nonzero in the lanes that use the result */
static void inv3(const V* m, V det, V* r) {
    V inv_det = dv(splat(1.0f), det);
    r[0] = sub(mul(m[4], m[8]), mul(m[5], m[7]));
    r[1] = sub(mul(m[2], m[7]), mul(m[1], m[8]));
    r[2] = sub(mul(m[1], m[5]), mul(m[2], m[4]));
    r[3] = sub(mul(m[5], m[6]), mul(m[3], m[8]));
    r[4] = sub(mul(m[0], m[8]), mul(m[2], m[6]));
    r[5] = sub(mul(m[2], m[3]), mul(m[0], m[5]));
    r[6] = sub(mul(m[3], m[7]), mul(m[4], m[6]));
    r[7] = sub(mul(m[1], m[6]), mul(m[0], m[7]));
    r[8] = sub(mul(m[0], m[4]), mul(m[1], m[3]));
    for (int k = 0; k < 9; ++k) r[k] = mul(r[k], inv_det);
}

static void identity(V* r) {
    for (int k = 0; k < 9; ++k) r[k] = splat(0.0f);
    r[0] = r[4] = r[8] = splat(1.0f);
}

static void block_empty(block_t* b) {
    memset(b, 0, sizeof *b);
    for (int l = 0; l < LANES; ++l) {
        b->youngs[l] = 1.0f;
        b->trans[l] = 1.0f;
        for (int i = 0; i < NKNOT; ++i) b->knot_c[i][l] = __builtin_inff();
        b->f_prev_inv[0][l] = b->f_prev_inv[4][l] = b->f_prev_inv[8][l] = 1.0f;
        b->rotation[0][l] = b->rotation[4][l] = b->rotation[8][l] = 1.0f;
        b->eroded[l] = 1;
    }
}

/* block_forces_impl: src/kernel_simd.rs:300-491. */
static int block_forces(block_t* b, const double* pos, const double* u) {
    float xs[24][LANES] __attribute__((aligned(32)));
    float us[24][LANES] __attribute__((aligned(32)));
    memset(xs, 0, sizeof xs);
    memset(us, 0, sizeof us);
    for (int l = 0; l < b->count; ++l)
        for (int i = 0; i < 8; ++i) {
            size_t n = (size_t)b->conn[l][i];
            for (int c = 0; c < 3; ++c) {
                xs[3 * i + c][l] = (float)(pos[3 * n + c] + u[3 * n + c]);
                us[3 * i + c][l] = (float)u[3 * n + c];
            }
        }
    /* F = sum_i x_i (x) gradN_i */
    V f[9];
    for (int k = 0; k < 9; ++k) f[k] = splat(0.0f);
    for (int i = 0; i < 8; ++i)
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) f[3 * a + c] = mad(ld(xs[3 * i + a]), ld(b->grad[3 * i + c]), f[3 * a + c]);
    V det = det3(f);
    float detl[LANES] __attribute__((aligned(32)));
    float actf[LANES] __attribute__((aligned(32)));
    st(detl, det);
    int newly = 0;
    for (int l = 0; l < LANES; ++l) {
        if (l < b->count && !b->eroded[l] && detl[l] <= 0.0f) {
            b->eroded[l] = 1;
            ++newly;
        }
        /* all-ones lane mask for active elements */
        uint32_t bits = b->eroded[l] ? 0u : 0xffffffffu;
        memcpy(&actf[l], &bits, 4);
    }
    V active = ld(actf);
    V I[9];
    identity(I);
    V one = splat(1.0f), zero = splat(0.0f);
    V det_safe = sel(active, det, one);
    V fs[9];
    for (int k = 0; k < 9; ++k) fs[k] = sel(active, f[k], I[k]);

    /* honeycomb rate update */
    V fpi[9], rot[9];
    for (int k = 0; k < 9; ++k) {
        fpi[k] = ld(b->f_prev_inv[k]);
        rot[k] = ld(b->rotation[k]);
    }
    V f_inc[9], d_inc[9], mid[9], mid_inv[9], g[9];
    mat_mul(fs, fpi, f_inc);
    for (int k = 0; k < 9; ++k) {
        d_inc[k] = sub(f_inc[k], I[k]);
        mid[k] = mad(d_inc[k], splat(0.5f), I[k]);
    }
    inv3(mid, det3(mid), mid_inv);
    mat_mul(d_inc, mid_inv, g);
    V d[9], w[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            d[3 * i + j] = mul(add(g[3 * i + j], g[3 * j + i]), splat(0.5f));
            w[3 * i + j] = mul(sub(g[3 * i + j], g[3 * j + i]), splat(0.5f));
        }
    /* Cayley rotation increment */
    V omega2 = mad(w[1], w[1], mad(w[2], w[2], mul(w[5], w[5])));
    V inv_den = dv(one, mad(omega2, splat(0.25f), one));
    V ww[9], dr[9], r[9];
    mat_mul(w, w, ww);
    for (int k = 0; k < 9; ++k) dr[k] = mad(mad(ww[k], splat(0.5f), w[k]), inv_den, I[k]);
    mat_mul(dr, rot, r);
    V rtd[9], de[9];
    mat_t_mul(r, d, rtd);
    mat_mul(rtd, r, de);
    V e = ld(b->youngs);
    V s[6];
    for (int k = 0; k < 6; ++k) s[k] = ld(b->stress[k]);
    s[0] = mad(e, de[0], s[0]);
    s[1] = mad(e, de[4], s[1]);
    s[2] = mad(e, de[8], s[2]);
    s[3] = mad(e, de[1], s[3]);
    s[4] = mad(e, de[5], s[4]);
    s[5] = mad(e, de[2], s[5]);
    /* yield stress and slope from the knot table */
    V comp = ld(b->compaction);
    V kc = ld(b->knot_c[0]), ks = ld(b->knot_s[0]), kh = ld(b->knot_h[0]);
    for (int i = 1; i < NKNOT; ++i) {
        V c_i = ld(b->knot_c[i]);
        V m = _mm256_cmp_ps(comp, c_i, _CMP_GE_OQ);
        kc = sel(m, c_i, kc);
        ks = sel(m, ld(b->knot_s[i]), ks);
        kh = sel(m, ld(b->knot_h[i]), kh);
    }
    V hard = kh;
    V sigma_yc = mad(hard, sub(comp, kc), ks);
    /* three fixed passes of the capped active set with implicit hardening */
    V tk[3] = {one, ld(b->trans), ld(b->trans)};
    V dc = zero;
    for (int pass = 0; pass < 3; ++pass) {
        V sigma_y = mad(hard, dc, sigma_yc);
        V sum = zero, t_active = zero;
        for (int k = 0; k < 3; ++k) {
            V a = mul(s[k], splat(-1.0f));
            V cap = mul(tk[k], sigma_y);
            V over = _mm256_cmp_ps(a, cap, _CMP_GT_OQ);
            sum = add(sum, sel(over, sub(a, mul(tk[k], sigma_yc)), zero));
            t_active = add(t_active, sel(over, tk[k], zero));
        }
        V den = mad(t_active, hard, e);
        V dc_new = dv(sum, den);
        V any = _mm256_cmp_ps(t_active, zero, _CMP_GT_OQ);
        dc = sel(any, dc_new, zero);
    }
    V sigma_y = vmax(mad(hard, dc, sigma_yc), zero);
    for (int k = 0; k < 3; ++k) {
        V cap = mul(tk[k], sigma_y);
        s[k] = clampv(s[k], mul(cap, splat(-1.0f)), cap);
    }
    V tau_y = mul(sigma_y, splat(0.5f));
    V neg_tau_y = mul(tau_y, splat(-1.0f));
    for (int k = 3; k < 6; ++k) s[k] = clampv(s[k], neg_tau_y, tau_y);
    /* tau = J R sigma_hat R^T */
    V sh[9] = {s[0], s[3], s[5], s[3], s[1], s[4], s[5], s[4], s[2]};
    V rs[9], tau[9];
    mat_mul(r, sh, rs);
    mat_mul_t(rs, r, tau);
    for (int k = 0; k < 9; ++k) tau[k] = mul(tau[k], det_safe);
    /* commit the state of active lanes */
    V f_inv[9];
    inv3(fs, det_safe, f_inv);
    for (int k = 0; k < 9; ++k) {
        st(b->f_prev_inv[k], sel(active, f_inv[k], fpi[k]));
        st(b->rotation[k], sel(active, r[k], rot[k]));
    }
    for (int k = 0; k < 6; ++k) st(b->stress[k], sel(active, s[k], ld(b->stress[k])));
    st(b->compaction, sel(active, add(comp, dc), comp));

    /* nodal forces f_i = tau F^-T gradN_i V0 */
    V mm[9];
    mat_mul_t(tau, f_inv, mm);
    V vol = ld(b->volume);
    for (int k = 0; k < 9; ++k) mm[k] = mul(mm[k], vol);
    V fo[24];
    for (int i = 0; i < 8; ++i)
        for (int a = 0; a < 3; ++a) {
            V acc = mul(mm[3 * a], ld(b->grad[3 * i]));
            acc = mad(mm[3 * a + 1], ld(b->grad[3 * i + 1]), acc);
            fo[3 * i + a] = mad(mm[3 * a + 2], ld(b->grad[3 * i + 2]), acc);
        }
    /* hourglass: q = U Gamma, rr = H q, f += rr Gamma^T */
    V q[12];
    for (int mode = 0; mode < 4; ++mode)
        for (int dir = 0; dir < 3; ++dir) {
            V acc = zero;
            for (int i = 0; i < 8; ++i) acc = mad(ld(us[3 * i + dir]), ld(b->gamma[4 * i + mode]), acc);
            q[3 * mode + dir] = acc;
        }
    V rr[12];
    for (int k = 0; k < 12; ++k) rr[k] = zero;
    for (int col = 0; col < 12; ++col)
        for (int row = 0; row < 12; ++row) rr[row] = mad(ld(b->h[12 * col + row]), q[col], rr[row]);
    for (int i = 0; i < 8; ++i)
        for (int dir = 0; dir < 3; ++dir) {
            V acc = fo[3 * i + dir];
            for (int mode = 0; mode < 4; ++mode) acc = mad(rr[3 * mode + dir], ld(b->gamma[4 * i + mode]), acc);
            st(b->out[3 * i + dir], sel(active, acc, zero));
        }
    return newly;
}

lane_blocks* lane_blocks_create(int n) {
    lane_blocks* b = (lane_blocks*)calloc(1, sizeof *b);
    if (!b) return NULL;
    b->n = n;
    b->nblocks = (n + LANES - 1) / LANES;
    size_t bytes = sizeof(block_t) * (size_t)(b->nblocks > 0 ? b->nblocks : 1);
    b->blocks = (block_t*)aligned_alloc(64, (bytes + 63) / 64 * 64);
    if (!b->blocks) {
        free(b);
        return NULL;
    }
    for (int k = 0; k < b->nblocks; ++k) block_empty(&b->blocks[k]);
    return b;
}

void lane_blocks_free(lane_blocks* b) {
    if (!b) return;
    free(b->blocks);
    free(b);
}

void lane_blocks_set(lane_blocks* b, int k, const int32_t* conn, const double* grad, double volume,
                     const double* gamma, const double* H, double youngs, const double* knots,
                     const double* cap, const double* state, int eroded) {
    block_t* blk = &b->blocks[k / LANES];
    int l = k % LANES;
    for (int i = 0; i < 8; ++i) blk->conn[l][i] = conn[i];
    for (int j = 0; j < 24; ++j) blk->grad[j][l] = (float)grad[j];
    blk->volume[l] = (float)volume;
    for (int j = 0; j < 32; ++j) blk->gamma[j][l] = (float)gamma[j];
    for (int col = 0; col < 12; ++col)
        for (int row = 0; row < 12; ++row) blk->h[12 * col + row][l] = (float)H[12 * row + col];
    blk->youngs[l] = (float)youngs;
    blk->trans[l] = (float)cap[1];
    for (int i = 0; i < NKNOT; ++i) {
        blk->knot_c[i][l] = (float)knots[3 * i];
        blk->knot_s[i][l] = (float)knots[3 * i + 1];
        blk->knot_h[i][l] = (float)knots[3 * i + 2];
    }
    for (int j = 0; j < 9; ++j) {
        blk->f_prev_inv[j][l] = (float)state[j];
        blk->rotation[j][l] = (float)state[9 + j];
    }
    for (int j = 0; j < 6; ++j) blk->stress[j][l] = (float)state[18 + j];
    blk->compaction[l] = (float)state[24];
    blk->eroded[l] = (uint8_t)(eroded != 0);
    if (blk->count < l + 1) blk->count = l + 1;
}

int lane_blocks_forces(lane_blocks* b, const double* positions, const double* u, double* global,
                       int threads) {
    int newly = 0;
    int nb = b->nblocks;
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads > 1 ? threads : 1) if (threads > 1) \
    reduction(+ : newly)
    for (int k = 0; k < nb; ++k) newly += block_forces(&b->blocks[k], positions, u);
    for (int k = 0; k < nb; ++k) {
        const block_t* blk = &b->blocks[k];
        for (int l = 0; l < blk->count; ++l)
            for (int i = 0; i < 8; ++i) {
                size_t n = (size_t)blk->conn[l][i];
                global[3 * n] += (double)blk->out[3 * i][l];
                global[3 * n + 1] += (double)blk->out[3 * i + 1][l];
                global[3 * n + 2] += (double)blk->out[3 * i + 2][l];
            }
    }
    return newly;
}

int lane_blocks_eroded(const lane_blocks* b, int k) { return b->blocks[k / LANES].eroded[k % LANES]; }

void lane_blocks_state(const lane_blocks* b, int k, double* s) {
    const block_t* blk = &b->blocks[k / LANES];
    int l = k % LANES;
    for (int j = 0; j < 9; ++j) {
        s[j] = blk->f_prev_inv[j][l];
        s[9 + j] = blk->rotation[j][l];
    }
    for (int j = 0; j < 6; ++j) s[18 + j] = blk->stress[j][l];
    s[24] = blk->compaction[l];
}

void lane_blocks_out(const lane_blocks* b, int k, float* out) {
    const block_t* blk = &b->blocks[k / LANES];
    for (int j = 0; j < 24; ++j) out[j] = blk->out[j][k % LANES];
}
