/*
 * crush_oracle.c — TEST INFRASTRUCTURE ONLY (see crush_oracle.h).
 *
 * f64 restatement of the reference's scalar solver path, written from the
 * Rust sources of osullivryan/crushrs.  Every function cites the file:line
 * it follows.  3x3 matrices are row-major double[9]; m[3*i+j] = m(i,j).
 */
#define _POSIX_C_SOURCE 199309L
#include "crush_oracle.h"
#include "crush_avx2.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---------------------------------------------------------------- tables */

/* C3D8 parametric node coordinates: src/element.rs:9-18. */
static const double XI[8][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
    {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

/* 1/sqrt(3): src/element.rs:20. */
static const double GAUSS = 0.5773502691896258;

/* Flanagan-Belytschko hourglass base vectors: src/element.rs:83-88. */
static const double MODES[4][8] = {
    {1, 1, -1, -1, -1, -1, 1, 1},
    {1, -1, -1, 1, -1, 1, 1, -1},
    {1, -1, 1, -1, 1, -1, 1, -1},
    {-1, 1, -1, 1, 1, -1, 1, -1}};

/* Contact constants: src/contact.rs:75,82,87 and the triangle tolerance
 * of :34. */
static const double MAX_ACCEL = 1.0e5;
static const double SOFT = 0.1;
static const double TRI_TOL = 0.1;

/* ------------------------------------------------------------ 3x3 algebra */

static double det3(const double* m) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

/* Cofactor inverse; returns 0 when the determinant is exactly zero
 * (nalgebra's try_inverse returns None then). */
static int inv3(const double* m, double* r) {
    double c00 = m[4] * m[8] - m[5] * m[7];
    double c01 = m[5] * m[6] - m[3] * m[8];
    double c02 = m[3] * m[7] - m[4] * m[6];
    double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    if (det == 0.0) return 0;
    double t[9];
    t[0] = c00 / det;
    t[1] = (m[2] * m[7] - m[1] * m[8]) / det;
    t[2] = (m[1] * m[5] - m[2] * m[4]) / det;
    t[3] = c01 / det;
    t[4] = (m[0] * m[8] - m[2] * m[6]) / det;
    t[5] = (m[2] * m[3] - m[0] * m[5]) / det;
    t[6] = c02 / det;
    t[7] = (m[1] * m[6] - m[0] * m[7]) / det;
    t[8] = (m[0] * m[4] - m[1] * m[3]) / det;
    memcpy(r, t, sizeof t);
    return 1;
}

static void mul3(const double* a, const double* b, double* r) {
    double t[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            t[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
    memcpy(r, t, sizeof t);
}

static void transpose3(const double* a, double* r) {
    double t[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) t[3 * i + j] = a[3 * j + i];
    memcpy(r, t, sizeof t);
}

static void ident3(double* r) {
    memset(r, 0, 9 * sizeof(double));
    r[0] = r[4] = r[8] = 1.0;
}

/* General inverse by Gauss-Jordan with partial pivoting (n <= 24). */
static int invert_n(const double* a, double* r, int n) {
    double w[24 * 48];
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            w[i * 2 * n + j] = a[i * n + j];
            w[i * 2 * n + n + j] = (i == j) ? 1.0 : 0.0;
        }
    }
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int i = c + 1; i < n; ++i)
            if (fabs(w[i * 2 * n + c]) > fabs(w[p * 2 * n + c])) p = i;
        if (w[p * 2 * n + c] == 0.0) return 0;
        if (p != c)
            for (int j = 0; j < 2 * n; ++j) {
                double t = w[c * 2 * n + j];
                w[c * 2 * n + j] = w[p * 2 * n + j];
                w[p * 2 * n + j] = t;
            }
        double piv = w[c * 2 * n + c];
        for (int j = 0; j < 2 * n; ++j) w[c * 2 * n + j] /= piv;
        for (int i = 0; i < n; ++i) {
            if (i == c) continue;
            double f = w[i * 2 * n + c];
            if (f == 0.0) continue;
            for (int j = 0; j < 2 * n; ++j) w[i * 2 * n + j] -= f * w[c * 2 * n + j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) r[i * n + j] = w[i * 2 * n + n + j];
    return 1;
}

/* r (m x p) = a (m x n) * b (n x p), row-major. */
static void matmul(const double* a, const double* b, double* r, int m, int n, int p) {
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < p; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += a[i * n + k] * b[k * p + j];
            r[i * p + j] = s;
        }
}

/* ------------------------------------------------------------- geometry */

/* shape_derivatives: src/element.rs:23-32. */
static void shape_derivatives(double xi, double eta, double zeta, double dn[8][3]) {
    for (int i = 0; i < 8; ++i) {
        const double* s = XI[i];
        dn[i][0] = 0.125 * s[0] * (1.0 + s[1] * eta) * (1.0 + s[2] * zeta);
        dn[i][1] = 0.125 * s[1] * (1.0 + s[0] * xi) * (1.0 + s[2] * zeta);
        dn[i][2] = 0.125 * s[2] * (1.0 + s[0] * xi) * (1.0 + s[1] * eta);
    }
}

/* shape_functions: src/element.rs:35-40. */
static void shape_functions(double xi, double eta, double zeta, double n[8]) {
    for (int i = 0; i < 8; ++i) {
        const double* s = XI[i];
        n[i] = 0.125 * (1.0 + s[0] * xi) * (1.0 + s[1] * eta) * (1.0 + s[2] * zeta);
    }
}

/* reference_gradients: src/element.rs:43-47.  J = dN^T x; grad = dN J^-T. */
static int reference_gradients(const double* x, double dn[8][3], double g[8][3], double* det_j) {
    double j[9] = {0}, ji[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double s = 0.0;
            for (int i = 0; i < 8; ++i) s += dn[i][a] * x[3 * i + b];
            j[3 * a + b] = s;
        }
    if (!inv3(j, ji)) return -1;
    for (int i = 0; i < 8; ++i)
        for (int c = 0; c < 3; ++c)
            g[i][c] = dn[i][0] * ji[3 * c + 0] + dn[i][1] * ji[3 * c + 1] + dn[i][2] * ji[3 * c + 2];
    *det_j = det3(j);
    return 0;
}

/* b_matrix: src/element.rs:50-65 (6x24, engineering shear). */
static void b_matrix(double g[8][3], double* b) {
    memset(b, 0, 6 * 24 * sizeof(double));
    for (int i = 0; i < 8; ++i) {
        double n0 = g[i][0], n1 = g[i][1], n2 = g[i][2];
        b[0 * 24 + 3 * i] = n0;
        b[1 * 24 + 3 * i + 1] = n1;
        b[2 * 24 + 3 * i + 2] = n2;
        b[3 * 24 + 3 * i] = n1;
        b[3 * 24 + 3 * i + 1] = n0;
        b[4 * 24 + 3 * i + 1] = n2;
        b[4 * 24 + 3 * i + 2] = n1;
        b[5 * 24 + 3 * i] = n2;
        b[5 * 24 + 3 * i + 2] = n0;
    }
}

/* elastic_matrix: src/material.rs:662-673. */
static void elastic_matrix(double e, double nu, double* c) {
    double lam = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double mu = e / (2.0 * (1.0 + nu));
    memset(c, 0, 36 * sizeof(double));
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) c[i * 6 + j] = (i == j) ? lam + 2.0 * mu : lam;
        c[(i + 3) * 6 + i + 3] = mu;
    }
}

/* The 8-point Gauss loop of HexGeometry::new: src/element.rs:141-161.
 * Fills center grad, K (optional), volume, consistent-mass row sums. */
static int hex_integrate(const double* x, double youngs, double poisson, double density,
                         double hourglass, double gc[8][3], double* K, double* volume,
                         double* lumped) {
    double dn[8][3];
    double det;
    shape_derivatives(0.0, 0.0, 0.0, dn);
    if (reference_gradients(x, dn, gc, &det)) return -1;
    double cs[36];
    elastic_matrix(hourglass * youngs, poisson, cs);
    double bc[6 * 24];
    b_matrix(gc, bc);
    if (K) memset(K, 0, 24 * 24 * sizeof(double));
    double vol = 0.0;
    double mass[64];
    memset(mass, 0, sizeof mass);
    for (int q = 0; q < 8; ++q) {
        double xi = XI[q][0] * GAUSS, eta = XI[q][1] * GAUSS, zeta = XI[q][2] * GAUSS;
        double d[8][3], g[8][3], dj;
        shape_derivatives(xi, eta, zeta, d);
        if (reference_gradients(x, d, g, &dj)) return -1;
        if (K) {
            double db[6 * 24], cdb[6 * 24];
            b_matrix(g, db);
            for (int k = 0; k < 6 * 24; ++k) db[k] -= bc[k];
            matmul(cs, db, cdb, 6, 6, 24);
            for (int i = 0; i < 24; ++i)
                for (int j = 0; j < 24; ++j) {
                    double s = 0.0;
                    for (int k = 0; k < 6; ++k) s += db[k * 24 + i] * cdb[k * 24 + j];
                    K[i * 24 + j] += s * dj;
                }
        }
        vol += dj;
        double n[8];
        shape_functions(xi, eta, zeta, n);
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 8; ++j) mass[i * 8 + j] += density * n[i] * n[j] * dj;
    }
    *volume = vol;
    for (int i = 0; i < 8; ++i) {
        double s = 0.0;
        for (int j = 0; j < 8; ++j) s += mass[i * 8 + j];
        lumped[i] = s;
    }
    return 0;
}

int oracle_dense_hourglass_k(const double* x, double youngs, double poisson, double hourglass,
                             double* K) {
    double gc[8][3], vol, lumped[8];
    return hex_integrate(x, youngs, poisson, 1.0, hourglass, gc, K, &vol, lumped);
}

int oracle_hex_geometry(const double* x, double youngs, double poisson, double density,
                        double hourglass, double* grad, double* volume, double* lumped,
                        double* gamma, double* H) {
    double gc[8][3];
    double K[24 * 24];
    if (hex_integrate(x, youngs, poisson, density, hourglass, gc, K, volume, lumped)) return -1;
    for (int i = 0; i < 8; ++i)
        for (int c = 0; c < 3; ++c) grad[3 * i + c] = gc[i][c];
    /* Hourglass::from_dense: src/element.rs:91-114. */
    for (int a = 0; a < 4; ++a) {
        double hx[3] = {0, 0, 0};
        for (int j = 0; j < 8; ++j)
            for (int d = 0; d < 3; ++d) hx[d] += MODES[a][j] * x[3 * j + d];
        for (int i = 0; i < 8; ++i)
            gamma[4 * i + a] = MODES[a][i] - (hx[0] * gc[i][0] + hx[1] * gc[i][1] + hx[2] * gc[i][2]);
    }
    double G[12 * 24];
    memset(G, 0, sizeof G);
    for (int a = 0; a < 4; ++a)
        for (int i = 0; i < 8; ++i)
            for (int d = 0; d < 3; ++d) G[(3 * a + d) * 24 + 3 * i + d] = gamma[4 * i + a];
    double Gt[24 * 12];
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 24; ++j) Gt[j * 12 + i] = G[i * 24 + j];
    double ggt[144], ggti[144];
    matmul(G, Gt, ggt, 12, 24, 12);
    if (!invert_n(ggt, ggti, 12)) return -1;
    double t1[12 * 24], t2[12 * 24], t3[144];
    matmul(ggti, G, t1, 12, 12, 24);
    matmul(t1, K, t2, 12, 24, 24);
    matmul(t2, Gt, t3, 12, 24, 12);
    matmul(t3, ggti, H, 12, 12, 12);
    return 0;
}

/* characteristic_length_and_volume: src/element.rs:176-186. */
void oracle_char_length(const double* p, double* length, double* volume) {
    static const int TETS[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6},
                                   {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};
    static const int FACES[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                    {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};
    double vol = 0.0;
    for (int t = 0; t < 6; ++t) {
        const double* a = p + 3 * TETS[t][0];
        double u[3], v[3], w[3];
        for (int d = 0; d < 3; ++d) {
            u[d] = p[3 * TETS[t][1] + d] - a[d];
            v[d] = p[3 * TETS[t][2] + d] - a[d];
            w[d] = p[3 * TETS[t][3] + d] - a[d];
        }
        double cx = u[1] * v[2] - u[2] * v[1], cy = u[2] * v[0] - u[0] * v[2],
               cz = u[0] * v[1] - u[1] * v[0];
        vol += (cx * w[0] + cy * w[1] + cz * w[2]) / 6.0;
    }
    double amax = 0.0;
    for (int f = 0; f < 6; ++f) {
        double d1[3], d2[3];
        for (int d = 0; d < 3; ++d) {
            d1[d] = p[3 * FACES[f][2] + d] - p[3 * FACES[f][0] + d];
            d2[d] = p[3 * FACES[f][3] + d] - p[3 * FACES[f][1] + d];
        }
        double cx = d1[1] * d2[2] - d1[2] * d2[1], cy = d1[2] * d2[0] - d1[0] * d2[2],
               cz = d1[0] * d2[1] - d1[1] * d2[0];
        double a = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
        amax = fmax(amax, a);
    }
    *volume = vol;
    *length = (vol <= 0.0 || amax <= 0.0) ? INFINITY : vol / amax;
}

/* -------------------------------------------------------------- materials */

/* Plasticity::check (src/material.rs:116-151) for the honeycomb model, then
 * HoneycombParams::from_material (src/material.rs:430-449). */
int oracle_honeycomb_params(const crush_material* m, double* knots, double* cap) {
    int n = m->n_curve;
    int extra = m->has_densification ? 1 : 0;
    if (m->has_densification) {
        if (m->densification[0] <= 0.0 || m->densification[1] < 0.0) return -1;
        if (n > 0 && m->curve[n - 1][0] >= m->densification[0]) return -1;
    }
    if (n > 0) {
        if (n < 2 || n + extra > CRUSH_MAX_KNOTS) return -1;
        if (m->curve[0][0] != 0.0) return -1;
        for (int i = 1; i < n; ++i)
            if (m->curve[i][0] <= m->curve[i - 1][0]) return -1;
        for (int i = 0; i < n; ++i)
            if (m->curve[i][1] <= 0.0) return -1;
    }
    for (int i = 0; i < CRUSH_MAX_KNOTS; ++i) {
        knots[3 * i] = INFINITY;
        knots[3 * i + 1] = 0.0;
        knots[3 * i + 2] = 0.0;
    }
    int used;
    if (n < 2) {
        knots[0] = 0.0;
        knots[1] = m->yield_stress;
        knots[2] = m->hardening;
        used = 1;
    } else {
        for (int i = 0; i < n; ++i) {
            double h = (i + 1 < n) ? (m->curve[i + 1][1] - m->curve[i][1]) /
                                         (m->curve[i + 1][0] - m->curve[i][0])
                                   : knots[3 * (i - 1) + 2];
            knots[3 * i] = m->curve[i][0];
            knots[3 * i + 1] = m->curve[i][1];
            knots[3 * i + 2] = h;
        }
        used = n;
    }
    if (m->has_densification) {
        const double* last = knots + 3 * (used - 1);
        double cl = m->densification[0], kl = m->densification[1];
        knots[3 * used] = cl;
        knots[3 * used + 1] = last[1] + last[2] * (cl - last[0]);
        knots[3 * used + 2] = last[2] + kl;
    }
    cap[0] = 1.0;
    cap[1] = m->transverse_factor;
    cap[2] = m->transverse_factor;
    return 0;
}

/* HoneycombParams::yield_at: src/material.rs:459-467. */
void oracle_knot_yield_at(const double* knots, double c, double* sy, double* h) {
    const double* k = knots;
    for (int i = 1; i < CRUSH_MAX_KNOTS; ++i)
        if (c >= knots[3 * i]) k = knots + 3 * i;
    *sy = k[1] + k[2] * (c - k[0]);
    *h = k[2];
}

/* Plasticity::yield_at: src/material.rs:83-102. */
void oracle_plasticity_yield_at(const crush_material* m, double c, double* sy_out, double* h_out) {
    double sy, h;
    if (m->n_curve < 2) {
        sy = m->yield_stress + m->hardening * c;
        h = m->hardening;
    } else {
        int i = 0;
        while (i + 2 < m->n_curve && c >= m->curve[i + 1][0]) ++i;
        h = (m->curve[i + 1][1] - m->curve[i][1]) / (m->curve[i + 1][0] - m->curve[i][0]);
        sy = m->curve[i][1] + h * (c - m->curve[i][0]);
    }
    if (m->has_densification && c > m->densification[0]) {
        sy += m->densification[1] * (c - m->densification[0]);
        h += m->densification[1];
    }
    *sy_out = sy;
    *h_out = h;
}

/* honeycomb_rate_update: src/material.rs:472-535. */
void oracle_honeycomb_update(double e, const double* knots, const double* t, const double* F,
                             double* state, double* tau) {
    double* fpi = state;
    double* rot = state + 9;
    double* sig = state + 18;
    double c = state[24];
    double i3[9], f_inc[9], mid[9], mid_inv[9], g[9], dm[9], d[9], w[9];
    ident3(i3);
    mul3(F, fpi, f_inc);
    for (int k = 0; k < 9; ++k) mid[k] = 0.5 * (f_inc[k] + i3[k]);
    if (!inv3(mid, mid_inv)) ident3(mid_inv);
    for (int k = 0; k < 9; ++k) dm[k] = f_inc[k] - i3[k];
    mul3(dm, mid_inv, g);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            d[3 * i + j] = 0.5 * (g[3 * i + j] + g[3 * j + i]);
            w[3 * i + j] = 0.5 * (g[3 * i + j] - g[3 * j + i]);
        }
    /* Cayley transform in closed form (src/material.rs:481-485). */
    double omega2 = w[1] * w[1] + w[2] * w[2] + w[5] * w[5];
    double ww[9], dr[9], r[9];
    mul3(w, w, ww);
    for (int k = 0; k < 9; ++k) dr[k] = i3[k] + (w[k] + 0.5 * ww[k]) / (1.0 + 0.25 * omega2);
    mul3(dr, rot, r);
    double rt[9], tmp[9], de[9];
    transpose3(r, rt);
    mul3(rt, d, tmp);
    mul3(tmp, r, de);
    double s[6];
    for (int k = 0; k < 6; ++k) s[k] = sig[k];
    s[0] += e * de[0];
    s[1] += e * de[4];
    s[2] += e * de[8];
    s[3] += e * de[1];
    s[4] += e * de[5];
    s[5] += e * de[2];
    /* Implicit-hardening caps (src/material.rs:496-519). */
    double syc, hard;
    oracle_knot_yield_at(knots, c, &syc, &hard);
    double dc = 0.0;
    for (int it = 0; it < 3; ++it) {
        double sy = syc + hard * dc;
        double sum = 0.0, ta = 0.0;
        for (int k = 0; k < 3; ++k) {
            if (-s[k] > t[k] * sy) {
                sum += -s[k] - t[k] * syc;
                ta += t[k];
            }
        }
        double dc_new = (ta == 0.0) ? 0.0 : sum / (e + ta * hard);
        if (fabs(dc_new - dc) <= 1e-14 * fmax(dc_new, 1e-300)) {
            dc = dc_new;
            break;
        }
        dc = dc_new;
    }
    double sy = fmax(syc + hard * dc, 0.0);
    for (int k = 0; k < 3; ++k) s[k] = fmin(fmax(s[k], -t[k] * sy), t[k] * sy);
    double ty = 0.5 * sy;
    for (int k = 3; k < 6; ++k) s[k] = fmin(fmax(s[k], -ty), ty);
    double sh[9] = {s[0], s[3], s[5], s[3], s[1], s[4], s[5], s[4], s[2]};
    double rs[9], rsrt[9];
    mul3(r, sh, rs);
    mul3(rs, rt, rsrt);
    double jdet = det3(F);
    for (int k = 0; k < 9; ++k) tau[k] = jdet * rsrt[k];
    double finv[9];
    if (!inv3(F, finv)) ident3(finv);
    memcpy(fpi, finv, sizeof finv);
    memcpy(rot, r, sizeof r);
    for (int k = 0; k < 6; ++k) sig[k] = s[k];
    state[24] = c + dc;
}

/* symmetric_eigen3: src/material.rs:570-659. */
static void cross(const double* a, const double* b, double* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

typedef struct { double val; double vec[3]; int key; } eig_pair;

static void sort_pairs(eig_pair* p) {
    /* stable insertion sort by value */
    for (int i = 1; i < 3; ++i) {
        eig_pair x = p[i];
        int j = i - 1;
        while (j >= 0 && p[j].val > x.val) {
            p[j + 1] = p[j];
            --j;
        }
        p[j + 1] = x;
    }
}

void oracle_eigen3(const double* a, double* vals, double* vecs) {
    double a00 = a[0], a01 = a[1], a02 = a[2], a11 = a[4], a12 = a[5], a22 = a[8];
    double max_abs = fmax(fmax(fmax(fabs(a00), fabs(a01)), fmax(fabs(a02), fabs(a11))),
                          fmax(fabs(a12), fabs(a22)));
    if (max_abs == 0.0) {
        vals[0] = vals[1] = vals[2] = 0.0;
        ident3(vecs);
        return;
    }
    double inv = 1.0 / max_abs;
    a00 *= inv; a01 *= inv; a02 *= inv; a11 *= inv; a12 *= inv; a22 *= inv;
    double norm = a01 * a01 + a02 * a02 + a12 * a12;
    if (norm == 0.0) {
        eig_pair p[3] = {{a00, {0}, 0}, {a11, {0}, 1}, {a22, {0}, 2}};
        sort_pairs(p);
        memset(vecs, 0, 9 * sizeof(double));
        for (int c = 0; c < 3; ++c) {
            vecs[3 * p[c].key + c] = 1.0;
            vals[c] = p[c].val * max_abs;
        }
        return;
    }
    double q = (a00 + a11 + a22) / 3.0;
    double b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
    double p = sqrt((b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * norm) / 6.0);
    double c00 = b11 * b22 - a12 * a12;
    double c01 = a01 * b22 - a12 * a02;
    double c02 = a01 * a12 - b11 * a02;
    double det = (b00 * c00 - a01 * c01 + a02 * c02) / (p * p * p);
    double half_det = fmin(fmax(0.5 * det, -1.0), 1.0);
    double angle = acos(half_det) / 3.0;
    const double TWO_THIRDS_PI = 2.0943951023931955;
    double beta2 = 2.0 * cos(angle);
    double beta0 = 2.0 * cos(angle + TWO_THIRDS_PI);
    double beta1 = -(beta0 + beta2);
    double ev[3] = {q + p * beta0, q + p * beta1, q + p * beta2};
    double am[9] = {a00, a01, a02, a01, a11, a12, a02, a12, a22};
    /* eigenvector of the well-separated root from the largest row cross
     * product */
    double e = (half_det >= 0.0) ? ev[2] : ev[0];
    double r0[3] = {a00 - e, a01, a02}, r1[3] = {a01, a11 - e, a12}, r2[3] = {a02, a12, a22 - e};
    double x01[3], x02[3], x12[3];
    cross(r0, r1, x01);
    cross(r0, r2, x02);
    cross(r1, r2, x12);
    double d0 = dot(x01, x01), d1 = dot(x02, x02), d2 = dot(x12, x12);
    double w[3];
    if (d0 >= d1 && d0 >= d2) {
        double s = sqrt(d0);
        for (int k = 0; k < 3; ++k) w[k] = x01[k] / s;
    } else if (d1 >= d2) {
        double s = sqrt(d1);
        for (int k = 0; k < 3; ++k) w[k] = x02[k] / s;
    } else {
        double s = sqrt(d2);
        for (int k = 0; k < 3; ++k) w[k] = x12[k] / s;
    }
    /* exact Jacobi rotation in the plane orthogonal to w */
    double u[3];
    if (fabs(w[0]) > fabs(w[1])) {
        double s = sqrt(w[0] * w[0] + w[2] * w[2]);
        u[0] = -w[2] / s; u[1] = 0.0; u[2] = w[0] / s;
    } else {
        double s = sqrt(w[1] * w[1] + w[2] * w[2]);
        u[0] = 0.0; u[1] = w[2] / s; u[2] = -w[1] / s;
    }
    double v[3];
    cross(w, u, v);
    double au[3], av[3];
    for (int i = 0; i < 3; ++i) {
        au[i] = am[3 * i] * u[0] + am[3 * i + 1] * u[1] + am[3 * i + 2] * u[2];
        av[i] = am[3 * i] * v[0] + am[3 * i + 1] * v[1] + am[3 * i + 2] * v[2];
    }
    double m00 = dot(u, au), m01 = dot(u, av), m11 = dot(v, av);
    double theta = 0.5 * atan2(2.0 * m01, m00 - m11);
    double sn = sin(theta), cs = cos(theta);
    double p1[3], p2[3];
    for (int k = 0; k < 3; ++k) {
        p1[k] = u[k] * cs + v[k] * sn;
        p2[k] = v[k] * cs - u[k] * sn;
    }
    /* Rayleigh-quotient eigenvalues, sorted ascending */
    eig_pair pr[3];
    const double* vv[3] = {w, p1, p2};
    for (int k = 0; k < 3; ++k) {
        double t[3];
        for (int i = 0; i < 3; ++i) t[i] = am[3 * i] * vv[k][0] + am[3 * i + 1] * vv[k][1] + am[3 * i + 2] * vv[k][2];
        pr[k].val = dot(vv[k], t);
        memcpy(pr[k].vec, vv[k], 3 * sizeof(double));
        pr[k].key = k;
    }
    sort_pairs(pr);
    for (int c = 0; c < 3; ++c) {
        vals[c] = pr[c].val * max_abs;
        for (int i = 0; i < 3; ++i) vecs[3 * i + c] = pr[c].vec[i];
    }
}

/* finite_strain_return, elastic (src/material.rs:708-771). */
void oracle_hencky_elastic(double e, double nu, const double* F, double* tau) {
    double g = e / (2.0 * (1.0 + nu));
    double k = e / (3.0 * (1.0 - 2.0 * nu));
    double ft[9], be[9], bs[9];
    transpose3(F, ft);
    mul3(F, ft, be);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) bs[3 * i + j] = 0.5 * (be[3 * i + j] + be[3 * j + i]);
    double lam[3], vec[9];
    oracle_eigen3(bs, lam, vec);
    double eps[3];
    for (int a = 0; a < 3; ++a) eps[a] = 0.5 * log(fmax(lam[a], 1e-300));
    double vol = eps[0] + eps[1] + eps[2];
    double pressure = k * vol;
    memset(tau, 0, 9 * sizeof(double));
    for (int a = 0; a < 3; ++a) {
        double dev = eps[a] - vol / 3.0;
        double coef = pressure + 2.0 * g * dev;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) tau[3 * i + j] += coef * vec[3 * i + a] * vec[3 * j + a];
    }
}

/* --------------------------------------------------------- element force */

/* Elements::element_force: src/kernel.rs:116-156. */
int oracle_element_force(const double* x, const double* u, const double* grad, double volume,
                         const double* gamma, const double* H, int kind, double youngs,
                         double poisson, const double* knots, const double* cap, double* state,
                         double* out) {
    double F[9] = {0};
    for (int i = 0; i < 8; ++i)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) F[3 * a + b] += x[3 * i + a] * grad[3 * i + b];
    if (det3(F) <= 0.0) {
        memset(out, 0, 24 * sizeof(double));
        return 0;
    }
    double tau[9];
    if (kind == CRUSH_MAT_HONEYCOMB)
        oracle_honeycomb_update(youngs, knots, cap, F, state, tau);
    else
        oracle_hencky_elastic(youngs, poisson, F, tau);
    double finv[9], fit[9], m[9];
    inv3(F, finv);
    transpose3(finv, fit);
    mul3(tau, fit, m);
    for (int k = 0; k < 9; ++k) m[k] *= volume;
    for (int i = 0; i < 8; ++i)
        for (int a = 0; a < 3; ++a)
            out[3 * i + a] = m[3 * a] * grad[3 * i] + m[3 * a + 1] * grad[3 * i + 1] +
                             m[3 * a + 2] * grad[3 * i + 2];
    /* hourglass: q = U Gamma, r = H q, f += R Gamma^T */
    double q[12];
    for (int a = 0; a < 4; ++a)
        for (int d = 0; d < 3; ++d) {
            double s = 0.0;
            for (int i = 0; i < 8; ++i) s += u[3 * i + d] * gamma[4 * i + a];
            q[3 * a + d] = s;
        }
    double r[12];
    for (int row = 0; row < 12; ++row) {
        double s = 0.0;
        for (int col = 0; col < 12; ++col) s += H[row * 12 + col] * q[col];
        r[row] = s;
    }
    for (int i = 0; i < 8; ++i)
        for (int d = 0; d < 3; ++d) {
            double s = 0.0;
            for (int a = 0; a < 4; ++a) s += r[3 * a + d] * gamma[4 * i + a];
            out[3 * i + d] += s;
        }
    return 1;
}

/* ---------------------------------------------------------------- contact */

/* point_in_triangle: src/contact.rs:29-43. */
static int point_in_triangle(const double* p, const double* a, const double* b, const double* c) {
    double v0[3], v1[3], v2[3];
    for (int d = 0; d < 3; ++d) {
        v0[d] = c[d] - a[d];
        v1[d] = b[d] - a[d];
        v2[d] = p[d] - a[d];
    }
    double d00 = dot(v0, v0), d01 = dot(v0, v1), d02 = dot(v0, v2), d11 = dot(v1, v1),
           d12 = dot(v1, v2);
    double inv = 1.0 / (d00 * d11 - d01 * d01);
    double u = (d11 * d02 - d01 * d12) * inv;
    double v = (d00 * d12 - d01 * d02) * inv;
    return u >= -TRI_TOL && v >= -TRI_TOL && u + v <= 1.0 + TRI_TOL;
}

typedef struct {
    int n;            /* kept nodes */
    int* nodes;
    double* scales;
    double depth_cap;
} contact_rt;

/* ContactRuntime::new: src/contact.rs:48-70 (static broad phase). */
static void contact_rt_build(const crush_model* m, const crush_contact* c, contact_rt* rt) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int f = 0; f < c->n_faces; ++f)
        for (int k = 0; k < 4; ++k) {
            const double* p = m->nodes + 3 * m->faces[4 * c->faces[f] + k];
            for (int d = 0; d < 3; ++d) {
                lo[d] = fmin(lo[d], p[d]);
                hi[d] = fmax(hi[d], p[d]);
            }
        }
    for (int d = 0; d < 3; ++d) {
        lo[d] -= c->max_distance;
        hi[d] += c->max_distance;
    }
    rt->nodes = (int*)malloc(sizeof(int) * (c->n_nodes + 1));
    rt->scales = (double*)malloc(sizeof(double) * (c->n_nodes + 1));
    rt->n = 0;
    for (int i = 0; i < c->n_nodes; ++i) {
        const double* p = m->nodes + 3 * c->nodes[i];
        if (p[0] >= lo[0] && p[1] >= lo[1] && p[2] >= lo[2] && p[0] <= hi[0] && p[1] <= hi[1] &&
            p[2] <= hi[2]) {
            rt->nodes[rt->n] = c->nodes[i];
            rt->scales[rt->n] = c->node_scale ? c->node_scale[i] : 1.0;
            rt->n++;
        }
    }
    rt->depth_cap = 0.1 * c->max_distance;
}

static void contact_rt_free(contact_rt* rt) {
    free(rt->nodes);
    free(rt->scales);
}

/* ContactRuntime::apply: src/contact.rs:92-160.  Returns the number of
 * penetrating nodes; hit[k] = face position in c->faces or -1. */
static int contact_apply_one(const crush_model* m, const crush_contact* c, const contact_rt* rt,
                             const double* u, const double* masses, double dt, double* force,
                             int32_t* hit) {
    int nf = c->n_faces;
    double* fx = (double*)malloc(sizeof(double) * (size_t)nf * 24);
    for (int f = 0; f < nf; ++f) {
        double* rec = fx + 24 * f; /* x[12] normal[3] center[3] lo[3] hi[3] */
        const int32_t* q = m->faces + 4 * c->faces[f];
        for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) rec[3 * k + d] = m->nodes[3 * q[k] + d] + u[3 * q[k] + d];
        double e1[3], e2[3], n[3];
        for (int d = 0; d < 3; ++d) {
            e1[d] = rec[3 + d] - rec[d];
            e2[d] = rec[6 + d] - rec[d];
        }
        cross(e1, e2, n);
        double nn = sqrt(dot(n, n));
        for (int d = 0; d < 3; ++d) {
            rec[12 + d] = n[d] / nn;
            rec[15 + d] = (rec[d] + rec[3 + d] + rec[6 + d] + rec[9 + d]) / 4.0;
            double lo = rec[d], hi = rec[d];
            for (int k = 1; k < 4; ++k) {
                lo = fmin(lo, rec[3 * k + d]);
                hi = fmax(hi, rec[3 * k + d]);
            }
            rec[18 + d] = lo - c->max_distance;
            rec[21 + d] = hi + c->max_distance;
        }
    }
    int count = 0;
    for (int k = 0; k < rt->n; ++k) {
        int node = rt->nodes[k];
        double p[3];
        for (int d = 0; d < 3; ++d) p[d] = m->nodes[3 * node + d] + u[3 * node + d];
        if (hit) hit[k] = -1;
        for (int f = 0; f < nf; ++f) {
            const double* rec = fx + 24 * f;
            if (p[0] < rec[18] || p[0] > rec[21] || p[1] < rec[19] || p[1] > rec[22] ||
                p[2] < rec[20] || p[2] > rec[23])
                continue;
            const double* nrm = rec + 12;
            double cp[3] = {rec[15] - p[0], rec[16] - p[1], rec[17] - p[2]};
            double signed_d = -dot(nrm, cp);
            if (signed_d > 0.0 || signed_d < -c->max_distance) continue;
            double pr[3];
            for (int d = 0; d < 3; ++d) pr[d] = p[d] - nrm[d] * signed_d;
            const double *x0 = rec, *x1 = rec + 3, *x2 = rec + 6, *x3 = rec + 9;
            int inside = point_in_triangle(pr, x0, x1, x2) || point_in_triangle(pr, x0, x2, x3) ||
                         point_in_triangle(pr, x1, x2, x3) || point_in_triangle(pr, x1, x3, x0);
            if (!inside) continue;
            double depth = fmin(-signed_d, rt->depth_cap);
            double kk = fmin(c->stiffness * rt->scales[k], SOFT * masses[node] / (dt * dt));
            double mag = fmin(kk * depth + 0.0, masses[node] * MAX_ACCEL);
            double fv[3] = {mag * nrm[0], mag * nrm[1], mag * nrm[2]};
            for (int d = 0; d < 3; ++d) force[3 * node + d] += fv[d];
            const int32_t* q = m->faces + 4 * c->faces[f];
            for (int j = 0; j < 4; ++j)
                for (int d = 0; d < 3; ++d) force[3 * q[j] + d] -= 0.25 * fv[d];
            if (hit) hit[k] = f;
            ++count;
            break;
        }
    }
    free(fx);
    return count;
}

int oracle_contact_apply(const crush_model* m, const double* u, const double* masses, double dt,
                         double* force, int32_t* hit) {
    int total = 0, off = 0;
    for (int c = 0; c < m->n_contacts; ++c) {
        contact_rt rt;
        contact_rt_build(m, &m->contacts[c], &rt);
        total += contact_apply_one(m, &m->contacts[c], &rt, u, masses, dt, force, hit ? hit + off : NULL);
        off += rt.n;
        contact_rt_free(&rt);
    }
    return total;
}

/* ------------------------------------------------------------------ solve */

typedef struct {
    int n, e;
    double* grad;   /* [24e] */
    double* vol;    /* [e] */
    double* gamma;  /* [32e] */
    double* H;      /* [144e] */
    double* lumped; /* [8e] */
    double* dt_factor;
    double* state;  /* [25e] */
    uint8_t* eroded;
    int* kind;
    double* knots;  /* [24 * parts] */
    double* cap;    /* [3 * parts] */
    /* lane path (Settings.simd = true): honeycomb elements in f32 blocks of
     * 8, the other elements scalar f64 (src/kernel.rs:76-90) */
    lane_blocks* lanes;
    int* lane_ids;   /* element of the k-th lane slot */
    int n_lane;
    int* scalar_ids; /* elements on the scalar path, in element order */
    int n_scalar;
    double* scalar_out; /* [24 * n_scalar] */
    int* scalar_ok;
    int threads;
} elements_t;

static double wave_speed(const crush_material* m) {
    double e = m->youngs_modulus, v = m->poisson_ratio;
    double mm = e * (1.0 - v) / ((1.0 + v) * (1.0 - 2.0 * v));
    return sqrt(mm / m->density);
}

static void elements_free(elements_t* el) {
    free(el->grad); free(el->vol); free(el->gamma); free(el->H); free(el->lumped);
    free(el->dt_factor); free(el->state); free(el->eroded); free(el->kind); free(el->knots);
    free(el->cap);
    lane_blocks_free(el->lanes);
    free(el->lane_ids); free(el->scalar_ids); free(el->scalar_out); free(el->scalar_ok);
}

/* Elements::build: src/kernel.rs:53-91 (scalar path). */
static int elements_build(const crush_model* m, elements_t* el) {
    int ne = m->n_hexes;
    el->n = m->n_nodes;
    el->e = ne;
    el->grad = (double*)malloc(sizeof(double) * 24 * (size_t)ne);
    el->vol = (double*)malloc(sizeof(double) * (size_t)ne);
    el->gamma = (double*)malloc(sizeof(double) * 32 * (size_t)ne);
    el->H = (double*)malloc(sizeof(double) * 144 * (size_t)ne);
    el->lumped = (double*)malloc(sizeof(double) * 8 * (size_t)ne);
    el->dt_factor = (double*)malloc(sizeof(double) * (size_t)ne);
    el->state = (double*)malloc(sizeof(double) * 25 * (size_t)ne);
    el->eroded = (uint8_t*)calloc((size_t)ne, 1);
    el->kind = (int*)malloc(sizeof(int) * (size_t)ne);
    el->knots = (double*)malloc(sizeof(double) * 24 * (size_t)(m->n_parts + 1));
    el->cap = (double*)malloc(sizeof(double) * 3 * (size_t)(m->n_parts + 1));
    for (int p = 0; p < m->n_parts; ++p) {
        const crush_material* mat = &m->materials[p];
        if (mat->model == CRUSH_MAT_HONEYCOMB) {
            if (oracle_honeycomb_params(mat, el->knots + 24 * p, el->cap + 3 * p)) return CRUSH_E_INPUT;
        } else if (mat->model != CRUSH_MAT_ELASTIC) {
            return CRUSH_E_UNSUPPORTED;
        }
    }
    for (int e = 0; e < ne; ++e) {
        const crush_material* mat = &m->materials[m->hex_part[e]];
        double x[24];
        for (int i = 0; i < 8; ++i)
            for (int d = 0; d < 3; ++d) x[3 * i + d] = m->nodes[3 * m->hexes[8 * e + i] + d];
        if (oracle_hex_geometry(x, mat->youngs_modulus, mat->poisson_ratio, mat->density,
                                m->settings.hourglass, el->grad + 24 * (size_t)e, el->vol + e,
                                el->lumped + 8 * (size_t)e, el->gamma + 32 * (size_t)e,
                                el->H + 144 * (size_t)e))
            return CRUSH_E_INPUT;
        el->dt_factor[e] = (1.0 - 0.36 * mat->poisson_ratio) / wave_speed(mat);
        el->kind[e] = mat->model;
        double* st = el->state + 25 * (size_t)e;
        memset(st, 0, 25 * sizeof(double));
        st[0] = st[4] = st[8] = 1.0;
        st[9] = st[13] = st[17] = 1.0;
    }
    return CRUSH_OK;
}

/* Elements::build with simd = true (src/kernel.rs:76-90) and Blocks::build
 * (src/kernel_simd.rs:523-546): honeycomb elements, in element order, are
 * packed into lane blocks; every other element stays on the scalar path. */
static int elements_build_lanes(const crush_model* m, elements_t* el, int threads) {
    int ne = el->e;
    el->threads = threads;
    el->lane_ids = (int*)malloc(sizeof(int) * (size_t)(ne + 1));
    el->scalar_ids = (int*)malloc(sizeof(int) * (size_t)(ne + 1));
    for (int e = 0; e < ne; ++e) {
        if (el->kind[e] == CRUSH_MAT_HONEYCOMB)
            el->lane_ids[el->n_lane++] = e;
        else
            el->scalar_ids[el->n_scalar++] = e;
    }
    el->scalar_out = (double*)malloc(sizeof(double) * 24 * (size_t)(el->n_scalar + 1));
    el->scalar_ok = (int*)malloc(sizeof(int) * (size_t)(el->n_scalar + 1));
    el->lanes = lane_blocks_create(el->n_lane);
    if (!el->lanes) return CRUSH_E_OOM;
    for (int k = 0; k < el->n_lane; ++k) {
        int e = el->lane_ids[k];
        int part = m->hex_part[e];
        lane_blocks_set(el->lanes, k, m->hexes + 8 * (size_t)e, el->grad + 24 * (size_t)e, el->vol[e],
                        el->gamma + 32 * (size_t)e, el->H + 144 * (size_t)e,
                        m->materials[part].youngs_modulus, el->knots + 24 * part, el->cap + 3 * part,
                        el->state + 25 * (size_t)e, el->eroded[e]);
    }
    return CRUSH_OK;
}

/* Elements::compute_forces (scalar): src/kernel.rs:160-204.  Returns the
 * number of newly eroded elements. */
static int compute_forces_lanes(const crush_model* m, elements_t* el, const double* u, double* global);

/* Test hook: element forces of the kinds in g_hook_kinds computed by an
 * outside function (e.g. another implementation's element arithmetic), the
 * rest of the solve unchanged.  Process-global; not thread-safe. */
static oracle_element_hook g_hook;
static void* g_hook_ctx;
static int g_hook_kinds;

void oracle_set_element_hook(oracle_element_hook fn, void* ctx, int32_t kinds_mask) {
    g_hook = fn;
    g_hook_ctx = ctx;
    g_hook_kinds = kinds_mask;
}

static int compute_forces(const crush_model* m, elements_t* el, const double* u, double* global) {
    if (el->lanes) return compute_forces_lanes(m, el, u, global);
    int newly = 0;
    for (int e = 0; e < el->e; ++e) {
        if (el->eroded[e]) continue;
        double x[24], ul[24], out[24];
        const int32_t* h = m->hexes + 8 * (size_t)e;
        for (int i = 0; i < 8; ++i)
            for (int d = 0; d < 3; ++d) {
                x[3 * i + d] = m->nodes[3 * h[i] + d] + u[3 * h[i] + d];
                ul[3 * i + d] = u[3 * h[i] + d];
            }
        int part = m->hex_part[e];
        const crush_material* mat = &m->materials[part];
        int ok = (g_hook && (g_hook_kinds & (1 << el->kind[e]))) ? g_hook(g_hook_ctx, e, el->kind[e], ul, out) : -1;
        if (ok < 0)
            ok = oracle_element_force(x, ul, el->grad + 24 * (size_t)e, el->vol[e],
                                            el->gamma + 32 * (size_t)e, el->H + 144 * (size_t)e,
                                            el->kind[e], mat->youngs_modulus, mat->poisson_ratio,
                                            el->knots + 24 * part, el->cap + 3 * part,
                                            el->state + 25 * (size_t)e, out);
        if (!ok) {
            el->eroded[e] = 1;
            ++newly;
            continue;
        }
        for (int i = 0; i < 8; ++i)
            for (int d = 0; d < 3; ++d) global[3 * h[i] + d] += out[3 * i + d];
    }
    return newly;
}

/* Elements::compute_forces with simd = true (src/kernel.rs:160-204): the
 * lane blocks first (forces, scatter, erosion mirrored back), then the
 * scalar elements computed independently (in parallel when threads > 1)
 * and scattered in element order. */
static int compute_forces_lanes(const crush_model* m, elements_t* el, const double* u, double* global) {
    int newly = lane_blocks_forces(el->lanes, m->nodes, u, global, el->threads);
    for (int k = 0; k < el->n_lane; ++k) el->eroded[el->lane_ids[k]] = (uint8_t)lane_blocks_eroded(el->lanes, k);
    int ns = el->n_scalar;
#pragma omp parallel for schedule(dynamic, 32) num_threads(el->threads > 1 ? el->threads : 1) if (el->threads > 1)
    for (int k = 0; k < ns; ++k) {
        int e = el->scalar_ids[k];
        double* out = el->scalar_out + 24 * (size_t)k;
        el->scalar_ok[k] = 1;
        if (el->eroded[e]) {
            memset(out, 0, 24 * sizeof(double));
            continue;
        }
        double x[24], ul[24];
        const int32_t* h = m->hexes + 8 * (size_t)e;
        for (int i = 0; i < 8; ++i)
            for (int d = 0; d < 3; ++d) {
                x[3 * i + d] = m->nodes[3 * h[i] + d] + u[3 * h[i] + d];
                ul[3 * i + d] = u[3 * h[i] + d];
            }
        int part = m->hex_part[e];
        const crush_material* mat = &m->materials[part];
        el->scalar_ok[k] = oracle_element_force(x, ul, el->grad + 24 * (size_t)e, el->vol[e],
                                                el->gamma + 32 * (size_t)e, el->H + 144 * (size_t)e,
                                                el->kind[e], mat->youngs_modulus, mat->poisson_ratio,
                                                el->knots + 24 * part, el->cap + 3 * part,
                                                el->state + 25 * (size_t)e, out);
    }
    for (int k = 0; k < ns; ++k) {
        int e = el->scalar_ids[k];
        if (!el->scalar_ok[k]) {
            el->eroded[e] = 1;
            ++newly;
            continue;
        }
        if (el->eroded[e]) continue;
        const double* out = el->scalar_out + 24 * (size_t)k;
        const int32_t* h = m->hexes + 8 * (size_t)e;
        for (int i = 0; i < 8; ++i)
            for (int d = 0; d < 3; ++d) global[3 * h[i] + d] += out[3 * i + d];
    }
    return newly;
}

/* Elements::element_time_steps: src/kernel.rs:214-237. */
static void element_time_steps(const crush_model* m, const elements_t* el, const double* u,
                               double* dts) {
    for (int e = 0; e < el->e; ++e) {
        if (el->eroded[e]) {
            dts[e] = INFINITY;
            continue;
        }
        double p[24];
        const int32_t* h = m->hexes + 8 * (size_t)e;
        for (int i = 0; i < 8; ++i)
            for (int d = 0; d < 3; ++d) p[3 * i + d] = m->nodes[3 * h[i] + d] + u[3 * h[i] + d];
        double l, vol;
        oracle_char_length(p, &l, &vol);
        dts[e] = (isfinite(l) && vol > 0.0) ? l * el->dt_factor[e] : INFINITY;
    }
}

typedef struct {
    int32_t n_steps;
    double* f_int;
    double* f_ext;
    int32_t* hit;
    uint8_t* eroded;
    double* dt;
    int32_t steps_done;
    double* u;
} taps_t;

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static void group_means(const crush_model* m, const crush_group* g, const double* masses,
                        const double* u, const double* v, double* xm, double* vm) {
    double mt = 0.0, sx[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
    for (int i = 0; i < g->n_nodes; ++i) {
        int n = g->nodes[i];
        mt += masses[n];
        for (int d = 0; d < 3; ++d) {
            sx[d] += masses[n] * (m->nodes[3 * n + d] + u[3 * n + d]);
            sv[d] += masses[n] * v[3 * n + d];
        }
    }
    for (int d = 0; d < 3; ++d) {
        xm[d] = mt > 0 ? sx[d] / mt : 0.0;
        vm[d] = mt > 0 ? sv[d] / mt : 0.0;
    }
}

static double group_crush(const crush_group* g, const double* u) {
    if (g->accel_node < 0 || g->n_front <= 0) return 0.0;
    double s = 0.0;
    for (int i = 0; i < g->n_front; ++i) s += u[3 * g->front[i] + g->axis];
    return fabs(u[3 * g->accel_node + g->axis] - s / g->n_front);
}

/* ---------------------------------------------------------------- signal */

/* cfc_filter (src/signal.rs:9-39). */
static void cfc_pass(const double* x, int n, double dt, double cfc, const double* a, const double* b, double* out) {
    int pad = (int)ceil(3.0 / (cfc * dt));
    if (pad > n - 1) pad = n - 1;
    if (pad < 1) pad = 1;
    int m = n + 2 * pad;
    double* ext = (double*)malloc(sizeof(double) * (size_t)m);
    double* y = (double*)malloc(sizeof(double) * (size_t)m);
    for (int i = 0; i < pad; ++i) ext[i] = 2.0 * x[0] - x[pad - i];
    for (int i = 0; i < n; ++i) ext[pad + i] = x[i];
    for (int i = 1; i <= pad; ++i) ext[pad + n - 1 + i] = 2.0 * x[n - 1] - x[n - 1 - i];
    y[0] = ext[0];
    y[1] = ext[1];
    for (int i = 2; i < m; ++i)
        y[i] = a[0] * ext[i] + a[1] * ext[i - 1] + a[2] * ext[i - 2] + b[0] * y[i - 1] + b[1] * y[i - 2];
    for (int i = 0; i < n; ++i) out[i] = y[pad + i];
    free(ext);
    free(y);
}

void oracle_cfc_filter(const double* x, int32_t n, double dt, double cfc, double* out) {
    if (n < 4) {
        for (int i = 0; i < n; ++i) out[i] = x[i];
        return;
    }
    double wd = 2.0 * 3.141592653589793 * cfc * 2.0775;
    double wa = sin(wd * dt / 2.0) / cos(wd * dt / 2.0);
    double s2 = sqrt(2.0);
    double den = 1.0 + s2 * wa + wa * wa;
    double a0 = wa * wa / den;
    double a[3] = {a0, 2.0 * a0, a0};
    double b[2] = {-2.0 * (wa * wa - 1.0) / den, (-1.0 + s2 * wa - wa * wa) / den};
    double* f = (double*)malloc(sizeof(double) * (size_t)n);
    double* r = (double*)malloc(sizeof(double) * (size_t)n);
    cfc_pass(x, n, dt, cfc, a, b, f);
    for (int i = 0; i < n; ++i) r[i] = f[n - 1 - i];
    cfc_pass(r, n, dt, cfc, a, b, f);
    for (int i = 0; i < n; ++i) out[i] = f[n - 1 - i];
    free(f);
    free(r);
}

/* resample (src/signal.rs:43-64); returns the count, writes at most cap. */
int32_t oracle_resample(const double* t, const double* y, int32_t n, double t0, double dt, double t_end, double* out,
                        int32_t cap) {
    int32_t m = (int32_t)floor((t_end - t0) / dt) + 1;
    int j = 0;
    for (int k = 0; k < m && k < cap; ++k) {
        double tk = t0 + (double)k * dt;
        while (j + 1 < n && t[j + 1] < tk) ++j;
        double v;
        if (j + 1 >= n)
            v = y[n - 1];
        else if (tk <= t[0])
            v = y[0];
        else {
            double w = (tk - t[j]) / (t[j + 1] - t[j]);
            v = y[j] + w * (y[j + 1] - y[j]);
        }
        out[k] = v;
    }
    return m;
}

/* integrate (src/signal.rs:67-76). */
void oracle_integrate(const double* t, const double* y, int32_t n, double y0, double* out) {
    double acc = y0;
    if (n <= 0) return;
    out[0] = acc;
    for (int i = 1; i < n; ++i) {
        acc += 0.5 * (y[i] + y[i - 1]) * (t[i] - t[i - 1]);
        out[i] = acc;
    }
}

/* Local-x acceleration of group g's accelerometer node: sample_accelerometers
 * (src/solver.rs:150-187) with the acceleration of :333-342 and the frame
 * axis of LocalFrame::axes (src/model.rs:45-61). */
static double accel_local_x(const crush_model* m, const crush_group* g, const double* u, const double* f_ext,
                            const double* f_int, const double* inv_mass, const uint8_t* is_fixed) {
    int n = g->accel_node;
    if (n < 0) return 0.0;
    double a[3];
    for (int d = 0; d < 3; ++d) a[d] = is_fixed[n] ? 0.0 : (f_ext[3 * n + d] - f_int[3 * n + d]) * inv_mass[n];
    if (g->frame_origin < 0) return a[0];
    double e[3];
    for (int d = 0; d < 3; ++d)
        e[d] = (m->nodes[3 * g->frame_x + d] + u[3 * g->frame_x + d]) -
               (m->nodes[3 * g->frame_origin + d] + u[3 * g->frame_origin + d]);
    double en = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    if (en < 1e-300) return a[0];
    for (int d = 0; d < 3; ++d) e[d] = e[d] / en;
    return e[0] * a[0] + e[1] * a[1] + e[2] * a[2];
}

/* ------------------------------------------------ histories and frames */

/* Growable record store of the histories: per sample the time, the step,
 * part_states (src/solver.rs:114-144) and sample_accelerometers
 * (src/solver.rs:150-187). */
typedef struct {
    int n_parts, n_acc;
    int* part_ptr; /* [n_parts + 1] */
    int* part_nodes;
    size_t n, cap, rec;
    double* data;
} history_t;

static int cmp_int(const void* a, const void* b) {
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

/* Mesh::part_nodes (src/mesh.rs:74-82): nodes of each part's elements,
 * ascending, without repeats. */
static void history_init(const crush_model* m, history_t* h) {
    memset(h, 0, sizeof *h);
    h->n_parts = m->n_parts;
    h->n_acc = m->n_groups < 8 ? m->n_groups : 8;
    h->rec = 2 + (size_t)CRUSH_PART_FIELDS * h->n_parts + (size_t)CRUSH_ACCEL_FIELDS * h->n_acc;
    h->part_ptr = (int*)calloc((size_t)m->n_parts + 1, sizeof(int));
    h->part_nodes = (int*)malloc(sizeof(int) * (8 * (size_t)m->n_hexes + 1));
    int* tmp = (int*)malloc(sizeof(int) * (8 * (size_t)m->n_hexes + 1));
    int total = 0;
    for (int p = 0; p < m->n_parts; ++p) {
        int k = 0;
        for (int e = 0; e < m->n_hexes; ++e)
            if (m->hex_part[e] == p)
                for (int i = 0; i < 8; ++i) tmp[k++] = m->hexes[8 * e + i];
        qsort(tmp, (size_t)k, sizeof(int), cmp_int);
        for (int i = 0; i < k; ++i)
            if (i == 0 || tmp[i] != tmp[i - 1]) h->part_nodes[total++] = tmp[i];
        h->part_ptr[p + 1] = total;
    }
    free(tmp);
    h->cap = 256;
    h->data = (double*)malloc(sizeof(double) * h->rec * h->cap);
}

static void history_free(history_t* h) {
    free(h->part_ptr);
    free(h->part_nodes);
    free(h->data);
}

/* LocalFrame::axes (src/model.rs:44-58), rows e1, e2, e3. */
static void frame_axes(const double* xo, const double* xx, const double* xp, double* ax) {
    double e1[3], p[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < 3; ++d) e1[d] = xx[d] - xo[d];
    if (xp)
        for (int d = 0; d < 3; ++d) p[d] = xp[d] - xo[d];
    double n1 = sqrt(dot(e1, e1));
    if (n1 < 1e-300) {
        e1[0] = 1.0;
        e1[1] = e1[2] = 0.0;
        n1 = 1.0;
    }
    for (int d = 0; d < 3; ++d) e1[d] /= n1;
    double c[3];
    cross(p, e1, c);
    if (sqrt(dot(c, c)) < 1e-12 * fmax(sqrt(dot(p, p)), 1e-300)) {
        int xa = fabs(e1[0]) < 0.9;
        p[0] = xa ? 1.0 : 0.0;
        p[1] = xa ? 0.0 : 1.0;
        p[2] = 0.0;
    }
    double e3[3], e2[3];
    cross(e1, p, e3);
    double n3 = sqrt(dot(e3, e3));
    for (int d = 0; d < 3; ++d) e3[d] /= n3;
    cross(e3, e1, e2);
    for (int d = 0; d < 3; ++d) {
        ax[d] = e1[d];
        ax[3 + d] = e2[d];
        ax[6 + d] = e3[d];
    }
}

static void history_record(const crush_model* m, history_t* h, double t, long step, const double* masses,
                           const double* u, const double* v, const double* acc) {
    if (h->n == h->cap) {
        h->cap *= 2;
        h->data = (double*)realloc(h->data, sizeof(double) * h->rec * h->cap);
    }
    double* r = h->data + h->n * h->rec;
    r[0] = t;
    r[1] = (double)step;
    for (int p = 0; p < h->n_parts; ++p) {
        double mt = 0.0, mv[3] = {0, 0, 0}, mu[3] = {0, 0, 0}, ma[3] = {0, 0, 0}, ke = 0.0;
        for (int i = h->part_ptr[p]; i < h->part_ptr[p + 1]; ++i) {
            int n = h->part_nodes[i];
            mt += masses[n];
            for (int d = 0; d < 3; ++d) {
                mv[d] += masses[n] * v[3 * n + d];
                mu[d] += masses[n] * u[3 * n + d];
                ma[d] += masses[n] * acc[3 * n + d];
                ke += 0.5 * masses[n] * v[3 * n + d] * v[3 * n + d];
            }
        }
        double inv = mt > 0.0 ? 1.0 / mt : 0.0;
        double* q = r + 2 + (size_t)CRUSH_PART_FIELDS * p;
        q[0] = mt;
        for (int d = 0; d < 3; ++d) {
            q[1 + d] = mu[d] * inv;
            q[4 + d] = mv[d] * inv;
            q[7 + d] = ma[d] * inv;
        }
        q[10] = ke;
    }
    for (int g = 0; g < h->n_acc; ++g) {
        const crush_group* gr = &m->groups[g];
        double* q = r + 2 + (size_t)CRUSH_PART_FIELDS * h->n_parts + (size_t)CRUSH_ACCEL_FIELDS * g;
        memset(q, 0, sizeof(double) * CRUSH_ACCEL_FIELDS);
        int n = gr->accel_node;
        if (n < 0) {
            q[15] = q[19] = q[23] = 1.0;
            continue;
        }
        for (int d = 0; d < 3; ++d) {
            q[d] = m->nodes[3 * n + d] + u[3 * n + d];
            q[3 + d] = u[3 * n + d];
            q[6 + d] = v[3 * n + d];
            q[9 + d] = acc[3 * n + d];
        }
        if (gr->frame_origin >= 0) {
            double xo[3], xx[3], xp[3];
            for (int d = 0; d < 3; ++d) {
                xo[d] = m->nodes[3 * gr->frame_origin + d] + u[3 * gr->frame_origin + d];
                xx[d] = m->nodes[3 * gr->frame_x + d] + u[3 * gr->frame_x + d];
                if (gr->frame_plane >= 0) xp[d] = m->nodes[3 * gr->frame_plane + d] + u[3 * gr->frame_plane + d];
            }
            for (int d = 0; d < 3; ++d) q[12 + d] = xo[d];
            frame_axes(xo, xx, gr->frame_plane >= 0 ? xp : NULL, q + 15);
        } else {
            for (int d = 0; d < 3; ++d) q[12 + d] = q[d];
            q[15] = q[19] = q[23] = 1.0;
        }
    }
    ++h->n;
}

/* Frame capture (src/solver.rs:265-275): displacement, plastic strain
 * (compaction) and Cauchy stress (Elements::cauchy_stress,
 * src/kernel.rs:281-320) per element, eroded flags. */
typedef struct {
    int cap, n;
    double* time;
    float *u, *plastic, *stress;
    uint8_t* eroded;
} frames_t;

static void frame_capture(const crush_model* m, const elements_t* el, const double* u, double t, frames_t* f) {
    if (f->n >= f->cap) {
        ++f->n;
        return;
    }
    int k = f->n++;
    size_t n3 = 3 * (size_t)m->n_nodes, ne = (size_t)m->n_hexes;
    f->time[k] = t;
    for (size_t i = 0; i < n3; ++i) f->u[k * n3 + i] = (float)u[i];
    /* the lane path keeps honeycomb states in its blocks (sync_states,
     * src/kernel.rs:261-267) */
    int* lane_of = NULL;
    if (el->lanes) {
        lane_of = (int*)malloc(sizeof(int) * (ne + 1));
        for (size_t e = 0; e < ne; ++e) lane_of[e] = -1;
        for (int q = 0; q < el->n_lane; ++q) lane_of[el->lane_ids[q]] = q;
    }
    for (int e = 0; e < m->n_hexes; ++e) {
        double st[25];
        const double* s = el->state + 25 * (size_t)e;
        if (lane_of && lane_of[e] >= 0) {
            lane_blocks_state(el->lanes, lane_of[e], st);
            s = st;
        }
        float* sig = f->stress + (k * ne + (size_t)e) * 6;
        f->plastic[k * ne + (size_t)e] = el->kind[e] == CRUSH_MAT_HONEYCOMB ? (float)s[24] : 0.0f;
        f->eroded[k * ne + (size_t)e] = el->eroded[e];
        for (int q = 0; q < 6; ++q) sig[q] = 0.0f;
        if (el->eroded[e]) continue;
        double c[9];
        if (el->kind[e] == CRUSH_MAT_HONEYCOMB) {
            const double* r = s + 9;
            const double* sg = s + 18;
            double sh[9] = {sg[0], sg[3], sg[5], sg[3], sg[1], sg[4], sg[5], sg[4], sg[2]};
            double rt[9], rs[9];
            transpose3(r, rt);
            mul3(r, sh, rs);
            mul3(rs, rt, c);
        } else {
            double F[9] = {0};
            const int32_t* hx = m->hexes + 8 * (size_t)e;
            for (int i = 0; i < 8; ++i)
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        F[3 * a + b] += (m->nodes[3 * hx[i] + a] + u[3 * hx[i] + a]) * el->grad[24 * (size_t)e + 3 * i + b];
            double j = det3(F);
            if (j <= 0.0) continue;
            const crush_material* mat = &m->materials[m->hex_part[e]];
            double tau[9];
            oracle_hencky_elastic(mat->youngs_modulus, mat->poisson_ratio, F, tau);
            for (int q = 0; q < 9; ++q) c[q] = tau[q] / j;
        }
        sig[0] = (float)c[0];
        sig[1] = (float)c[4];
        sig[2] = (float)c[8];
        sig[3] = (float)c[1];
        sig[4] = (float)c[5];
        sig[5] = (float)c[2];
    }
    free(lane_of);
}

/* solver::run: src/solver.rs:190-381; simd selects the element path
 * (Settings.simd, src/model.rs:100-102). */
static int solve(const crush_model* m, crush_results* out, taps_t* taps, int simd, int threads) {
    double t_start = now_ms();
    const crush_settings* s = &m->settings;
    int nn = m->n_nodes, n = 3 * nn;
    elements_t el;
    memset(&el, 0, sizeof el);
    int rc = elements_build(m, &el);
    if (!rc && simd) rc = elements_build_lanes(m, &el, threads);
    if (rc) {
        elements_free(&el);
        return rc;
    }
    double* masses = (double*)calloc((size_t)nn, sizeof(double));
    for (int e = 0; e < m->n_hexes; ++e)
        for (int i = 0; i < 8; ++i) masses[m->hexes[8 * e + i]] += el.lumped[8 * e + i];
    double* element_mass = (double*)malloc(sizeof(double) * (size_t)(m->n_hexes + 1));
    for (int e = 0; e < m->n_hexes; ++e) element_mass[e] = el.vol[e] * m->materials[m->hex_part[e]].density;
    double* added = (double*)calloc((size_t)nn, sizeof(double));
    double* inv_mass = (double*)malloc(sizeof(double) * (size_t)(nn + 1));
    for (int i = 0; i < nn; ++i) {
        double mm = masses[i] + added[i];
        inv_mass[i] = mm > 0.0 ? 1.0 / mm : 0.0;
    }
    contact_rt* rts = (contact_rt*)calloc((size_t)m->n_contacts + 1, sizeof(contact_rt));
    int total_cn = 0;
    for (int c = 0; c < m->n_contacts; ++c) {
        contact_rt_build(m, &m->contacts[c], &rts[c]);
        total_cn += rts[c].n;
    }
    double* u = (double*)calloc((size_t)n, sizeof(double));
    double contact_dt = INFINITY;
    for (int c = 0; c < m->n_contacts; ++c) {
        const crush_contact* ct = &m->contacts[c];
        for (int i = 0; i < ct->n_nodes; ++i) {
            int nd = ct->nodes[i];
            if (masses[nd] > 0.0 && ct->stiffness > 0.0)
                contact_dt = fmin(contact_dt, 2.0 * sqrt(masses[nd] / ct->stiffness));
        }
    }
    double* dts = (double*)malloc(sizeof(double) * (size_t)(m->n_hexes + 1));
    element_time_steps(m, &el, u, dts);
    double stable = INFINITY;
    for (int e = 0; e < m->n_hexes; ++e) stable = fmin(stable, dts[e]);
    if (!isfinite(stable)) {
        static const int EDGES[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                         {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
        double min_edge = INFINITY, cmax = 0.0;
        for (int e = 0; e < m->n_hexes; ++e)
            for (int k = 0; k < 12; ++k) {
                const double* p = m->nodes + 3 * m->hexes[8 * e + EDGES[k][0]];
                const double* q = m->nodes + 3 * m->hexes[8 * e + EDGES[k][1]];
                double d = sqrt((p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) +
                                (p[2] - q[2]) * (p[2] - q[2]));
                min_edge = fmin(min_edge, d);
            }
        for (int e = 0; e < m->n_hexes; ++e) cmax = fmax(cmax, wave_speed(&m->materials[m->hex_part[e]]));
        stable = min_edge / (sqrt(3.0) * cmax);
    }
    stable = fmin(stable, contact_dt);
    double dt_floor = s->mass_scaling > 0.0 ? s->mass_scaling * stable : 0.0;
    double added_total = 0.0;
    double dt = s->time_step > 0.0 ? s->time_step : stable * s->dt_scale;
    int adaptive = !(s->time_step > 0.0) && s->adaptive_check_steps > 0;
    double* v = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    memcpy(v, m->initial_velocity, sizeof(double) * (size_t)n);
    double* v_half = (double*)calloc((size_t)n + 1, sizeof(double));
    double* f_int = (double*)calloc((size_t)n + 1, sizeof(double));
    double* f_ext = (double*)calloc((size_t)n + 1, sizeof(double));
    int32_t* hits = (int32_t*)malloc(sizeof(int32_t) * (size_t)(total_cn + 1));
    int* fixed = (int*)malloc(sizeof(int) * (size_t)(m->n_fixed + 1));
    for (int i = 0; i < m->n_fixed; ++i) fixed[i] = m->fixed_nodes[i];

    compute_forces(m, &el, u, f_int);
    for (int c = 0; c < m->n_contacts; ++c)
        contact_apply_one(m, &m->contacts[c], &rts[c], u, masses, dt, f_ext, NULL);
    /* accelerometer history (history_steps = 1): t and local-x
     * acceleration of each group at t = 0 and after every step */
    uint8_t* is_fixed = (uint8_t*)calloc((size_t)nn + 1, 1);
    for (int i = 0; i < m->n_fixed; ++i) is_fixed[m->fixed_nodes[i]] = 1;
    int ng_p = s->pulse_dt > 0.0 ? (m->n_groups < 8 ? m->n_groups : 8) : 0;
    size_t h_cap = 1024, h_n = 0;
    double* h_t = (double*)malloc(sizeof(double) * h_cap);
    double* h_a = (double*)malloc(sizeof(double) * h_cap * 8);
    if (ng_p > 0) {
        h_t[0] = 0.0;
        for (int g = 0; g < ng_p; ++g) h_a[g] = accel_local_x(m, &m->groups[g], u, f_ext, f_int, inv_mass, is_fixed);
        h_n = 1;
    }
    /* histories (history_steps) and frames (frame_steps) */
    double* acc = (double*)calloc((size_t)n + 1, sizeof(double));
    for (int k = 0; k < n; ++k) acc[k] = is_fixed[k / 3] ? 0.0 : (f_ext[k] - f_int[k]) * inv_mass[k / 3];
    history_t hist;
    int hs = s->history_steps > 0 ? s->history_steps : 0;
    if (hs) {
        history_init(m, &hist);
        history_record(m, &hist, 0.0, 0, masses, u, v, acc);
    }
    int fs = s->frame_steps > 0 ? s->frame_steps : 0;
    frames_t frames;
    memset(&frames, 0, sizeof frames);
    if (fs && out && out->frame_capacity > 0) {
        frames.cap = out->frame_capacity;
        size_t n3 = (size_t)n, ne = (size_t)m->n_hexes, fc = (size_t)frames.cap;
        frames.time = (double*)calloc(fc, sizeof(double));
        frames.u = (float*)calloc(fc * n3, sizeof(float));
        frames.plastic = (float*)calloc(fc * ne, sizeof(float));
        frames.stress = (float*)calloc(fc * ne * 6, sizeof(float));
        frames.eroded = (uint8_t*)calloc(fc * ne, 1);
    }
    if (fs) frame_capture(m, &el, u, 0.0, &frames);

    double t = 0.0;
    long step = 0;
    long eroded_total = 0;
    double dt_cfl = dt;
    int chk = s->adaptive_check_steps > 0 ? s->adaptive_check_steps : 1;
    double crush_max[8] = {0};
    int had_pair = 0;
    long quiet = 0;
    int separated = 0;
    rc = CRUSH_OK;
    while (t < s->end_time) {
        if (taps && step >= taps->n_steps) break;
        if (adaptive && step % s->adaptive_check_steps == 0) {
            element_time_steps(m, &el, u, dts);
            double dt_stable = INFINITY;
            for (int e = 0; e < m->n_hexes; ++e) dt_stable = fmin(dt_stable, dts[e]);
            if (isfinite(dt_stable)) {
                double dt_use = dt_stable;
                if (dt_floor > 0.0 && dt_stable < dt_floor) {
                    /* Elements::mass_scaling: src/kernel.rs:243-258 */
                    memset(added, 0, sizeof(double) * (size_t)nn);
                    double total = 0.0;
                    for (int e = 0; e < m->n_hexes; ++e) {
                        if (dts[e] < dt_floor) {
                            double r = dt_floor / dts[e];
                            double extra = (fmin(r * r, 100.0) - 1.0) * element_mass[e];
                            for (int i = 0; i < 8; ++i) added[m->hexes[8 * e + i]] += extra / 8.0;
                            total += extra;
                        }
                    }
                    for (int i = 0; i < nn; ++i) {
                        double mm = masses[i] + added[i];
                        inv_mass[i] = mm > 0.0 ? 1.0 / mm : 0.0;
                    }
                    added_total = fmax(added_total, total);
                    dt_use = dt_floor;
                }
                dt_cfl = s->dt_scale * fmin(dt_use, contact_dt);
            }
        }
        dt = fmax(fmin(dt_cfl, s->end_time - t), 1e-15);
        if (step % 20 == 0) {
            for (int k = 0; k < n; ++k)
                if (!isfinite(u[k])) {
                    rc = CRUSH_E_NONFINITE;
                    goto done;
                }
        }
        memset(f_ext, 0, sizeof(double) * (size_t)n);
        int pair_hits = 0, off = 0;
        for (int c = 0; c < m->n_contacts; ++c) {
            int cnt = contact_apply_one(m, &m->contacts[c], &rts[c], u, masses, dt, f_ext, hits + off);
            if (m->contacts[c].pair) pair_hits += cnt;
            off += rts[c].n;
        }
        for (int k = 0; k < n; ++k) {
            double a = (f_ext[k] - f_int[k]) * inv_mass[k / 3];
            v_half[k] = v[k] + 0.5 * dt * a;
            u[k] += dt * v_half[k];
        }
        for (int i = 0; i < m->n_fixed; ++i)
            for (int d = 0; d < 3; ++d) {
                int k = 3 * fixed[i] + d;
                u[k] = 0.0;
                v[k] = 0.0;
                v_half[k] = 0.0;
            }
        memset(f_int, 0, sizeof(double) * (size_t)n);
        eroded_total += compute_forces(m, &el, u, f_int);
        for (int k = 0; k < n; ++k) {
            double a = (f_ext[k] - f_int[k]) * inv_mass[k / 3];
            v[k] = (v_half[k] + 0.5 * dt * a) * s->velocity_damping;
        }
        for (int i = 0; i < m->n_fixed; ++i)
            for (int d = 0; d < 3; ++d) v[3 * fixed[i] + d] = 0.0;
        if (taps) {
            size_t st = (size_t)step;
            if (taps->f_int) memcpy(taps->f_int + st * n, f_int, sizeof(double) * (size_t)n);
            if (taps->f_ext) memcpy(taps->f_ext + st * n, f_ext, sizeof(double) * (size_t)n);
            if (taps->hit) memcpy(taps->hit + st * total_cn, hits, sizeof(int32_t) * (size_t)total_cn);
            if (taps->eroded) memcpy(taps->eroded + st * m->n_hexes, el.eroded, (size_t)m->n_hexes);
            if (taps->dt) taps->dt[st] = dt;
            if (taps->u) memcpy(taps->u + st * n, u, sizeof(double) * (size_t)n);
        }
        step += 1;
        t += dt;
        if (ng_p > 0) {
            if (h_n == h_cap) {
                h_cap *= 2;
                h_t = (double*)realloc(h_t, sizeof(double) * h_cap);
                h_a = (double*)realloc(h_a, sizeof(double) * h_cap * 8);
            }
            h_t[h_n] = t;
            for (int g = 0; g < ng_p; ++g)
                h_a[8 * h_n + g] = accel_local_x(m, &m->groups[g], u, f_ext, f_int, inv_mass, is_fixed);
            ++h_n;
        }
        if (hs || fs) {
            for (int k = 0; k < n; ++k) acc[k] = is_fixed[k / 3] ? 0.0 : (f_ext[k] - f_int[k]) * inv_mass[k / 3];
            if (hs && step % hs == 0) history_record(m, &hist, t, step, masses, u, v, acc);
            if (fs && step % fs == 0) frame_capture(m, &el, u, t, &frames);
        }
        for (int g = 0; g < m->n_groups && g < 8; ++g)
            crush_max[g] = fmax(crush_max[g], group_crush(&m->groups[g], u));
        /* retirement at separation (crushrs.h, crush_settings) */
        if (pair_hits > 0) {
            had_pair = 1;
            quiet = 0;
        } else {
            ++quiet;
        }
        if (s->separation_steps > 0 && m->n_groups >= 2 && had_pair && quiet >= s->separation_steps &&
            step % chk == 0) {
            double x0[3], v0[3], x1[3], v1[3];
            group_means(m, &m->groups[0], masses, u, v, x0, v0);
            group_means(m, &m->groups[1], masses, u, v, x1, v1);
            double sep = (v1[0] - v0[0]) * (x1[0] - x0[0]) + (v1[1] - v0[1]) * (x1[1] - x0[1]) +
                         (v1[2] - v0[2]) * (x1[2] - x0[2]);
            if (sep > 0.0) {
                separated = 1;
                break;
            }
        }
    }
done:
    if (taps) taps->steps_done = (int32_t)step;
    if (rc == CRUSH_OK) {
        if (hs && step % hs != 0) history_record(m, &hist, t, step, masses, u, v, acc);
        if (fs && step % fs != 0) frame_capture(m, &el, u, t, &frames);
    }
    if (out) {
        out->history_samples = hs ? (int32_t)hist.n : 0;
        for (size_t k = 0; hs && k < hist.n && (int32_t)k < out->history_capacity; ++k) {
            const double* r = hist.data + k * hist.rec;
            if (out->history_time) out->history_time[k] = r[0];
            if (out->history_step) out->history_step[k] = (int64_t)r[1];
            if (out->history)
                memcpy(out->history + k * (size_t)hist.n_parts * CRUSH_PART_FIELDS, r + 2,
                       sizeof(double) * CRUSH_PART_FIELDS * (size_t)hist.n_parts);
            if (out->accel_history && hist.n_acc > 0)
                memcpy(out->accel_history + k * (size_t)hist.n_acc * CRUSH_ACCEL_FIELDS,
                       r + 2 + (size_t)CRUSH_PART_FIELDS * hist.n_parts, sizeof(double) * CRUSH_ACCEL_FIELDS * (size_t)hist.n_acc);
        }
        out->frame_samples = fs ? frames.n : 0;
        int fk = frames.n < frames.cap ? frames.n : frames.cap;
        size_t n3 = (size_t)n, ne = (size_t)m->n_hexes;
        if (fk > 0) {
            if (out->frame_time) memcpy(out->frame_time, frames.time, sizeof(double) * (size_t)fk);
            if (out->frame_displacement) memcpy(out->frame_displacement, frames.u, sizeof(float) * n3 * (size_t)fk);
            if (out->frame_plastic_strain) memcpy(out->frame_plastic_strain, frames.plastic, sizeof(float) * ne * (size_t)fk);
            if (out->frame_stress) memcpy(out->frame_stress, frames.stress, sizeof(float) * 6 * ne * (size_t)fk);
            if (out->frame_eroded) memcpy(out->frame_eroded, frames.eroded, ne * (size_t)fk);
        }
    }
    if (hs) history_free(&hist);
    free(frames.time); free(frames.u); free(frames.plastic); free(frames.stress); free(frames.eroded);
    free(acc);
    if (out) {
        out->time = t;
        out->steps = step;
        out->eroded_elements = eroded_total;
        out->added_mass_max = added_total;
        out->separated = separated;
        out->wall_ms = now_ms() - t_start;
        out->device_ms = 0.0;
        if (out->displacement) memcpy(out->displacement, u, sizeof(double) * (size_t)n);
        if (out->velocity) memcpy(out->velocity, v, sizeof(double) * (size_t)n);
        if (out->masses) memcpy(out->masses, masses, sizeof(double) * (size_t)nn);
        out->n_groups = m->n_groups;
        if (out->groups) {
            for (int g = 0; g < m->n_groups; ++g) {
                const crush_group* gr = &m->groups[g];
                double mt = 0.0, p1[3] = {0, 0, 0}, p0[3] = {0, 0, 0};
                for (int i = 0; i < gr->n_nodes; ++i) {
                    int nd = gr->nodes[i];
                    mt += masses[nd];
                    for (int d = 0; d < 3; ++d) {
                        p1[d] += masses[nd] * v[3 * nd + d];
                        p0[d] += masses[nd] * m->initial_velocity[3 * nd + d];
                    }
                }
                crush_group_result* r = &out->groups[g];
                r->mass = mt;
                for (int d = 0; d < 3; ++d) r->delta_v[d] = mt > 0 ? p1[d] / mt - p0[d] / mt : 0.0;
                r->crush_end = group_crush(gr, u);
                r->crush_max = g < 8 ? crush_max[g] : r->crush_end;
                r->pulse_delta_v = 0.0;
            }
        }
        /* Pulse::from_history + filtered(60) + velocity (src/signal.rs:
         * 127-146, 184-193; src/headon.rs:70-78) */
        out->pulse_samples = 0;
        if (ng_p > 0 && h_n >= 1) {
            double* ys = (double*)malloc(sizeof(double) * h_n);
            int32_t ns = (int32_t)floor((h_t[h_n - 1] - 0.0) / s->pulse_dt) + 1;
            double* pr = (double*)malloc(sizeof(double) * (size_t)ns);
            double* pf = (double*)malloc(sizeof(double) * (size_t)ns);
            double* tg = (double*)malloc(sizeof(double) * (size_t)ns);
            double* vel = (double*)malloc(sizeof(double) * (size_t)ns);
            for (int k = 0; k < ns; ++k) tg[k] = 0.0 + (double)k * s->pulse_dt;
            for (int g = 0; g < ng_p; ++g) {
                for (size_t i = 0; i < h_n; ++i) ys[i] = h_a[8 * i + g];
                oracle_resample(h_t, ys, (int32_t)h_n, 0.0, s->pulse_dt, h_t[h_n - 1], pr, ns);
                if (out->pulse && out->pulse_capacity > 0)
                    for (int k = 0; k < ns && k < out->pulse_capacity; ++k)
                        out->pulse[(size_t)g * out->pulse_capacity + k] = pr[k];
                oracle_cfc_filter(pr, ns, s->pulse_dt, 60.0, pf);
                oracle_integrate(tg, pf, ns, 0.0, vel);
                if (out->groups) {
                    int hd = m->groups[g].heading;
                    out->groups[g].pulse_delta_v = hd < 0 ? vel[ns - 1] : -vel[ns - 1];
                }
            }
            out->pulse_samples = ns;
            free(ys); free(pr); free(pf); free(tg); free(vel);
        }
    }
    free(h_t);
    free(h_a);
    free(is_fixed);
    for (int c = 0; c < m->n_contacts; ++c) contact_rt_free(&rts[c]);
    free(rts); free(u); free(v); free(v_half); free(f_int); free(f_ext); free(hits); free(fixed);
    free(masses); free(element_mass); free(added); free(inv_mass); free(dts);
    elements_free(&el);
    return rc;
}

int oracle_run(const crush_model* model, crush_results* out) { return solve(model, out, NULL, 0, 1); }

int oracle_run_lanes(const crush_model* model, int32_t threads, crush_results* out) {
    return solve(model, out, NULL, 1, threads);
}

int oracle_run_steps_lanes(const crush_model* model, int32_t n_steps, int32_t threads, double* f_int,
                           double* f_ext, int32_t* hit, uint8_t* eroded, double* dt,
                           int32_t* steps_done) {
    taps_t t = {n_steps, f_int, f_ext, hit, eroded, dt, 0, NULL};
    int rc = solve(model, NULL, &t, 1, threads);
    if (steps_done) *steps_done = t.steps_done;
    return rc;
}

int oracle_run_steps(const crush_model* model, int32_t n_steps, double* f_int, double* f_ext,
                     int32_t* hit, uint8_t* eroded, double* dt, int32_t* steps_done) {
    taps_t t = {n_steps, f_int, f_ext, hit, eroded, dt, 0, NULL};
    int rc = solve(model, NULL, &t, 0, 1);
    if (steps_done) *steps_done = t.steps_done;
    return rc;
}

int oracle_run_displacements(const crush_model* model, int32_t n_steps, int32_t simd, double* u,
                             int32_t* steps_done) {
    taps_t t = {n_steps, NULL, NULL, NULL, NULL, NULL, 0, u};
    int rc = solve(model, NULL, &t, simd ? 1 : 0, 1);
    if (steps_done) *steps_done = t.steps_done;
    return rc;
}
