// Device-side building blocks shared by the one-solve kernels and the
// batched kernels: the f32 element force (honeycomb rate form and Hencky
// elastic), the f64 node-to-face contact test and the f64 characteristic
// length.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>

#include "prep.hpp"

// Element math is also compiled for the host so that tests can exercise it
// on the CPU; device code uses the hardware reciprocal square root.
#define CRUSH_HD __host__ __device__ __forceinline__

namespace crush {

CRUSH_HD float rsqrt_f(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return rsqrtf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}

// The largest float <= x and the smallest float >= x (NaN stays NaN): the
// nearest float, stepped one ulp outward when it lies on the wrong side.
// Boxes rounded so stay supersets of their f64 originals.
CRUSH_HD float f32_next(float r, bool up) {
    if (r == 0.0f) return up ? __builtin_bit_cast(float, 1u) : -__builtin_bit_cast(float, 1u);
    const unsigned b = __builtin_bit_cast(unsigned, r);
    return __builtin_bit_cast(float, (r > 0.0f) == up ? b + 1u : b - 1u);
}
CRUSH_HD float f32_down(double x) {
    const float r = (float)x;
    return (double)r > x ? f32_next(r, false) : r;
}
CRUSH_HD float f32_up(double x) {
    const float r = (float)x;
    return (double)r < x ? f32_next(r, true) : r;
}

// Per-solve control block in device memory.
struct StepCtl {
    double t, dt, dt_cfl, end_time;
    double contact_dt, dt_floor, dt_scale, damping;
    double added_total_max;
    long long step;
    unsigned long long eroded_total;
    int done, error, scale_now, had_pair;
    int quiet, separated, pair_hits, chk;
    int sep_steps, adaptive;
    int phase;        // batched path: PH_* work of the current tick
    int eroded_step;  // batched path: elements eroded by the open step's force pass
    int bad_step;     // batched path: BAD_* a node became non-finite
    int reserved3;
    double crush_max[8];
    double crush_end[8];
    unsigned long long t_retire;  // device wall clock at retirement (batched path)
    unsigned long long reserved2;
    // accelerometer pulse (groups < 8): the last sample's time and local-x
    // accelerations, and the next grid index k of t = k pulse_dt to fill
    double pulse_t;
    double pulse_a[8];
    int pulse_k;
    int hist_k;  // history records taken after the initial one (may exceed the capacity)
};

// Batched path, per instance and tick (StepCtl::phase): PH_HALF runs the
// contact, first half step and element forces of the step after the last
// closed one; PH_GATHER gathers the forces of the step in flight (with
// PH_HALF in the same node pass, before that half step).  StepCtl::bad_step:
// the half step (BAD_HALF) or the gather (BAD_GATHER) left a non-finite
// value.
constexpr int PH_HALF = 1, PH_GATHER = 2;
constexpr int BAD_HALF = 1, BAD_GATHER = 2;

// Face geometry record width (doubles) in the SoA face array:
// x0..x3 (12), normal (3), centre (3), inflated lo (3), hi (3),
// candidate lo (3), candidate hi (3), size (1).
constexpr int FG = 37;
constexpr int FG_N = 12, FG_C = 15, FG_LO = 18, FG_HI = 21, FG_CLO = 24, FG_CHI = 27, FG_SZ = 30;
// batched path: the candidate box a face had when its contact's grid was
// last built (a reused grid stays valid while every box stays within it
// plus the grid's margin)
constexpr int FG_BLO = 31, FG_BHI = 34;

// Contact constants: src/contact.rs:75 (MAX_ACCEL), :87 (SOFT), :34 (edge
// tolerance of the triangle test).
constexpr double C_MAX_ACCEL = 1.0e5;
constexpr double C_SOFT = 0.1;
constexpr double C_TRI_TOL = 0.1;

// ------------------------------------------------------------ f32 3x3

// The f32 element arithmetic is compiled without contraction and states
// its fused multiply-adds explicitly: every kernel that inlines it, on
// either path of a class-uniformity branch, and the host build round the
// same way, so results do not depend on how the compiler schedules a
// kernel.
CRUSH_HD float fmf(float a, float b, float c) { return fmaf(a, b, c); }
// a0 b0 + a1 b1 + a2 b2, accumulated left to right
CRUSH_HD float dot3f(float a0, float b0, float a1, float b1, float a2, float b2) {
#pragma clang fp contract(off)
    return fmf(a2, b2, fmf(a1, b1, a0 * b0));
}
// a b - c d
CRUSH_HD float cross2f(float a, float b, float c, float d) {
#pragma clang fp contract(off)
    return fmf(a, b, -(c * d));
}

CRUSH_HD float det3f(const float* m) {
#pragma clang fp contract(off)
    const float c0 = cross2f(m[4], m[8], m[5], m[7]);
    const float c1 = cross2f(m[3], m[8], m[5], m[6]);
    const float c2 = cross2f(m[3], m[7], m[4], m[6]);
    return fmf(m[2], c2, fmf(-m[1], c1, m[0] * c0));
}

// Cofactor inverse with a given determinant.
CRUSH_HD void inv3f(const float* m, float det, float* r) {
#pragma clang fp contract(off)
    float id = 1.0f / det;
    r[0] = cross2f(m[4], m[8], m[5], m[7]) * id;
    r[1] = cross2f(m[2], m[7], m[1], m[8]) * id;
    r[2] = cross2f(m[1], m[5], m[2], m[4]) * id;
    r[3] = cross2f(m[5], m[6], m[3], m[8]) * id;
    r[4] = cross2f(m[0], m[8], m[2], m[6]) * id;
    r[5] = cross2f(m[2], m[3], m[0], m[5]) * id;
    r[6] = cross2f(m[3], m[7], m[4], m[6]) * id;
    r[7] = cross2f(m[1], m[6], m[0], m[7]) * id;
    r[8] = cross2f(m[0], m[4], m[1], m[3]) * id;
}

CRUSH_HD void mul3f(const float* a, const float* b, float* r) {
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r[3 * i + j] = dot3f(a[3 * i], b[j], a[3 * i + 1], b[3 + j], a[3 * i + 2], b[6 + j]);
}

// ------------------------------------------------- element force (f32)

// Honeycomb rate-form stress update (src/material.rs:472-535, the f32
// lane form of src/kernel_simd.rs:336-443).  hd = F - I is the current
// displacement gradient; st[25] = H_prev (9, the displacement gradient at
// the last committed step), R (9), sigma_hat (6), compaction.  The
// reference's F_inc - I = F F_prev^-1 - I is formed here as
// (H - H_prev) F_prev^-1, the same quantity without the cancellation
// against I that limits an f32 F_prev^-1.  Updates st and returns the
// Kirchhoff stress tau (row-major).
CRUSH_HD void honeycomb_update(const MatParams& mp, const float* hd, float* st, float* tau) {
#pragma clang fp contract(off)
    float fprev[9], fpinv[9], dh[9], mid[9], midinv[9], g[9], dinc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        fprev[k] = st[k] + ((k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f);
        dh[k] = hd[k] - st[k];
    }
    inv3f(fprev, det3f(fprev), fpinv);
    mul3f(dh, fpinv, dinc);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        float id = (k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f;
        mid[k] = fmf(0.5f, dinc[k], id);
    }
    inv3f(mid, det3f(mid), midinv);
    mul3f(dinc, midinv, g);
    float d[9], w[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            d[3 * i + j] = 0.5f * (g[3 * i + j] + g[3 * j + i]);
            w[3 * i + j] = 0.5f * (g[3 * i + j] - g[3 * j + i]);
        }
    // Cayley rotation increment I + (W + W^2/2) / (1 + |omega|^2 / 4)
    float omega2 = dot3f(w[1], w[1], w[2], w[2], w[5], w[5]);
    float inv_den = 1.0f / fmf(0.25f, omega2, 1.0f);
    float ww[9], dr[9], r[9];
    mul3f(w, w, ww);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        float id = (k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f;
        dr[k] = fmf(fmf(0.5f, ww[k], w[k]), inv_den, id);
    }
    mul3f(dr, st + 9, r);
    // strain increment in the corotational frame: R^T d R
    float rtd[9], de[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) rtd[3 * i + j] = dot3f(r[i], d[j], r[3 + i], d[3 + j], r[6 + i], d[6 + j]);
    mul3f(rtd, r, de);
    const float e = mp.youngs;
    float s[6];
    s[0] = fmf(e, de[0], st[18]);
    s[1] = fmf(e, de[4], st[19]);
    s[2] = fmf(e, de[8], st[20]);
    s[3] = fmf(e, de[1], st[21]);
    s[4] = fmf(e, de[5], st[22]);
    s[5] = fmf(e, de[2], st[23]);
    const float c = st[24];
    // Active knot: the last i with c >= knot_c[i] (src/material.rs:459-467);
    // knots strictly increase with unused ones at +inf, so it is the count
    // of knots i >= 1 at or below c.
    int ki = 0;
#pragma unroll
    for (int i = 1; i < 8; ++i) ki += (c >= mp.knot_c[i]) ? 1 : 0;
    const float kc = mp.knot_c[ki], ks = mp.knot_s[ki], kh = mp.knot_h[ki];
    const float hard = kh;
    const float syc = fmf(hard, c - kc, ks);
    const float tk[3] = {1.0f, mp.trans, mp.trans};
    // Normal caps with implicit hardening: three passes of the active-set
    // solve dc = sum_A(-s_k - t_k syc) / (E + H sum_A t_k).
    float dc = 0.0f;
#pragma unroll
    for (int it = 0; it < 3; ++it) {
        float sy = fmf(hard, dc, syc);
        float sum = 0.0f, ta = 0.0f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            bool over = -s[k] > tk[k] * sy;
            sum += over ? fmf(-tk[k], syc, -s[k]) : 0.0f;
            ta += over ? tk[k] : 0.0f;
        }
        dc = ta > 0.0f ? sum / fmf(ta, hard, e) : 0.0f;
    }
    const float sy = fmaxf(fmf(hard, dc, syc), 0.0f);
#pragma unroll
    for (int k = 0; k < 3; ++k) s[k] = fminf(fmaxf(s[k], -tk[k] * sy), tk[k] * sy);
    const float ty = 0.5f * sy;
#pragma unroll
    for (int k = 3; k < 6; ++k) s[k] = fminf(fmaxf(s[k], -ty), ty);
    // tau = J R sigma_hat R^T
    const float sh[9] = {s[0], s[3], s[5], s[3], s[1], s[4], s[5], s[4], s[2]};
    float rs[9];
    mul3f(r, sh, rs);
    float F[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = hd[k] + ((k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f);
    const float J = det3f(F);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            tau[3 * i + j] = J * dot3f(rs[3 * i], r[3 * j], rs[3 * i + 1], r[3 * j + 1], rs[3 * i + 2], r[3 * j + 2]);
    // commit
#pragma unroll
    for (int k = 0; k < 9; ++k) st[k] = hd[k];
#pragma unroll
    for (int k = 0; k < 9; ++k) st[9 + k] = r[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) st[18 + k] = s[k];
    st[24] = c + dc;
}

CRUSH_HD float rsqrt_t(float x) { return rsqrt_f(x); }
CRUSH_HD double rsqrt_t(double x) { return 1.0 / sqrt(x); }

// Symmetric 3x3 eigen-decomposition (vals ascending irrelevant here):
// trigonometric root for the well-separated eigenvalue, its eigenvector
// from the largest row cross product, one exact Jacobi rotation in the
// orthogonal plane and Rayleigh-quotient eigenvalues (the method of
// src/material.rs:570-659).  Works on any scale; vecs column-major
// (vecs[3*c + i] = component i of eigenvector c).
template <class T>
CRUSH_HD void sym_eigen3(const T* a, T* vals, T* vecs) {
    const T eps2 = std::numeric_limits<T>::epsilon() * std::numeric_limits<T>::epsilon();
    T a00 = a[0], a01 = a[1], a02 = a[2], a11 = a[4], a12 = a[5], a22 = a[8];
    T mx = fmax(fmax(fmax(fabs(a00), fabs(a01)), fmax(fabs(a02), fabs(a11))), fmax(fabs(a12), fabs(a22)));
    if (!(mx >= std::numeric_limits<T>::min() * T(1e8))) {
        vals[0] = vals[1] = vals[2] = T(0.0);
#pragma unroll
        for (int k = 0; k < 9; ++k) vecs[k] = (k == 0 || k == 4 || k == 8) ? T(1.0) : T(0.0);
        return;
    }
    const T inv = T(1.0) / mx;
    a00 *= inv; a01 *= inv; a02 *= inv; a11 *= inv; a12 *= inv; a22 *= inv;
    const T off = a01 * a01 + a02 * a02 + a12 * a12;
    // off-diagonals below the resolution of the diagonal: treat as diagonal
    if (off < eps2) {
        vals[0] = a00 * mx; vals[1] = a11 * mx; vals[2] = a22 * mx;
#pragma unroll
        for (int k = 0; k < 9; ++k) vecs[k] = (k == 0 || k == 4 || k == 8) ? T(1.0) : T(0.0);
        return;
    }
    const T q = (a00 + a11 + a22) * (T(1.0) / T(3.0));
    const T b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
    const T p = sqrt((b00 * b00 + b11 * b11 + b22 * b22 + T(2.0) * off) * (T(1.0) / T(6.0)));
    const T c00 = b11 * b22 - a12 * a12;
    const T c01 = a01 * b22 - a12 * a02;
    const T c02 = a01 * a12 - b11 * a02;
    const T det = (b00 * c00 - a01 * c01 + a02 * c02) / (p * p * p);
    const T hd = fmin(fmax(T(0.5) * det, -T(1.0)), T(1.0));
    const T ang = acos(hd) * (T(1.0) / T(3.0));
    // separated root: the largest when hd >= 0, else the smallest
    const T beta = hd >= T(0.0) ? T(2.0) * cos(ang) : T(2.0) * cos(ang + T(2.0943951023931955));
    const T ev = q + p * beta;
    T r0[3] = {a00 - ev, a01, a02}, r1[3] = {a01, a11 - ev, a12}, r2[3] = {a02, a12, a22 - ev};
    T x01[3] = {r0[1] * r1[2] - r0[2] * r1[1], r0[2] * r1[0] - r0[0] * r1[2], r0[0] * r1[1] - r0[1] * r1[0]};
    T x02[3] = {r0[1] * r2[2] - r0[2] * r2[1], r0[2] * r2[0] - r0[0] * r2[2], r0[0] * r2[1] - r0[1] * r2[0]};
    T x12[3] = {r1[1] * r2[2] - r1[2] * r2[1], r1[2] * r2[0] - r1[0] * r2[2], r1[0] * r2[1] - r1[1] * r2[0]};
    T d0 = x01[0] * x01[0] + x01[1] * x01[1] + x01[2] * x01[2];
    T d1 = x02[0] * x02[0] + x02[1] * x02[1] + x02[2] * x02[2];
    T d2 = x12[0] * x12[0] + x12[1] * x12[1] + x12[2] * x12[2];
    T w[3];
    if (!(fmax(d0, fmax(d1, d2)) > eps2 * eps2)) {
        vals[0] = a00 * mx; vals[1] = a11 * mx; vals[2] = a22 * mx;
#pragma unroll
        for (int k = 0; k < 9; ++k) vecs[k] = (k == 0 || k == 4 || k == 8) ? T(1.0) : T(0.0);
        return;
    }
    if (d0 >= d1 && d0 >= d2) {
        T s = rsqrt_t(d0);
        w[0] = x01[0] * s; w[1] = x01[1] * s; w[2] = x01[2] * s;
    } else if (d1 >= d2) {
        T s = rsqrt_t(d1);
        w[0] = x02[0] * s; w[1] = x02[1] * s; w[2] = x02[2] * s;
    } else {
        T s = rsqrt_t(d2);
        w[0] = x12[0] * s; w[1] = x12[1] * s; w[2] = x12[2] * s;
    }
    T u[3];
    if (fabs(w[0]) > fabs(w[1])) {
        T s = rsqrt_t(w[0] * w[0] + w[2] * w[2]);
        u[0] = -w[2] * s; u[1] = T(0.0); u[2] = w[0] * s;
    } else {
        T s = rsqrt_t(w[1] * w[1] + w[2] * w[2]);
        u[0] = T(0.0); u[1] = w[2] * s; u[2] = -w[1] * s;
    }
    T v[3] = {w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2], w[0] * u[1] - w[1] * u[0]};
    const T am[9] = {a00, a01, a02, a01, a11, a12, a02, a12, a22};
    T au[3], av[3], aw[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        au[i] = am[3 * i] * u[0] + am[3 * i + 1] * u[1] + am[3 * i + 2] * u[2];
        av[i] = am[3 * i] * v[0] + am[3 * i + 1] * v[1] + am[3 * i + 2] * v[2];
        aw[i] = am[3 * i] * w[0] + am[3 * i + 1] * w[1] + am[3 * i + 2] * w[2];
    }
    const T m00 = u[0] * au[0] + u[1] * au[1] + u[2] * au[2];
    const T m01 = u[0] * av[0] + u[1] * av[1] + u[2] * av[2];
    const T m11 = v[0] * av[0] + v[1] * av[1] + v[2] * av[2];
    const T th = T(0.5) * atan2(T(2.0) * m01, m00 - m11);
    T sn, cs;
    sn = sin(th);
    cs = cos(th);
    T p1[3], p2[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        p1[k] = u[k] * cs + v[k] * sn;
        p2[k] = v[k] * cs - u[k] * sn;
    }
    // Rayleigh quotients; cs^2 m00 + 2 sn cs m01 + sn^2 m11 etc.
    const T l0 = w[0] * aw[0] + w[1] * aw[1] + w[2] * aw[2];
    const T l1 = cs * cs * m00 + T(2.0) * sn * cs * m01 + sn * sn * m11;
    const T l2 = sn * sn * m00 - T(2.0) * sn * cs * m01 + cs * cs * m11;
    vals[0] = l0 * mx; vals[1] = l1 * mx; vals[2] = l2 * mx;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        vecs[k] = w[k];
        vecs[3 + k] = p1[k];
        vecs[6 + k] = p2[k];
    }
}

CRUSH_HD void sym_eigen3f(const float* a, float* vals, float* vecs) { sym_eigen3<float>(a, vals, vecs); }

// A 3x3 matrix passed and returned by value (in registers).
template <class T>
struct Mat3v {
    T a[9];
};

// Large-strain Hencky strain eps = ln(F F^T)/2 from principal stretches in
// f64, where the smallest eigenvalues of b keep their relative accuracy.
// Kept out of line: rare, and its f64 temporaries would otherwise raise the
// register budget of every caller.  It takes and returns its matrices by
// value: a caller array whose address went to an out-of-line call would
// live in scratch memory on every path.
template <class T>
__host__ __device__ __attribute__((noinline)) inline Mat3v<T> hencky_log_strain_large(const Mat3v<T> hv) {
#pragma clang fp contract(off)
    const T* hd = hv.a;
    Mat3v<T> ev;
    T* eps = ev.a;
    double bd[9], lam[3], vec[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double h0 = hd[3 * i], h1 = hd[3 * i + 1], h2 = hd[3 * i + 2];
            bd[3 * i + j] = (double)hd[3 * i + j] + (double)hd[3 * j + i] +
                            (h0 * hd[3 * j] + h1 * hd[3 * j + 1] + h2 * hd[3 * j + 2]);
        }
    sym_eigen3<double>(bd, lam, vec);
    double e3[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) e3[a] = 0.5 * log1p(fmax(lam[a], -1.0 + 1e-300));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            eps[3 * i + j] = (T)(e3[0] * vec[i] * vec[j] + e3[1] * vec[3 + i] * vec[3 + j] + e3[2] * vec[6 + i] * vec[6 + j]);
    return ev;
}

CRUSH_HD float fmt(float a, float b, float c) { return fmaf(a, b, c); }
CRUSH_HD double fmt(double a, double b, double c) { return fma(a, b, c); }

template <class T>
CRUSH_HD void mul3t(const T* a, const T* b, T* r) {
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r[3 * i + j] = fmt(a[3 * i + 2], b[6 + j], fmt(a[3 * i + 1], b[3 + j], a[3 * i] * b[j]));
}

// Hencky elastic Kirchhoff stress (src/material.rs:708-771 with no
// plasticity): tau = K tr(eps) I + 2G dev(eps) with eps = ln(b)/2, b = F F^T.
// The small quantity A = b - I = Hd + Hd^T + Hd Hd^T is formed from the
// displacement gradient Hd = F - I.  For |A| below a threshold (0.05 in f32,
// 0.01 in f64: every elastic element of a vehicle model in practice)
// eps = ln(I + A)/2 is the matrix series to A^6 (truncation below 2e-10 |A|
// in f32, 2e-13 |A| in f64); otherwise principal stretches from an f64
// eigen-decomposition with eps_a = log1p(mu_a)/2.
template <class T>
CRUSH_HD void hencky_elastic_t(const MatParams& mp, const T* hd, T* tau) {
#pragma clang fp contract(off)
    T bm[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            bm[3 * i + j] = hd[3 * i + j] + hd[3 * j + i] +
                            fmt(hd[3 * i + 2], hd[3 * j + 2], fmt(hd[3 * i + 1], hd[3 * j + 1], hd[3 * i] * hd[3 * j]));
    T mx = T(0);
#pragma unroll
    for (int k = 0; k < 9; ++k) mx = fmax(mx, fabs(bm[k]));
    T eps[9];
    if (mx < (sizeof(T) == 4 ? T(0.05) : T(0.01))) {
        // ln(I + A) = A (I - A (I/2 - A (I/3 - A (I/4 - A (I/5 - A/6)))))
        const T c[6] = {T(1), T(0.5), T(1) / T(3), T(0.25), T(0.2), T(1) / T(6)};
        T acc[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) acc[k] = fmt(-c[5], bm[k], (k == 0 || k == 4 || k == 8) ? c[4] : T(0));
#pragma unroll
        for (int n = 3; n >= 0; --n) {
            T t[9];
            mul3t(bm, acc, t);
#pragma unroll
            for (int k = 0; k < 9; ++k) acc[k] = ((k == 0 || k == 4 || k == 8) ? c[n] : T(0)) - t[k];
        }
        // acc = I - A/2 + A^2/3 - ...; ln(I + A) = A acc
        T l[9];
        mul3t(bm, acc, l);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) eps[3 * i + j] = T(0.25) * (l[3 * i + j] + l[3 * j + i]);
    } else {
        Mat3v<T> hv;
#pragma unroll
        for (int k = 0; k < 9; ++k) hv.a[k] = hd[k];
        const Mat3v<T> ev = hencky_log_strain_large<T>(hv);
#pragma unroll
        for (int k = 0; k < 9; ++k) eps[k] = ev.a[k];
    }
    const T vol = eps[0] + eps[4] + eps[8];
    const T pressure = (T)mp.bulk * vol;
    const T g2 = T(2) * (T)mp.shear;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const bool dg = (k == 0 || k == 4 || k == 8);
        tau[k] = fmt(g2, eps[k] - (dg ? vol * (T(1) / T(3)) : T(0)), dg ? pressure : T(0));
    }
}

CRUSH_HD void hencky_elastic(const MatParams& mp, const float* hd, float* tau) { hencky_elastic_t<float>(mp, hd, tau); }

// Nodal forces of one element from tau (src/kernel.rs:141-154,
// src/kernel_simd.rs:445-484): f_i = tau F^-T grad_i V + hourglass.
// dl[8][3]: displacements relative to node 0; g, gam, hs read by the
// caller's accessors.
template <class GradF, class GamF, class HF>
CRUSH_HD void element_nodal_forces(const float* tau, const float* F, float J, float vol,
                                                     const float (*dl)[3], GradF grad, GamF gam, HF hsym,
                                                     float (*fo)[3]) {
#pragma clang fp contract(off)
    float finv[9];
    inv3f(F, J, finv);
    // M = tau F^-T V
    float M[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            M[3 * i + j] = vol * dot3f(tau[3 * i], finv[3 * j], tau[3 * i + 1], finv[3 * j + 1], tau[3 * i + 2], finv[3 * j + 2]);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float g0 = grad(3 * i), g1 = grad(3 * i + 1), g2 = grad(3 * i + 2);
#pragma unroll
        for (int a = 0; a < 3; ++a) fo[i][a] = dot3f(M[3 * a], g0, M[3 * a + 1], g1, M[3 * a + 2], g2);
    }
    // hourglass: q = U Gamma, r = H q, f += r Gamma^T (node 0 has dl = 0)
    float q[12];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            float s = 0.0f;
#pragma unroll
            for (int i = 1; i < 8; ++i) s = fmf(dl[i][d], gam(4 * i + a), s);
            q[3 * a + d] = s;
        }
    float rr[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) rr[k] = 0.0f;
    int kk = 0;
#pragma unroll
    for (int r = 0; r < 12; ++r) {
#pragma unroll
        for (int c = r; c < 12; ++c) {
            const float h = hsym(kk++);
            rr[r] = fmf(h, q[c], rr[r]);
            if (c != r) rr[c] = fmf(h, q[r], rr[c]);
        }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            float s = fo[i][d];
#pragma unroll
            for (int a = 0; a < 4; ++a) s = fmf(rr[3 * a + d], gam(4 * i + a), s);
            fo[i][d] = s;
        }
}

// ------------------------------------------------------ contact (f64)

// The reference's acceptance test of node p against one face
// (src/contact.rs:116-128): AABB reject, signed distance window, both
// diagonal splits with the 10 % barycentric tolerance.  Returns true and
// the signed distance when accepted.  Compiled without FMA contraction so
// the decision is the IEEE f64 result of the reference's expression.
CRUSH_HD bool point_in_tri(const double* p, const double* a, const double* b, const double* c) {
#pragma clang fp contract(off)
    double v0[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double v1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double v2[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    double d00 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2];
    double d01 = v0[0] * v1[0] + v0[1] * v1[1] + v0[2] * v1[2];
    double d02 = v0[0] * v2[0] + v0[1] * v2[1] + v0[2] * v2[2];
    double d11 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
    double d12 = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
    double inv = 1.0 / (d00 * d11 - d01 * d01);
    double u = (d11 * d02 - d01 * d12) * inv;
    double v = (d00 * d12 - d01 * d02) * inv;
    return u >= -C_TRI_TOL && v >= -C_TRI_TOL && u + v <= 1.0 + C_TRI_TOL;
}

CRUSH_HD bool face_accepts(const double* __restrict__ fg, int F, int f, const double* p,
                                             double max_d, double* signed_out) {
#pragma clang fp contract(off)
    auto at = [&](int k) { return fg[(size_t)k * F + f]; };
    if (p[0] < at(FG_LO) || p[0] > at(FG_HI) || p[1] < at(FG_LO + 1) || p[1] > at(FG_HI + 1) ||
        p[2] < at(FG_LO + 2) || p[2] > at(FG_HI + 2))
        return false;
    const double n[3] = {at(FG_N), at(FG_N + 1), at(FG_N + 2)};
    const double cp[3] = {at(FG_C) - p[0], at(FG_C + 1) - p[1], at(FG_C + 2) - p[2]};
    const double sd = -(n[0] * cp[0] + n[1] * cp[1] + n[2] * cp[2]);
    if (sd > 0.0 || sd < -max_d) return false;
    const double pr[3] = {p[0] - n[0] * sd, p[1] - n[1] * sd, p[2] - n[2] * sd};
    double x[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) x[k][d] = at(3 * k + d);
    bool inside = point_in_tri(pr, x[0], x[1], x[2]) || point_in_tri(pr, x[0], x[2], x[3]) ||
                  point_in_tri(pr, x[1], x[2], x[3]) || point_in_tri(pr, x[1], x[3], x[0]);
    *signed_out = sd;
    return inside;
}

// Face geometry for the current configuration (src/contact.rs:95-110) plus
// the candidate box used by the bucket search: a superset of every point
// the acceptance test can accept.  For each of the four triangles the
// region of projected points it accepts is the barycentric-enlarged
// triangle mapped onto the face plane along the triangle normal (an affine
// map, so the corners suffice); the node itself lies up to max_distance
// behind that.  Ill-conditioned faces fall back to the inflated AABB, which
// the acceptance test applies first anyway.  In three parts, so that the
// four triangles can also be formed on four lanes: face_geometry_base (the
// corners, normal, centre, inflated AABB and size; returns whether the
// normal is usable), face_triangle_region (triangle t's part of the tight
// region, min/max-accumulated into tlo/thi; returns false when the triangle
// is degenerate or too oblique) and face_geometry_finish (the candidate
// box).  Min and max do not depend on the order the triangles are taken in.
CRUSH_HD bool face_geometry_base(const double (*x)[3], double max_d, double* out /* [FG_BLO] */) {
#pragma clang fp contract(off)
    double e1[3], e2[3], n[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        e1[d] = x[1][d] - x[0][d];
        e2[d] = x[2][d] - x[0][d];
    }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) out[3 * k + d] = x[k][d];
    double size = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        n[d] = n[d] / nn;
        out[FG_N + d] = n[d];
        out[FG_C + d] = (x[0][d] + x[1][d] + x[2][d] + x[3][d]) / 4.0;
        double lo = fmin(fmin(x[0][d], x[1][d]), fmin(x[2][d], x[3][d]));
        double hi = fmax(fmax(x[0][d], x[1][d]), fmax(x[2][d], x[3][d]));
        out[FG_LO + d] = lo - max_d;
        out[FG_HI + d] = hi + max_d;
        size = fmax(size, hi - lo);
    }
    out[FG_SZ] = size;
    return __builtin_isfinite(nn) && nn > 0.0;
}

CRUSH_HD bool face_triangle_region(const double (*x)[3], const double* out, double max_d, int t, double* tlo,
                                   double* thi) {
#pragma clang fp contract(off)
    const int tri[4][3] = {{0, 1, 2}, {0, 2, 3}, {1, 2, 3}, {1, 3, 0}};
    const double cu[3] = {-C_TRI_TOL, 1.0 + 2.0 * C_TRI_TOL, -C_TRI_TOL};
    const double cv[3] = {-C_TRI_TOL, -C_TRI_TOL, 1.0 + 2.0 * C_TRI_TOL};
    const double n[3] = {out[FG_N], out[FG_N + 1], out[FG_N + 2]};
    const double* a = x[tri[t][0]];
    const double* b = x[tri[t][1]];
    const double* c = x[tri[t][2]];
    double v0[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double v1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double nt[3] = {v1[1] * v0[2] - v1[2] * v0[1], v1[2] * v0[0] - v1[0] * v0[2], v1[0] * v0[1] - v1[1] * v0[0]};
    // the triangle is degenerate when |nt| <= 1e-6 |v0| |v1|, and too oblique
    // to the face when |n . nt| <= 0.05 |nt|: both tested on squares, so
    // without square roots (a region only widens the search, and at the
    // boundary either answer leaves the box a superset of what the face
    // accepts)
    const double ntn2 = nt[0] * nt[0] + nt[1] * nt[1] + nt[2] * nt[2];
    const double l02 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2];
    const double l12 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
    if (!(ntn2 > 1e-12 * l02 * l12)) return false;
    const double dot = n[0] * nt[0] + n[1] * nt[1] + n[2] * nt[2];
    if (!(dot * dot > 0.0025 * ntn2)) return false;
    for (int k = 0; k < 3; ++k) {
        double ep[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) ep[d] = a[d] + cu[k] * v0[d] + cv[k] * v1[d];
        // the corner's projection onto the face plane along nt: ep + tt nt
        double tt = (n[0] * (out[FG_C] - ep[0]) + n[1] * (out[FG_C + 1] - ep[1]) + n[2] * (out[FG_C + 2] - ep[2])) / dot;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double q = ep[d] + tt * nt[d];
            double q2 = q - max_d * n[d];
            tlo[d] = fmin(tlo[d], fmin(q, q2));
            thi[d] = fmax(thi[d], fmax(q, q2));
        }
    }
    return true;
}

CRUSH_HD void face_geometry_finish(bool valid, const double* tlo, const double* thi, double* out) {
#pragma clang fp contract(off)
    const double margin = 1e-6 * out[FG_SZ] + 1e-12;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double lo = out[FG_LO + d], hi = out[FG_HI + d];
        if (valid) {
            lo = fmax(lo, tlo[d] - margin);
            hi = fmin(hi, thi[d] + margin);
        }
        out[FG_CLO + d] = lo;
        out[FG_CHI + d] = hi;
    }
}

CRUSH_HD void face_geometry(const double (*x)[3], double max_d, double* out /* [FG_BLO] */) {
    double tlo[3] = {INFINITY, INFINITY, INFINITY}, thi[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool valid = face_geometry_base(x, max_d, out);
    for (int t = 0; t < 4 && valid; ++t) valid = face_triangle_region(x, out, max_d, t, tlo, thi);
    face_geometry_finish(valid, tlo, thi, out);
}

// V / A_max of a hexahedron from 8 positions (src/element.rs:176-186), f64.
CRUSH_HD void char_length_d(const double (*p)[3], double* length, double* volume) {
#pragma clang fp contract(off)
    const int tets[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};
    const int faces[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};
    double vol = 0.0;
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        double u[3], v[3], w[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            u[d] = p[tets[t][1]][d] - p[tets[t][0]][d];
            v[d] = p[tets[t][2]][d] - p[tets[t][0]][d];
            w[d] = p[tets[t][3]][d] - p[tets[t][0]][d];
        }
        double cx = u[1] * v[2] - u[2] * v[1], cy = u[2] * v[0] - u[0] * v[2], cz = u[0] * v[1] - u[1] * v[0];
        vol += (cx * w[0] + cy * w[1] + cz * w[2]) / 6.0;
    }
    // the largest face area: x -> 0.5 sqrt(x) is non-decreasing, so the
    // largest of the areas is 0.5 sqrt of the largest squared norm, exactly
    double nmax = 0.0;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        double a[3], b[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            a[d] = p[faces[f][2]][d] - p[faces[f][0]][d];
            b[d] = p[faces[f][3]][d] - p[faces[f][1]][d];
        }
        double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
        nmax = fmax(nmax, cx * cx + cy * cy + cz * cz);
    }
    const double amax = 0.5 * sqrt(nmax);
    *volume = vol;
    *length = (vol <= 0.0 || amax <= 0.0) ? INFINITY : vol / amax;
}

}  // namespace crush
