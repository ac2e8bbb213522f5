// Test-only host build of the product's f32 element math
// (crushrs_amd/csrc/device_common.hpp), so the per-element arithmetic the
// GPU kernels run can be checked against the f64 oracle on the CPU.
#include <hip/hip_runtime.h>

#include "../../crushrs_amd/csrc/device_common.hpp"

using namespace crush;

static MatParams make(float youngs, float nu, const float* knots, float trans) {
    MatParams mp{};
    mp.youngs = youngs;
    mp.bulk = youngs / (3.0f * (1.0f - 2.0f * nu));
    mp.shear = youngs / (2.0f * (1.0f + nu));
    mp.trans = trans;
    for (int i = 0; i < 8; ++i) {
        mp.knot_c[i] = knots ? knots[3 * i] : INFINITY;
        mp.knot_s[i] = knots ? knots[3 * i + 1] : 0.0f;
        mp.knot_h[i] = knots ? knots[3 * i + 2] : 0.0f;
    }
    return mp;
}

extern "C" void hk_hencky(const float* hd, float youngs, float nu, float* tau) {
    MatParams mp = make(youngs, nu, nullptr, 1.0f);
    hencky_elastic(mp, hd, tau);
}

extern "C" void hk_honeycomb(const float* hd, float youngs, const float* knots, float trans, float* st, float* tau) {
    MatParams mp = make(youngs, 0.0f, knots, trans);
    honeycomb_update(mp, hd, st, tau);
}

extern "C" void hk_eigen(const float* a, float* vals, float* vecs) { sym_eigen3f(a, vals, vecs); }

// The stable step's characteristic length and volume of n hexes (p: 24
// doubles each).
extern "C" void hk_char_length(const double* p, int n, double* length, double* volume) {
    for (int e = 0; e < n; ++e) {
        double q[8][3];
        for (int i = 0; i < 8; ++i)
            for (int d = 0; d < 3; ++d) q[i][d] = p[24 * (size_t)e + 3 * i + d];
        char_length_d(q, length + e, volume + e);
    }
}

// Face geometry of n faces (x: 12 doubles each) into g (FG_BLO doubles
// each), then for each of m points per face (p: 3 doubles each) whether the
// acceptance test accepts it (acc) and whether it lies in the f64 candidate
// box (in64) and in the f32 box record (in32).
extern "C" void hk_face_check(const double* x, int n, double max_d, const double* p, int m, double* g, int* acc,
                              int* in64, int* in32) {
    for (int f = 0; f < n; ++f) {
        double q[4][3];
        for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) q[k][d] = x[12 * (size_t)f + 3 * k + d];
        double* gf = g + (size_t)FG_BLO * f;
        face_geometry(q, max_d, gf);
        for (int j = 0; j < m; ++j) {
            const double* pp = p + 3 * ((size_t)f * m + j);
            double sd;
            const size_t o = (size_t)f * m + j;
            acc[o] = face_accepts(gf, 1, 0, pp, max_d, &sd) ? 1 : 0;
            bool a = true, b = true;
            for (int d = 0; d < 3; ++d) {
                a = a && pp[d] >= gf[FG_CLO + d] && pp[d] <= gf[FG_CHI + d];
                b = b && pp[d] >= (double)f32_down(gf[FG_CLO + d]) && pp[d] <= (double)f32_up(gf[FG_CHI + d]);
            }
            in64[o] = a ? 1 : 0;
            in32[o] = b ? 1 : 0;
        }
    }
}

// The search's outward f32 rounding of candidate boxes, over n values.
extern "C" void hk_f32_round(const double* x, int n, float* down, float* up) {
    for (int i = 0; i < n; ++i) {
        down[i] = f32_down(x[i]);
        up[i] = f32_up(x[i]);
    }
}

// The batched/one-model force_element (crushrs_amd/csrc/phases.hpp) restated
// over explicit per-element arrays, as an element hook for the oracle's
// solve loop (oracle_set_element_hook): relative displacements cast to f32,
// F = I + sum_i (u_i - u_0) (x) grad_i, hourglass r = H q, the f32 material
// update and f_i = tau F^-T grad_i V + r gamma_i^T.
#include <vector>

struct HookCtx {
    int E = 0;
    std::vector<float> grad, gam, hs, vol, st;
    std::vector<MatParams> mp;
    std::vector<uint8_t> on;  // elements this hook computes (others: -1)
};

extern "C" void* hk_hook_create(int n_elems) {
    HookCtx* c = new HookCtx;
    c->E = n_elems;
    c->grad.assign(24 * (size_t)n_elems, 0.0f);
    c->gam.assign(32 * (size_t)n_elems, 0.0f);
    c->hs.assign(78 * (size_t)n_elems, 0.0f);
    c->vol.assign(n_elems, 0.0f);
    c->st.assign(25 * (size_t)n_elems, 0.0f);
    c->mp.resize(n_elems);
    c->on.assign(n_elems, 1);
    for (int e = 0; e < n_elems; ++e)
        for (int k : {9, 13, 17}) c->st[25 * (size_t)e + k] = 1.0f;
    return c;
}

extern "C" void hk_hook_free(void* ctx) { delete static_cast<HookCtx*>(ctx); }

// H144 row-major (already scaled by E), knots[24] = [c, sigma, slope] x 8.
extern "C" void hk_hook_set(void* ctx, int e, const double* grad, double vol, const double* gamma, const double* H144,
                            float youngs, float nu, const double* knots, float trans) {
    HookCtx* c = static_cast<HookCtx*>(ctx);
    for (int k = 0; k < 24; ++k) c->grad[24 * (size_t)e + k] = (float)grad[k];
    for (int k = 0; k < 32; ++k) c->gam[32 * (size_t)e + k] = (float)gamma[k];
    int kk = 0;
    for (int r = 0; r < 12; ++r)
        for (int q = r; q < 12; ++q) c->hs[78 * (size_t)e + kk++] = (float)(0.5 * (H144[12 * r + q] + H144[12 * q + r]));
    c->vol[e] = (float)vol;
    float kf[24];
    for (int k = 0; k < 24; ++k) kf[k] = (float)knots[k];
    c->mp[e] = make(youngs, nu, kf, trans);
}

extern "C" void hk_hook_enable(void* ctx, int e, int on) { static_cast<HookCtx*>(ctx)->on[e] = (uint8_t)(on != 0); }

extern "C" int hk_hook(void* ctx, int e, int kind, const double* u, double* out) {
    HookCtx* c = static_cast<HookCtx*>(ctx);
    if (!c->on[e]) return -1;
    const float* g = &c->grad[24 * (size_t)e];
    const float* ga = &c->gam[32 * (size_t)e];
    const float* hs = &c->hs[78 * (size_t)e];
    float hd[9] = {0}, q[12] = {0};
    for (int i = 1; i < 8; ++i) {
        const float d[3] = {(float)(u[3 * i] - u[0]), (float)(u[3 * i + 1] - u[1]), (float)(u[3 * i + 2] - u[2])};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) hd[3 * a + b] += d[a] * g[3 * i + b];
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 3; ++k) q[3 * a + k] += d[k] * ga[4 * i + a];
    }
    float F[9];
    for (int k = 0; k < 9; ++k) F[k] = hd[k] + ((k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f);
    const float J = det3f(F);
    if (!(J > 0.0f)) {
        for (int k = 0; k < 24; ++k) out[k] = 0.0;
        return 0;
    }
    float rr[12] = {0};
    int kk = 0;
    for (int r = 0; r < 12; ++r)
        for (int cc = r; cc < 12; ++cc) {
            const float h = hs[kk++];
            rr[r] += h * q[cc];
            if (cc != r) rr[cc] += h * q[r];
        }
    float tau[9];
    if (kind == CRUSH_MAT_HONEYCOMB)
        honeycomb_update(c->mp[e], hd, &c->st[25 * (size_t)e], tau);
    else
        hencky_elastic(c->mp[e], hd, tau);
    float finv[9], M[9];
    inv3f(F, J, finv);
    const float vol = c->vol[e];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            M[3 * i + j] = vol * (tau[3 * i] * finv[3 * j] + tau[3 * i + 1] * finv[3 * j + 1] + tau[3 * i + 2] * finv[3 * j + 2]);
    for (int i = 0; i < 8; ++i)
        for (int d = 0; d < 3; ++d) {
            float f = M[3 * d] * g[3 * i] + M[3 * d + 1] * g[3 * i + 1] + M[3 * d + 2] * g[3 * i + 2];
            for (int a = 0; a < 4; ++a) f += rr[3 * a + d] * ga[4 * i + a];
            out[3 * i + d] = f;
        }
    return 1;
}

// The same element force in f64 arithmetic over the f32-stored constants
// (elastic only): relative displacements, F, hourglass, Hencky stress and
// nodal forces in double; round_out rounds each force to f32 as an f32
// per-corner force array would.
struct Hook64 {
    HookCtx* c;
    int round_out;
};

extern "C" void* hk_hook64_create(void* ctx, int round_out) { return new Hook64{static_cast<HookCtx*>(ctx), round_out}; }
extern "C" void hk_hook64_free(void* h) { delete static_cast<Hook64*>(h); }

extern "C" int hk_hook64(void* hp, int e, int kind, const double* u, double* out) {
    Hook64* h = static_cast<Hook64*>(hp);
    HookCtx* c = h->c;
    if (!c->on[e] || kind != CRUSH_MAT_ELASTIC) return -1;
    const float* g = &c->grad[24 * (size_t)e];
    const float* ga = &c->gam[32 * (size_t)e];
    const float* hs = &c->hs[78 * (size_t)e];
    double hd[9] = {0}, q[12] = {0};
    for (int i = 1; i < 8; ++i) {
        const double d[3] = {u[3 * i] - u[0], u[3 * i + 1] - u[1], u[3 * i + 2] - u[2]};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) hd[3 * a + b] += d[a] * (double)g[3 * i + b];
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 3; ++k) q[3 * a + k] += d[k] * (double)ga[4 * i + a];
    }
    double F[9];
    for (int k = 0; k < 9; ++k) F[k] = hd[k] + ((k == 0 || k == 4 || k == 8) ? 1.0 : 0.0);
    const double J = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) + F[2] * (F[3] * F[7] - F[4] * F[6]);
    if (!(J > 0.0)) {
        for (int k = 0; k < 24; ++k) out[k] = 0.0;
        return 0;
    }
    double rr[12] = {0};
    int kk = 0;
    for (int r = 0; r < 12; ++r)
        for (int cc = r; cc < 12; ++cc) {
            const double hv = hs[kk++];
            rr[r] += hv * q[cc];
            if (cc != r) rr[cc] += hv * q[r];
        }
    double tau[9];
    hencky_elastic_t<double>(c->mp[e], hd, tau);
    double finv[9];
    const double id = 1.0 / J;
    finv[0] = (F[4] * F[8] - F[5] * F[7]) * id;
    finv[1] = (F[2] * F[7] - F[1] * F[8]) * id;
    finv[2] = (F[1] * F[5] - F[2] * F[4]) * id;
    finv[3] = (F[5] * F[6] - F[3] * F[8]) * id;
    finv[4] = (F[0] * F[8] - F[2] * F[6]) * id;
    finv[5] = (F[2] * F[3] - F[0] * F[5]) * id;
    finv[6] = (F[3] * F[7] - F[4] * F[6]) * id;
    finv[7] = (F[1] * F[6] - F[0] * F[7]) * id;
    finv[8] = (F[0] * F[4] - F[1] * F[3]) * id;
    double M[9];
    const double vol = c->vol[e];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            M[3 * i + j] = vol * (tau[3 * i] * finv[3 * j] + tau[3 * i + 1] * finv[3 * j + 1] + tau[3 * i + 2] * finv[3 * j + 2]);
    for (int i = 0; i < 8; ++i)
        for (int d = 0; d < 3; ++d) {
            double f = M[3 * d] * g[3 * i] + M[3 * d + 1] * g[3 * i + 1] + M[3 * d + 2] * g[3 * i + 2];
            for (int a = 0; a < 4; ++a) f += rr[3 * a + d] * (double)ga[4 * i + a];
            out[3 * i + d] = h->round_out ? (double)(float)f : f;
        }
    return 1;
}
