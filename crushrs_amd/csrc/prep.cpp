// Host-side preparation of a crush_model for the device.
#include "prep.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <unordered_map>

#include "common.hpp"

namespace crush {

namespace {

// C3D8 node signs (src/element.rs:9-18) and Flanagan-Belytschko base
// vectors (src/element.rs:83-88): the element's node numbering and its
// four hourglass modes.
const double SGN[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                          {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
const double HGM[4][8] = {{1, 1, -1, -1, -1, -1, 1, 1},
                          {1, -1, -1, 1, -1, 1, 1, -1},
                          {1, -1, 1, -1, 1, -1, 1, -1},
                          {-1, 1, -1, 1, 1, -1, 1, -1}};

bool inverse3(const double* m, double* r, double* det_out) {
    double c0 = m[4] * m[8] - m[5] * m[7];
    double c1 = m[5] * m[6] - m[3] * m[8];
    double c2 = m[3] * m[7] - m[4] * m[6];
    double det = m[0] * c0 + m[1] * c1 + m[2] * c2;
    if (det_out) *det_out = det;
    if (det == 0.0) return false;
    double id = 1.0 / det;
    r[0] = c0 * id;
    r[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    r[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    r[3] = c1 * id;
    r[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    r[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    r[6] = c2 * id;
    r[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    r[8] = (m[0] * m[4] - m[1] * m[3]) * id;
    return true;
}

// dN/dxi at (xi, eta, zeta) and the spatial gradient dN/dX = dN/dxi J^-T.
bool spatial_gradient(const double* x, double xi, double eta, double zeta, double g[8][3], double* det_j,
                      double n_out[8]) {
    double dn[8][3];
    for (int i = 0; i < 8; ++i) {
        const double* s = SGN[i];
        dn[i][0] = 0.125 * s[0] * (1.0 + s[1] * eta) * (1.0 + s[2] * zeta);
        dn[i][1] = 0.125 * s[1] * (1.0 + s[0] * xi) * (1.0 + s[2] * zeta);
        dn[i][2] = 0.125 * s[2] * (1.0 + s[0] * xi) * (1.0 + s[1] * eta);
        if (n_out) n_out[i] = 0.125 * (1.0 + s[0] * xi) * (1.0 + s[1] * eta) * (1.0 + s[2] * zeta);
    }
    double j[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, ji[9];
    for (int i = 0; i < 8; ++i)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) j[3 * a + b] += dn[i][a] * x[3 * i + b];
    if (!inverse3(j, ji, det_j)) return false;
    for (int i = 0; i < 8; ++i)
        for (int c = 0; c < 3; ++c) g[i][c] = dn[i][0] * ji[3 * c] + dn[i][1] * ji[3 * c + 1] + dn[i][2] * ji[3 * c + 2];
    return true;
}

// 4x4 inverse by Gauss-Jordan with partial pivoting.
bool inverse4(const double* a, double* r) {
    double w[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            w[i][j] = a[4 * i + j];
            w[i][4 + j] = (i == j) ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int i = c + 1; i < 4; ++i)
            if (std::fabs(w[i][c]) > std::fabs(w[p][c])) p = i;
        if (w[p][c] == 0.0) return false;
        if (p != c)
            for (int j = 0; j < 8; ++j) std::swap(w[p][j], w[c][j]);
        double piv = w[c][c];
        for (int j = 0; j < 8; ++j) w[c][j] /= piv;
        for (int i = 0; i < 4; ++i) {
            if (i == c || w[i][c] == 0.0) continue;
            double f = w[i][c];
            for (int j = 0; j < 8; ++j) w[i][j] -= f * w[c][j];
        }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r[4 * i + j] = w[i][4 + j];
    return true;
}

double wave_speed(const crush_material& m) {
    double e = m.youngs_modulus, v = m.poisson_ratio;
    double mm = e * (1.0 - v) / ((1.0 + v) * (1.0 - 2.0 * v));
    return std::sqrt(mm / m.density);
}

}  // namespace

// The stabilisation stiffness K = sum_gp (B_gp - B_c)^T C (B_gp - B_c) detJ of
// src/element.rs:141-161, kept only through its rank-12 projection
// H = (A^-1 (x) I) (G K G^T) (A^-1 (x) I), A = Gamma^T Gamma, which is what
// src/element.rs:91-114 computes with the full 12x24 G.  G K G^T is formed
// per Gauss point as P^T C P with P = (B_gp - B_c) G^T (6x12).
void hex_geometry(const double* x, double nu, double hourglass, double* grad, double* vol, double* lumped,
                  double* gamma, double* hsym) {
    const double gp = 1.0 / std::sqrt(3.0);
    double gc[8][3], detc;
    if (!spatial_gradient(x, 0.0, 0.0, 0.0, gc, &detc, nullptr)) fail(CRUSH_E_INPUT, "degenerate hexahedron");
    for (int i = 0; i < 8; ++i)
        for (int c = 0; c < 3; ++c) grad[3 * i + c] = gc[i][c];
    for (int a = 0; a < 4; ++a) {
        double hx[3] = {0, 0, 0};
        for (int j = 0; j < 8; ++j)
            for (int d = 0; d < 3; ++d) hx[d] += HGM[a][j] * x[3 * j + d];
        for (int i = 0; i < 8; ++i)
            gamma[4 * i + a] = HGM[a][i] - (hx[0] * gc[i][0] + hx[1] * gc[i][1] + hx[2] * gc[i][2]);
    }
    // isotropic C for E = hourglass (scaled by the material's E later)
    double e = hourglass;
    double lam = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double mu = e / (2.0 * (1.0 + nu));
    double S[12][12];
    std::memset(S, 0, sizeof S);
    double v = 0.0;
    for (int i = 0; i < 8; ++i) lumped[i] = 0.0;
    for (int q = 0; q < 8; ++q) {
        double g[8][3], dj, nsh[8];
        if (!spatial_gradient(x, SGN[q][0] * gp, SGN[q][1] * gp, SGN[q][2] * gp, g, &dj, nsh))
            fail(CRUSH_E_INPUT, "degenerate hexahedron (singular Jacobian)");
        v += dj;
        double nsum = 0.0;
        for (int i = 0; i < 8; ++i) nsum += nsh[i];
        for (int i = 0; i < 8; ++i) lumped[i] += nsh[i] * nsum * dj;
        // Qm[c][a] = sum_i (g_i - gc_i)[c] gamma_ia
        double Qm[3][4];
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 4; ++a) {
                double s = 0.0;
                for (int i = 0; i < 8; ++i) s += (g[i][c] - gc[i][c]) * gamma[4 * i + a];
                Qm[c][a] = s;
            }
        // P (6x12): rows xx, yy, zz, xy, yz, xz (engineering shear)
        double P[6][12];
        std::memset(P, 0, sizeof P);
        for (int a = 0; a < 4; ++a) {
            P[0][3 * a + 0] = Qm[0][a];
            P[1][3 * a + 1] = Qm[1][a];
            P[2][3 * a + 2] = Qm[2][a];
            P[3][3 * a + 0] = Qm[1][a];
            P[3][3 * a + 1] = Qm[0][a];
            P[4][3 * a + 1] = Qm[2][a];
            P[4][3 * a + 2] = Qm[1][a];
            P[5][3 * a + 0] = Qm[2][a];
            P[5][3 * a + 2] = Qm[0][a];
        }
        // CP = C P
        double CP[6][12];
        for (int j = 0; j < 12; ++j) {
            double p0 = P[0][j], p1 = P[1][j], p2 = P[2][j];
            CP[0][j] = (lam + 2 * mu) * p0 + lam * p1 + lam * p2;
            CP[1][j] = lam * p0 + (lam + 2 * mu) * p1 + lam * p2;
            CP[2][j] = lam * p0 + lam * p1 + (lam + 2 * mu) * p2;
            CP[3][j] = mu * P[3][j];
            CP[4][j] = mu * P[4][j];
            CP[5][j] = mu * P[5][j];
        }
        for (int r = 0; r < 12; ++r)
            for (int c = 0; c < 12; ++c) {
                double s = 0.0;
                for (int k = 0; k < 6; ++k) s += P[k][r] * CP[k][c];
                S[r][c] += s * dj;
            }
    }
    *vol = v;
    double A[16], Ai[16];
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            double s = 0.0;
            for (int i = 0; i < 8; ++i) s += gamma[4 * i + a] * gamma[4 * i + b];
            A[4 * a + b] = s;
        }
    if (!inverse4(A, Ai)) fail(CRUSH_E_INPUT, "hourglass modes are not independent");
    // H[3a+d][3b+e] = sum_{a',b'} Ai[a][a'] S[3a'+d][3b'+e] Ai[b'][b]
    double T[12][12];
    for (int a = 0; a < 4; ++a)
        for (int d = 0; d < 3; ++d)
            for (int c = 0; c < 12; ++c) {
                double s = 0.0;
                for (int ap = 0; ap < 4; ++ap) s += Ai[4 * a + ap] * S[3 * ap + d][c];
                T[3 * a + d][c] = s;
            }
    int k = 0;
    for (int r = 0; r < 12; ++r)
        for (int c = r; c < 12; ++c) {
            int b = c / 3, ee = c % 3;
            double s = 0.0;
            for (int bp = 0; bp < 4; ++bp) s += T[r][3 * bp + ee] * Ai[4 * bp + b];
            // symmetrise with the transposed entry
            int b2 = r / 3, e2 = r % 3;
            double s2 = 0.0;
            for (int bp = 0; bp < 4; ++bp) s2 += T[c][3 * bp + e2] * Ai[4 * bp + b2];
            hsym[k++] = 0.5 * (s + s2);
        }
}

// V / A_max with the 6-tet volume and the largest face area
// (src/element.rs:176-186).
void hex_char_length(const double* p, double* length, double* volume) {
    static const int TETS[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};
    static const int FACES[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};
    double vol = 0.0;
    for (const auto& t : TETS) {
        double u[3], v[3], w[3];
        for (int d = 0; d < 3; ++d) {
            u[d] = p[3 * t[1] + d] - p[3 * t[0] + d];
            v[d] = p[3 * t[2] + d] - p[3 * t[0] + d];
            w[d] = p[3 * t[3] + d] - p[3 * t[0] + d];
        }
        vol += ((u[1] * v[2] - u[2] * v[1]) * w[0] + (u[2] * v[0] - u[0] * v[2]) * w[1] +
                (u[0] * v[1] - u[1] * v[0]) * w[2]) /
               6.0;
    }
    double amax = 0.0;
    for (const auto& f : FACES) {
        double a[3], b[3];
        for (int d = 0; d < 3; ++d) {
            a[d] = p[3 * f[2] + d] - p[3 * f[0] + d];
            b[d] = p[3 * f[3] + d] - p[3 * f[1] + d];
        }
        double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
        amax = std::max(amax, 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz));
    }
    *volume = vol;
    *length = (vol <= 0.0 || amax <= 0.0) ? INFINITY : vol / amax;
}

static void make_material(const crush_material& m, MatParams& mp) {
    std::memset(&mp, 0, sizeof mp);
    mp.kind = m.model;
    mp.youngs = (float)m.youngs_modulus;
    mp.bulk = (float)(m.youngs_modulus / (3.0 * (1.0 - 2.0 * m.poisson_ratio)));
    mp.shear = (float)(m.youngs_modulus / (2.0 * (1.0 + m.poisson_ratio)));
    mp.trans = 1.0f;
    for (int i = 0; i < 8; ++i) {
        mp.knot_c[i] = INFINITY;
        mp.knot_s[i] = 0.0f;
        mp.knot_h[i] = 0.0f;
    }
    if (m.model != CRUSH_MAT_HONEYCOMB) return;
    // Knot table [c, sigma, slope of the following segment]
    // (src/material.rs:430-449): bilinear = one knot; tabulated = one knot
    // per curve point, the last slope continuing; densification adds a knot
    // at c_lock whose slope gains k_lock.
    int n = m.n_curve;
    bool dens = m.has_densification != 0;
    if (n == 1 || n < 0 || n + (dens ? 1 : 0) > CRUSH_MAX_KNOTS)
        fail(CRUSH_E_INPUT, "yield curve needs 2..=%d knots (one fewer with densification)", CRUSH_MAX_KNOTS);
    double kc[8], ks[8], kh[8];
    for (int i = 0; i < 8; ++i) {
        kc[i] = INFINITY;
        ks[i] = 0.0;
        kh[i] = 0.0;
    }
    int used;
    if (n < 2) {
        kc[0] = 0.0;
        ks[0] = m.yield_stress;
        kh[0] = m.hardening;
        used = 1;
    } else {
        if (m.curve[0][0] != 0.0) fail(CRUSH_E_INPUT, "yield curve must start at compaction 0");
        for (int i = 0; i < n; ++i) {
            if (i > 0 && !(m.curve[i][0] > m.curve[i - 1][0]))
                fail(CRUSH_E_INPUT, "yield curve compaction must be strictly increasing");
            if (!(m.curve[i][1] > 0.0)) fail(CRUSH_E_INPUT, "yield curve stresses must be positive");
        }
        for (int i = 0; i < n; ++i) {
            kc[i] = m.curve[i][0];
            ks[i] = m.curve[i][1];
            kh[i] = (i + 1 < n) ? (m.curve[i + 1][1] - m.curve[i][1]) / (m.curve[i + 1][0] - m.curve[i][0]) : kh[i - 1];
        }
        used = n;
    }
    if (dens) {
        double cl = m.densification[0], kl = m.densification[1];
        if (!(cl > 0.0) || kl < 0.0) fail(CRUSH_E_INPUT, "densification needs c_lock > 0 and k_lock >= 0");
        if (n >= 2 && m.curve[n - 1][0] >= cl)
            fail(CRUSH_E_INPUT, "densification c_lock must lie beyond the last curve knot");
        kc[used] = cl;
        ks[used] = ks[used - 1] + kh[used - 1] * (cl - kc[used - 1]);
        kh[used] = kh[used - 1] + kl;
    }
    for (int i = 0; i < 8; ++i) {
        mp.knot_c[i] = (float)kc[i];
        mp.knot_s[i] = (float)ks[i];
        mp.knot_h[i] = (float)kh[i];
    }
    mp.trans = (float)m.transverse_factor;
}

void prepare(const crush_model& m, Prepared& p) {
    if (m.n_nodes <= 0 || m.n_hexes <= 0) fail(CRUSH_E_INPUT, "model needs nodes and hexahedra");
    if (!m.nodes || !m.hexes || !m.hex_part || !m.materials || !m.initial_velocity)
        fail(CRUSH_E_INPUT, "model arrays missing");
    if (m.n_parts <= 0) fail(CRUSH_E_INPUT, "model has no parts");
    const crush_settings& s = m.settings;
    if (!(s.end_time > 0.0) || !(s.dt_scale > 0.0) || !(s.hourglass >= 0.0))
        fail(CRUSH_E_INPUT, "settings: end_time, dt_scale must be positive, hourglass non-negative");
    p = Prepared();
    p.settings = s;
    const int N = m.n_nodes, E = m.n_hexes;
    p.n_nodes = N;
    p.n_elems = E;
    p.n_parts = m.n_parts;
    // materials
    p.mats.resize(m.n_parts);
    for (int q = 0; q < m.n_parts; ++q) {
        const crush_material& mt = m.materials[q];
        if (mt.model != CRUSH_MAT_ELASTIC && mt.model != CRUSH_MAT_HONEYCOMB)
            fail(CRUSH_E_UNSUPPORTED, "part %d: material model %d is not on the GPU path (elastic, honeycomb)", q,
                 mt.model);
        if (!(mt.density > 0.0) || !(mt.youngs_modulus > 0.0))
            fail(CRUSH_E_INPUT, "part %d: density and Young's modulus must be positive", q);
        if (!(mt.poisson_ratio > -1.0 && mt.poisson_ratio < 0.5)) fail(CRUSH_E_INPUT, "part %d: bad Poisson ratio", q);
        make_material(mt, p.mats[q]);
    }
    // element slots: honeycomb first
    for (int e = 0; e < E; ++e) {
        int part = m.hex_part[e];
        if (part < 0 || part >= m.n_parts) fail(CRUSH_E_INPUT, "hexahedron %d: part %d out of range", e, part);
        for (int k = 0; k < 8; ++k) {
            int n = m.hexes[8 * e + k];
            if (n < 0 || n >= N) fail(CRUSH_E_INPUT, "hexahedron %d: node %d out of range", e, n);
        }
    }
    p.slot_elem.reserve(E);
    for (int pass = 0; pass < 2; ++pass)
        for (int e = 0; e < E; ++e) {
            bool hc = m.materials[m.hex_part[e]].model == CRUSH_MAT_HONEYCOMB;
            if ((pass == 0) == hc) p.slot_elem.push_back(e);
            if (pass == 0 && hc) p.n_hc++;
        }
    p.elem_slot.assign(E, 0);
    for (int sl = 0; sl < E; ++sl) p.elem_slot[p.slot_elem[sl]] = sl;
    p.conn.resize(8 * (size_t)E);
    for (int sl = 0; sl < E; ++sl)
        for (int k = 0; k < 8; ++k) p.conn[(size_t)k * E + sl] = m.hexes[8 * p.slot_elem[sl] + k];
    // geometry classes: equal part and equal node offsets (to 1e-12 m)
    struct Key {
        int part;
        long long d[21];
        bool operator<(const Key& o) const {
            if (part != o.part) return part < o.part;
            for (int i = 0; i < 21; ++i)
                if (d[i] != o.d[i]) return d[i] < o.d[i];
            return false;
        }
    };
    std::map<Key, int> classes;
    p.ecls.resize(E);
    std::vector<double> lumped_all(8 * (size_t)E);
    std::vector<double> cls_lumped;  // [8][C] unit density
    std::vector<double> cls_vol64;
    p.X.assign(m.nodes, m.nodes + 3 * (size_t)N);
    for (int sl = 0; sl < E; ++sl) {
        int e = p.slot_elem[sl];
        int part = m.hex_part[e];
        double x[24];
        for (int i = 0; i < 8; ++i)
            for (int d = 0; d < 3; ++d) x[3 * i + d] = m.nodes[3 * m.hexes[8 * e + i] + d];
        Key key;
        key.part = part;
        for (int i = 1; i < 8; ++i)
            for (int d = 0; d < 3; ++d) key.d[3 * (i - 1) + d] = std::llround((x[3 * i + d] - x[d]) * 1e12);
        auto it = classes.find(key);
        int c;
        if (it == classes.end()) {
            c = (int)classes.size();
            classes.emplace(key, c);
            const crush_material& mt = m.materials[part];
            double grad[24], vol, lump[8], gam[32], hs[HSYM];
            for (int i = 7; i >= 0; --i)
                for (int d = 0; d < 3; ++d) x[3 * i + d] -= x[d];  // local coordinates
            hex_geometry(x, mt.poisson_ratio, s.hourglass, grad, &vol, lump, gam, hs);
            if (!(vol > 0.0)) fail(CRUSH_E_INPUT, "hexahedron %d has non-positive volume", e);
            float* pk = &*p.cls_pack.insert(p.cls_pack.end(), CPK, 0.0f);
            for (int i = 0; i < 8; ++i) {
                for (int b = 0; b < 3; ++b) pk[CPK_NODE * i + CPK_GRAD + b] = (float)grad[3 * i + b];
                for (int a = 0; a < 4; ++a) pk[CPK_NODE * i + CPK_GAM + a] = (float)gam[4 * i + a];
            }
            for (int k = 0; k < HSYM; ++k) pk[CPK_H + k] = (float)(hs[k] * mt.youngs_modulus);
            pk[CPK_VOL] = (float)vol;
            cls_vol64.push_back(vol);
            p.cls_mat.push_back(part);
            p.cls_dtf.push_back((1.0 - 0.36 * mt.poisson_ratio) / wave_speed(mt));
            for (int k = 0; k < 8; ++k) cls_lumped.push_back(lump[k] * mt.density);
        } else {
            c = it->second;
        }
        p.ecls[sl] = c;
    }
    const int C = (int)classes.size();
    p.n_classes = C;
    // lumped masses and element masses (src/model.rs:268-278, solver.rs:199)
    p.mass.assign(N, 0.0);
    p.elem_mass.resize(E);
    for (int e = 0; e < E; ++e) {
        int sl = p.elem_slot[e];
        int c = p.ecls[sl];
        for (int i = 0; i < 8; ++i) p.mass[m.hexes[8 * e + i]] += cls_lumped[8 * (size_t)c + i];
        p.elem_mass[sl] = cls_vol64[c] * m.materials[m.hex_part[e]].density;
    }
    p.v0.assign(m.initial_velocity, m.initial_velocity + 3 * (size_t)N);
    p.fixed.assign(N, 0);
    for (int i = 0; i < m.n_fixed; ++i) {
        int n = m.fixed_nodes[i];
        if (n < 0 || n >= N) fail(CRUSH_E_INPUT, "fixed node %d out of range", n);
        p.fixed[n] |= NODE_FIXED;
    }
    // node -> element incidence in original element order
    p.n2e_ptr.assign(N + 1, 0);
    for (int e = 0; e < E; ++e)
        for (int k = 0; k < 8; ++k) p.n2e_ptr[m.hexes[8 * e + k] + 1]++;
    for (int n = 0; n < N; ++n) p.n2e_ptr[n + 1] += p.n2e_ptr[n];
    p.n2e_idx.resize(8 * (size_t)E);
    {
        std::vector<int32_t> cur(p.n2e_ptr.begin(), p.n2e_ptr.end() - 1);
        for (int e = 0; e < E; ++e)
            for (int k = 0; k < 8; ++k) p.n2e_idx[cur[m.hexes[8 * e + k]]++] = (p.elem_slot[e] << 3) | k;
    }
    // contacts: static broad phase (src/contact.rs:48-70)
    for (int c = 0; c < m.n_contacts; ++c) {
        const crush_contact& ct = m.contacts[c];
        if (ct.n_nodes < 0 || ct.n_faces < 0 || (ct.n_nodes > 0 && !ct.nodes) || (ct.n_faces > 0 && !ct.faces))
            fail(CRUSH_E_INPUT, "contact %d: malformed", c);
        if (!(ct.max_distance >= 0.0) || !(ct.stiffness >= 0.0))
            fail(CRUSH_E_INPUT, "contact %d: stiffness and max_distance must be non-negative", c);
        ContactDesc d{};
        d.node_begin = (int32_t)p.cn_node.size();
        d.face_begin = (int32_t)p.cf_contact.size();
        d.n_faces = ct.n_faces;
        d.stiffness = ct.stiffness;
        d.max_distance = ct.max_distance;
        d.depth_cap = 0.1 * ct.max_distance;
        d.pair = ct.pair;
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int f = 0; f < ct.n_faces; ++f) {
            int fi = ct.faces[f];
            if (fi < 0 || fi >= m.n_faces) fail(CRUSH_E_INPUT, "contact %d: face %d out of range", c, fi);
            for (int k = 0; k < 4; ++k) {
                int n = m.faces[4 * fi + k];
                if (n < 0 || n >= N) fail(CRUSH_E_INPUT, "face %d: node out of range", fi);
                for (int dd = 0; dd < 3; ++dd) {
                    lo[dd] = std::min(lo[dd], m.nodes[3 * n + dd]);
                    hi[dd] = std::max(hi[dd], m.nodes[3 * n + dd]);
                }
            }
            p.cf_contact.push_back(c);
        }
        for (int dd = 0; dd < 3; ++dd) {
            lo[dd] -= ct.max_distance;
            hi[dd] += ct.max_distance;
        }
        for (int i = 0; i < ct.n_nodes; ++i) {
            int n = ct.nodes[i];
            if (n < 0 || n >= N) fail(CRUSH_E_INPUT, "contact %d: node %d out of range", c, n);
            const double* q = m.nodes + 3 * n;
            if (q[0] >= lo[0] && q[1] >= lo[1] && q[2] >= lo[2] && q[0] <= hi[0] && q[1] <= hi[1] && q[2] <= hi[2]) {
                p.cn_node.push_back(n);
                p.cn_contact.push_back(c);
                p.cn_scale.push_back(ct.node_scale ? ct.node_scale[i] : 1.0);
            }
        }
        d.n_nodes = (int32_t)p.cn_node.size() - d.node_begin;
        p.cdesc.push_back(d);
    }
    p.n_cnodes = (int)p.cn_node.size();
    p.n_cfaces = (int)p.cf_contact.size();
    p.cf_nodes.resize(4 * (size_t)p.n_cfaces);
    for (int c = 0; c < m.n_contacts; ++c) {
        const crush_contact& ct = m.contacts[c];
        for (int f = 0; f < ct.n_faces; ++f)
            for (int k = 0; k < 4; ++k)
                p.cf_nodes[(size_t)k * p.n_cfaces + p.cdesc[c].face_begin + f] = m.faces[4 * ct.faces[f] + k];
    }
    // faces of fixed nodes only (rigid walls and roads) keep their geometry
    p.cf_static.assign((size_t)p.n_cfaces, 0);
    for (int f = 0; f < p.n_cfaces; ++f) {
        bool st = true;
        for (int k = 0; k < 4; ++k) st = st && (p.fixed[p.cf_nodes[(size_t)k * p.n_cfaces + f]] & NODE_FIXED);
        p.cf_static[f] = st ? 1 : 0;
    }
    for (ContactDesc& d : p.cdesc) {
        d.all_static = 1;
        for (int f = 0; f < d.n_faces; ++f) d.all_static &= p.cf_static[d.face_begin + f];
    }
    // apply lists: per touched node, in contact order: own record, then
    // the faces containing it
    {
        std::vector<std::vector<int32_t>> lists(N);
        for (int c = 0; c < m.n_contacts; ++c) {
            const ContactDesc& d = p.cdesc[c];
            for (int k = 0; k < d.n_nodes; ++k) lists[p.cn_node[d.node_begin + k]].push_back(d.node_begin + k);
            for (int f = 0; f < d.n_faces; ++f)
                for (int k = 0; k < 4; ++k)
                    lists[p.cf_nodes[(size_t)k * p.n_cfaces + d.face_begin + f]].push_back(-(d.face_begin + f) - 1);
        }
        p.tn_ptr.push_back(0);
        p.tn_of.assign(N, -1);
        for (int n = 0; n < N; ++n) {
            if (lists[n].empty()) continue;
            p.tn_of[n] = (int32_t)p.tn_node.size();
            p.tn_node.push_back(n);
            p.fixed[n] |= NODE_CONTACT;
            p.tn_ent.insert(p.tn_ent.end(), lists[n].begin(), lists[n].end());
            p.tn_ptr.push_back((int32_t)p.tn_ent.size());
        }
    }
    for (int c = 0; c < m.n_contacts; ++c)
        for (int st = 0; st < p.cdesc[c].n_nodes; st += SEARCH_BLOCK) {
            p.sblk_contact.push_back(c);
            p.sblk_start.push_back(st);
        }
    // time step (src/solver.rs:214-232)
    double contact_dt = INFINITY;
    for (int c = 0; c < m.n_contacts; ++c) {
        const crush_contact& ct = m.contacts[c];
        for (int i = 0; i < ct.n_nodes; ++i) {
            int n = ct.nodes[i];
            if (p.mass[n] > 0.0 && ct.stiffness > 0.0) contact_dt = std::min(contact_dt, 2.0 * std::sqrt(p.mass[n] / ct.stiffness));
        }
    }
    double stable = INFINITY;
    for (int e = 0; e < E; ++e) {
        double x[24], l, v;
        for (int i = 0; i < 8; ++i)
            for (int d = 0; d < 3; ++d) x[3 * i + d] = m.nodes[3 * m.hexes[8 * e + i] + d];
        hex_char_length(x, &l, &v);
        if (std::isfinite(l) && v > 0.0) stable = std::min(stable, l * p.cls_dtf[p.ecls[p.elem_slot[e]]]);
    }
    if (!std::isfinite(stable)) {
        static const int EDGES[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                         {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
        double min_edge = INFINITY, cmax = 0.0;
        for (int e = 0; e < E; ++e) {
            for (const auto& ed : EDGES) {
                const double* a = m.nodes + 3 * m.hexes[8 * e + ed[0]];
                const double* b = m.nodes + 3 * m.hexes[8 * e + ed[1]];
                min_edge = std::min(min_edge, std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                                                        (a[2] - b[2]) * (a[2] - b[2])));
            }
            cmax = std::max(cmax, wave_speed(m.materials[m.hex_part[e]]));
        }
        stable = min_edge / (std::sqrt(3.0) * cmax);
    }
    stable = std::min(stable, contact_dt);
    p.contact_dt = contact_dt;
    p.stable = stable;
    p.dt_floor = s.mass_scaling > 0.0 ? s.mass_scaling * stable : 0.0;
    p.dt0 = s.time_step > 0.0 ? s.time_step : stable * s.dt_scale;
    p.adaptive = !(s.time_step > 0.0) && s.adaptive_check_steps > 0;
    if (!(p.dt0 > 0.0) || !std::isfinite(p.dt0)) fail(CRUSH_E_INPUT, "no finite stable time step");
    // groups
    p.n_groups = m.n_groups;
    p.g_nodes_ptr.push_back(0);
    p.g_front_ptr.push_back(0);
    for (int g = 0; g < m.n_groups; ++g) {
        const crush_group& gr = m.groups[g];
        double mt = 0.0, pv[3] = {0, 0, 0};
        for (int i = 0; i < gr.n_nodes; ++i) {
            int n = gr.nodes[i];
            if (n < 0 || n >= N) fail(CRUSH_E_INPUT, "group %d: node out of range", g);
            p.g_nodes.push_back(n);
            mt += p.mass[n];
            for (int d = 0; d < 3; ++d) pv[d] += p.mass[n] * p.v0[3 * n + d];
        }
        for (int i = 0; i < gr.n_front; ++i) {
            int n = gr.front[i];
            if (n < 0 || n >= N) fail(CRUSH_E_INPUT, "group %d: front node out of range", g);
            p.g_front.push_back(n);
        }
        if (gr.accel_node >= N || gr.axis < 0 || gr.axis > 2) fail(CRUSH_E_INPUT, "group %d: bad accelerometer", g);
        if (gr.frame_origin >= N || gr.frame_x >= N || gr.frame_plane >= N ||
            ((gr.frame_origin < 0) != (gr.frame_x < 0)) || (gr.frame_origin < 0 && gr.frame_plane >= 0) ||
            (gr.frame_origin >= 0 && gr.frame_origin == gr.frame_x))
            fail(CRUSH_E_INPUT, "group %d: bad accelerometer frame", g);
        p.g_frame.push_back(gr.frame_origin < 0 ? -1 : gr.frame_origin);
        p.g_frame.push_back(gr.frame_x < 0 ? -1 : gr.frame_x);
        p.g_frame.push_back(gr.frame_plane < 0 ? -1 : gr.frame_plane);
        p.g_heading.push_back(gr.heading > 0 ? 1 : (gr.heading < 0 ? -1 : 0));
        p.g_nodes_ptr.push_back((int32_t)p.g_nodes.size());
        p.g_front_ptr.push_back((int32_t)p.g_front.size());
        p.g_accel.push_back(gr.accel_node);
        p.g_axis.push_back(gr.axis);
        p.g_mass.push_back(mt);
        for (int d = 0; d < 3; ++d) p.g_v0.push_back(mt > 0 ? pv[d] / mt : 0.0);
    }
    // pulse grid: t = k pulse_dt up to the end time (the last sample lies at
    // or before end_time; one slot of slack for its rounding)
    p.n_pulse = 0;
    if (s.pulse_dt > 0.0 && m.n_groups > 0) {
        const double n = std::floor(s.end_time / s.pulse_dt) + 2.0;
        if (!(n <= 1e7)) fail(CRUSH_E_INPUT, "pulse_dt %g gives too many samples for end_time %g", s.pulse_dt, s.end_time);
        p.n_pulse = (int)n;
    } else if (s.pulse_dt < 0.0 || !(s.pulse_dt == s.pulse_dt)) {
        fail(CRUSH_E_INPUT, "settings: pulse_dt must be >= 0");
    }
    // part histories: the nodes of each part's elements, ascending
    // (Mesh::part_nodes, src/mesh.rs:74-82)
    if (s.history_steps < 0 || s.frame_steps < 0) fail(CRUSH_E_INPUT, "settings: history_steps and frame_steps must be >= 0");
    p.n_parts = m.n_parts;
    p.p_nodes_ptr.assign(1, 0);
    p.p_nodes.clear();
    p.n_hist = 0;
    if (s.history_steps > 0) {
        std::vector<std::vector<int32_t>> pn(m.n_parts);
        for (int e = 0; e < E; ++e)
            for (int i = 0; i < 8; ++i) pn[m.hex_part[e]].push_back(m.hexes[8 * e + i]);
        for (auto& v : pn) {
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            p.p_nodes.insert(p.p_nodes.end(), v.begin(), v.end());
            p.p_nodes_ptr.push_back((int32_t)p.p_nodes.size());
        }
        p.n_hist = history_capacity(p, s.history_steps);
    }
}

// Device history samples a solve can take: one per history_steps steps
// plus the last step's, with the step count bounded by the smallest step
// the solve can take (the fixed step, or dt_scale x min(dt_floor,
// contact_dt) under mass scaling); HISTORY_MAX_SAMPLES when unbounded.
int history_capacity(const Prepared& p, int every) {
    const crush_settings& s = p.settings;
    double dt_min = 0.0;
    if (s.time_step > 0.0 || !p.adaptive)
        dt_min = p.dt0;
    else if (p.dt_floor > 0.0)
        dt_min = s.dt_scale * std::min(p.dt_floor, p.contact_dt);
    double n = HISTORY_MAX_SAMPLES;
    if (dt_min > 0.0) n = std::min(n, std::ceil(s.end_time / dt_min) / (double)every + 2.0);
    return (int)std::max(n, 1.0);
}

}  // namespace crush
