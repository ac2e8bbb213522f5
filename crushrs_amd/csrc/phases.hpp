// Per-item device work of one explicit step, shared by the one-solve
// kernels (solver_plain.hip) and the batched persistent kernel (batch.hip):
// each function does the work of one element, node, face or contacting node.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace crush {

// Raw buffer access: one SGPR resource per array and 32-bit offsets, so a
// loop over uniform strides (SoA rows, class constants) shares one VGPR
// offset instead of materialising a 64-bit address per load.  Offsets past
// the resource's size read as zero and drop writes.
using rsrc_t = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ rsrc_t make_rsrc(const void* p, size_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)(unsigned)bytes, 0x00020000);
}
__device__ __forceinline__ float ld_f32(rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ int ld_i32(rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(int, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ double ld_f64(rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ void st_f32(rsrc_t r, float v, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}
typedef unsigned u32x3 __attribute__((ext_vector_type(3)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
// Three consecutive doubles in two loads (16 + 8 bytes).
__device__ __forceinline__ void ld_f64x3(rsrc_t r, unsigned voff, double* x) {
    const u32x4 a = __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0);
    const u32x2 b = __builtin_amdgcn_raw_buffer_load_b64(r, voff + 16u, 0, 0);
    x[0] = __builtin_bit_cast(double, (unsigned long long)a.x | ((unsigned long long)a.y << 32));
    x[1] = __builtin_bit_cast(double, (unsigned long long)a.z | ((unsigned long long)a.w << 32));
    x[2] = __builtin_bit_cast(double, b);
}
struct Float3 {
    float x, y, z;
};
struct Double3 {
    double x, y, z;
};
// Loads and stores through the global address space.  An instance's arrays
// are device memory reached through pointers in its table, which the
// compiler cannot place, so plain accesses become flat ones; a global
// access counts only against the vector-memory counter, so its waits do
// not also wait for scalar loads and LDS.
#define CRUSH_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T gld(const T* p) {
    return *(const CRUSH_GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ void gst(T* p, T v) {
    *(CRUSH_GLOBAL T*)p = v;
}
// Nodal vector n of an interleaved xyz array as one 24-byte access (the
// three loads and stores merge).
__device__ __forceinline__ Double3 ld_d3(const double* a, int n) {
    const double* p = a + 3 * (size_t)n;
    return Double3{gld(p), gld(p + 1), gld(p + 2)};
}
__device__ __forceinline__ void st_d3(double* a, int n, const Double3& v) {
    double* p = a + 3 * (size_t)n;
    gst(p, v.x);
    gst(p + 1, v.y);
    gst(p + 2, v.z);
}
// Three consecutive floats as one 12-byte access.
__device__ __forceinline__ Float3 ld_f3(const float* p) { return Float3{gld(p), gld(p + 1), gld(p + 2)}; }
// Current position X + u of node n.
__device__ __forceinline__ void pos3(const double* X, const double* u, int n, double* x) {
#pragma clang fp contract(off)
    const Double3 a = ld_d3(X, n), b = ld_d3(u, n);
    x[0] = a.x + b.x;
    x[1] = a.y + b.y;
    x[2] = a.z + b.z;
}
// Three consecutive floats in one 12-byte store.
__device__ __forceinline__ void st_f32x3(rsrc_t r, float a, float b, float c, unsigned voff, unsigned soff) {
    u32x3 v;
    v.x = __builtin_bit_cast(unsigned, a);
    v.y = __builtin_bit_cast(unsigned, b);
    v.z = __builtin_bit_cast(unsigned, c);
    __builtin_amdgcn_raw_buffer_store_b96(v, r, voff, soff, 0);
}

// Corner forces are stored by tiles of 64 element slots: each tile holds
// its 8 corners' xyz for its 64 slots (corner-major inside the tile), so a
// force pass wave stores each corner as one contiguous 768-byte run and the
// 8 corner forces a node gathers lie within a few tiles.  fe_floats is the
// array's size, fe_index the first float of (corner, slot).
__host__ __device__ __forceinline__ size_t fe_floats(int E) { return (size_t)24 * (((size_t)E + 63) & ~size_t(63)); }
__host__ __device__ __forceinline__ size_t fe_index(int corner, int slot) {
    return ((size_t)((slot >> 6) * 8 + corner) * 64 + (slot & 63)) * 3;
}

// Everything a step touches for one model (one instance), as device
// pointers.  The one-solve path fills it from its buffers; the batched path
// keeps one per instance in a table.
struct FaceGrid;

struct InstDev {
    int N, E, n_hc, C, F, K, T, n_contacts, n_groups, n_sblk;
    // elements
    const int32_t* conn;
    const int32_t* ecls;
    const int32_t* cls_mat;
    const float* cpk;  // [C][CPK] geometry constants per class
    const double* cdtf;
    const double* emass;
    const MatParams* mats;
    float* state;
    uint8_t* eroded;
    float* fe;
    double* extra;
    // nodes
    const double* X;
    double* u;
    double* v;
    double* f_int;
    double* f_ext;   // contact forces (batched path: of even steps)
    double* f_ext2;  // batched path: contact forces of odd steps (one-solve path: f_ext)
    double* inv_mass;
    const double* mass;
    double* added;
    const uint8_t* fixed;
    const int32_t* n2e_ptr;
    const int32_t* n2e_idx;
    // contacts
    const ContactDesc* cdesc;
    const int32_t* cn_node;
    const double* cn_scale;
    const int32_t* cf_nodes;
    const int32_t* cf_contact;
    const uint8_t* cf_static;
    uint8_t* cf_level;  // batched path: kept-grid margin of the face is margin x 2^level
    const int32_t* tn_node;
    const int32_t* tn_ptr;
    const int32_t* tn_ent;
    const int32_t* tn_of;  // batched path: apply list of each node (-1 if none)
    const int32_t* sblk_contact;
    const int32_t* sblk_start;
    double* fgeo;
    float4* fbox;  // [F][2]: each face's candidate box widened to f32 (lo, hi; w unused)
    int32_t* rec_face;
    double* rec_f;
    int32_t* fh_count;
    int32_t* fh_list;
    double* face_force;
    // groups
    const int32_t* g_nodes_ptr;
    const int32_t* g_nodes;
    const int32_t* g_front_ptr;
    const int32_t* g_front;
    const int32_t* g_accel;
    const int32_t* g_axis;
    const double* g_mass;
    const double* g_v0;   // [3G] initial mean velocities
    double* g_dv;         // [3G] delta-v written when the solve ends
    const int32_t* g_frame;  // [3G] accelerometer frame nodes (origin, x, plane), -1 = none
    // accelerometer pulses [min(G, 8)][n_pulse] at t = k pulse_dt (n_pulse 0: none)
    double* pulse;
    int n_pulse;
    double pulse_dt;
    // histories: nodes of each part (CSR) and n_hist records of
    // hist_record_len doubles, one per sample after the initial state
    const int32_t* p_nodes_ptr;
    const int32_t* p_nodes;
    double* hist;
    int n_parts, n_hist, hist_steps;
    StepCtl* ctl;
    // batched path: per-tile partials of the stable step (min dt_e, sum of
    // scaling masses), one slot per force tile of the instance
    double* dtp_min;
    double* dtp_sum;
    // batched path: each contact's face grid in device memory (cell starts
    // then entries, in the region ContactDesc::grid_off/grid_cells/
    // grid_entries describe) and its parameters
    int32_t* cgrid;
    FaceGrid* cgp;
    // batched path: nodes whose gathered state the pulses read (each
    // accelerometer node and its frame's origin and x node, groups < 8,
    // flagged NODE_PROBE), and per node a record {acceleration, position}
    // written by the gather of every step whose gather runs with the next
    // half step
    const int32_t* probe_node;
    double* probe;
    int n_probe;
    // per group < 8: record slots of the accelerometer node (-1: none, or
    // a fixed node), the frame's origin and its x node (-1: no frame)
    const int32_t* g_probe;
    // batched path: cell size of the instance's contact grids, as a
    // multiple of the mean face extent (a speed knob: the grid never
    // changes which face a node hits)
    double cell_scale;
};

// Contact forces of 1-based step q (two buffers by the parity of q on the
// batched path, where the contact of step q + 1 runs before the gather of
// step q; one buffer on the one-solve path).
__device__ __forceinline__ double* fext_step(const InstDev& D, long long q) { return (q & 1) ? D.f_ext2 : D.f_ext; }

__device__ __forceinline__ double step_dt(const StepCtl& c) { return fmax(fmin(c.dt_cfl, c.end_time - c.t), 1e-15); }

// Stable step and scaling mass of one element (src/kernel.rs:214-258).
__device__ __forceinline__ void dt_element(const InstDev& D, int s, double dt_floor, double* dte_out, double* ex_out) {
#pragma clang fp contract(off)
    double dte = INFINITY, ex = 0.0;
    if (!D.eroded[s]) {
        double p[8][3];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int n = D.conn[(size_t)i * D.E + s];
            pos3(D.X, D.u, n, p[i]);
        }
        double l, vol;
        char_length_d(p, &l, &vol);
        if (isfinite(l) && vol > 0.0) dte = l * D.cdtf[D.ecls[s]];
    }
    if (dte < dt_floor) {
        const double r = dt_floor / dte;
        ex = (fmin(r * r, 100.0) - 1.0) * D.emass[s];
    }
    D.extra[s] = ex;
    *dte_out = dte;
    *ex_out = ex;
}

// Added mass and inverse mass of one node (the inverse mass is returned).
__device__ __forceinline__ double mass_node(const InstDev& D, int n) {
#pragma clang fp contract(off)
    double a = 0.0;
    const int kb = gld(D.n2e_ptr + n), ke = gld(D.n2e_ptr + n + 1);
    for (int k = kb; k < ke; ++k) a += gld(D.extra + (gld(D.n2e_idx + k) >> 3)) / 8.0;
    gst(D.added + n, a);
    const double m = gld(D.mass + n) + a;
    const double im = m > 0.0 ? 1.0 / m : 0.0;
    gst(D.inv_mass + n, im);
    return im;
}

// Face record f: the f64 geometry (FG_BLO doubles, SoA) and the candidate
// box rounded outward to f32 in one 32-byte record, with which the search
// rejects what the f64 box rejects and a little less.
__device__ __forceinline__ void store_face(const InstDev& D, int f, const double* g) {
#pragma unroll
    for (int k = 0; k < FG_BLO; ++k) D.fgeo[(size_t)k * D.F + f] = g[k];
    D.fbox[2 * (size_t)f] = make_float4(f32_down(g[FG_CLO]), f32_down(g[FG_CLO + 1]), f32_down(g[FG_CLO + 2]), 0.0f);
    D.fbox[2 * (size_t)f + 1] = make_float4(f32_up(g[FG_CHI]), f32_up(g[FG_CHI + 1]), f32_up(g[FG_CHI + 2]), 0.0f);
}

// Geometry of contact face f for the current configuration.
__device__ __forceinline__ void contact_face(const InstDev& D, int f) {
#pragma clang fp contract(off)
    // a face of fixed nodes keeps the geometry of the first step
    if (D.cf_static[f] && D.ctl->step > 0) return;
    double x[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int n = D.cf_nodes[(size_t)k * D.F + f];
        pos3(D.X, D.u, n, x[k]);
    }
    double g[FG_BLO];
    face_geometry(x, D.cdesc[D.cf_contact[f]].max_distance, g);
    store_face(D, f, g);
}

// contact_face on the four lanes q = 0..3 of a quad of consecutive lanes,
// each forming one triangle's part of the candidate region; every lane of
// the quad calls it for the same face, and lane 0 writes the record (the
// same record as contact_face's).
__device__ __forceinline__ void contact_face_quad(const InstDev& D, int f, int q) {
#pragma clang fp contract(off)
    if (D.cf_static[f] && D.ctl->step > 0) return;
    double x[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int n = D.cf_nodes[(size_t)k * D.F + f];
        pos3(D.X, D.u, n, x[k]);
    }
    const double md = D.cdesc[D.cf_contact[f]].max_distance;
    double g[FG_BLO];
    double tlo[3] = {INFINITY, INFINITY, INFINITY}, thi[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool valid = face_geometry_base(x, md, g);
    if (valid) valid = face_triangle_region(x, g, md, q, tlo, thi);
#pragma unroll
    for (int off = 1; off < 4; off <<= 1) {
        valid = (__shfl_xor((int)valid, off, 64) != 0) && valid;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            tlo[d] = fmin(tlo[d], __shfl_xor(tlo[d], off, 64));
            thi[d] = fmax(thi[d], __shfl_xor(thi[d], off, 64));
        }
    }
    if (q != 0) return;
    face_geometry_finish(valid, tlo, thi, g);
    store_face(D, f, g);
}

// Grid over one contact's face candidate boxes (built in LDS).
struct FaceGrid {
    double glo[3];
    double h;
    int dim[3];
    int overflow;
    int stale;      // batched path, large contacts: a face left its kept box this step
    int built;      // the grid arrays hold a grid of these parameters
    double margin;  // face boxes were binned grown by this much (batched path)
};

__device__ __forceinline__ int grid_cell(const FaceGrid& g, double x, int d) {
    int ci = (int)floor((x - g.glo[d]) / g.h);
    return min(max(ci, 0), g.dim[d] - 1);
}

// Penalty force of contacting node gk (src/contact.rs:113-157): the
// lowest-index accepting face among the grid candidates (or, on grid
// overflow, the first accepting face in order).  Records the hit and
// returns 1 when the node penetrates.  With L > 1, L consecutive lanes of a
// wave share the node: lane `sub` tests every L-th group of candidates, the
// lanes' lowest accepting faces are reduced to the lowest, and lane 0 of the
// group records the hit.
template <int L = 1>
__device__ __forceinline__ int contact_node(const InstDev& D, const ContactDesc& cd, int k, const FaceGrid& g,
                                            const int* s_start, const int* s_entries, double dt, int sub = 0) {
#pragma clang fp contract(off)
    const int gk = cd.node_begin + k;
    const int node = D.cn_node[gk];
    const int F = D.F, fb = cd.face_begin;
    double p[3];
    pos3(D.X, D.u, node, p);
    int best = -1;
    double best_sd = 0.0;
    if (g.overflow) {
        for (int f = sub; f < cd.n_faces; f += L) {
            double sd;
            if (face_accepts(D.fgeo, F, fb + f, p, cd.max_distance, &sd)) {
                best = f;
                best_sd = sd;
                break;
            }
        }
    } else {
        bool inside = true;
        int ci[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double r = (p[d] - g.glo[d]) / g.h;
            inside = inside && r >= 0.0 && r < (double)g.dim[d] + 1.0;
            ci[d] = min(max((int)floor(r), 0), g.dim[d] - 1);
        }
        if (inside) {
            // candidates four at a time: their ids and candidate boxes are
            // loaded together, and a node outside a face's candidate box
            // (a superset of what the face can accept; here its f32 record,
            // rounded outward, so a superset of the f64 box) skips the test
            const int cell = (ci[2] * g.dim[1] + ci[1]) * g.dim[0] + ci[0];
            const int qe = s_start[cell + 1];
            for (int q0 = s_start[cell] + 4 * sub; q0 < qe; q0 += 4 * L) {
                int fs[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) fs[j] = q0 + j < qe ? s_entries[q0 + j] : -1;
                bool in_box[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int gf = fb + (fs[j] >= 0 ? fs[j] : 0);
                    const float4 lo = D.fbox[2 * (size_t)gf], hi = D.fbox[2 * (size_t)gf + 1];
                    in_box[j] = fs[j] >= 0 && p[0] >= (double)lo.x && p[1] >= (double)lo.y && p[2] >= (double)lo.z &&
                                p[0] <= (double)hi.x && p[1] <= (double)hi.y && p[2] <= (double)hi.z;
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int f = fs[j];
                    if (!in_box[j] || (best >= 0 && f >= best)) continue;
                    double sd;
                    if (face_accepts(D.fgeo, F, fb + f, p, cd.max_distance, &sd)) {
                        best = f;
                        best_sd = sd;
                    }
                }
            }
        }
    }
    if (L > 1) {
        // lowest accepting face of the group, with its signed distance
#pragma unroll
        for (int off = 1; off < L; off <<= 1) {
            const int ob = __shfl_xor(best, off, 64);
            const double osd = __shfl_xor(best_sd, off, 64);
            if (ob >= 0 && (best < 0 || ob < best)) {
                best = ob;
                best_sd = osd;
            }
        }
        if (sub != 0) return 0;
    }
    if (best < 0) {
        D.rec_face[gk] = -1;
        return 0;
    }
    const int gf = fb + best;
    const double m = D.mass[node];
    const double depth = fmin(-best_sd, cd.depth_cap);
    const double kk = fmin(cd.stiffness * D.cn_scale[gk], C_SOFT * m / (dt * dt));
    const double mag = fmin(kk * depth + 0.0, m * C_MAX_ACCEL);
#pragma unroll
    for (int d = 0; d < 3; ++d) D.rec_f[3 * (size_t)gk + d] = mag * D.fgeo[(size_t)(FG_N + d) * F + gf];
    D.rec_face[gk] = gf;
    const int slot = atomicAdd(&D.fh_count[gf], 1);
    if (slot < FACE_HITS_CAP) D.fh_list[(size_t)gf * FACE_HITS_CAP + slot] = gk;
    return 1;
}

// Sum of the quarter reactions face f receives, in contacting-node order.
__device__ __forceinline__ void face_reaction(const InstDev& D, int f, double* s) {
#pragma clang fp contract(off)
    const int cnt = D.fh_count[f];
    s[0] = s[1] = s[2] = 0.0;
    if (cnt > 0 && cnt <= FACE_HITS_CAP) {
        int ids[FACE_HITS_CAP];
        for (int i = 0; i < cnt; ++i) ids[i] = D.fh_list[(size_t)f * FACE_HITS_CAP + i];
        for (int i = 1; i < cnt; ++i) {
            int x = ids[i], j = i - 1;
            while (j >= 0 && ids[j] > x) {
                ids[j + 1] = ids[j];
                --j;
            }
            ids[j + 1] = x;
        }
        for (int i = 0; i < cnt; ++i)
#pragma unroll
            for (int d = 0; d < 3; ++d) s[d] += 0.25 * D.rec_f[3 * (size_t)ids[i] + d];
    } else if (cnt > FACE_HITS_CAP) {
        const ContactDesc cd = D.cdesc[D.cf_contact[f]];
        for (int k = cd.node_begin; k < cd.node_begin + cd.n_nodes; ++k)
            if (D.rec_face[k] == f)
#pragma unroll
                for (int d = 0; d < 3; ++d) s[d] += 0.25 * D.rec_f[3 * (size_t)k + d];
    }
}

// Reaction of face f, and its hit count cleared for the next step's search
// (the counts start at zero).
__device__ __forceinline__ void contact_face_force(const InstDev& D, int f) {
    double s[3];
    face_reaction(D, f, s);
#pragma unroll
    for (int d = 0; d < 3; ++d) D.face_force[3 * (size_t)f + d] = s[d];
    D.fh_count[f] = 0;
}

// contact_face_force for the faces of one wave's lanes (f, where valid),
// except that a face whose hits overflowed its list is summed by the whole
// wave: 64 contacting nodes at a time, each hit added in ascending node
// order as face_reaction's scan adds it, so the sums are bitwise the same.
// Every lane of the wave calls it.
__device__ __forceinline__ void contact_face_force_wave(const InstDev& D, int f, bool valid) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const bool over = valid && D.fh_count[f] > FACE_HITS_CAP;
    if (valid && !over) contact_face_force(D, f);
    unsigned long long m = __ballot(over);
    while (m) {
        const int src = __builtin_ctzll(m);
        m &= m - 1;
        const int ff = __shfl(f, src, 64);
        const ContactDesc cd = D.cdesc[D.cf_contact[ff]];
        const int end = cd.node_begin + cd.n_nodes;
        double s[3] = {0.0, 0.0, 0.0};
        for (int base = cd.node_begin; base < end; base += 64) {
            const int k = base + lane;
            const bool hit = k < end && D.rec_face[k] == ff;
            double r[3] = {0.0, 0.0, 0.0};
            if (hit)
#pragma unroll
                for (int d = 0; d < 3; ++d) r[d] = D.rec_f[3 * (size_t)k + d];
            unsigned long long h = __ballot(hit);
            while (h) {
                const int j = __builtin_ctzll(h);
                h &= h - 1;
#pragma unroll
                for (int d = 0; d < 3; ++d) s[d] += 0.25 * __shfl(r[d], j, 64);
            }
        }
        if (lane == src) {
#pragma unroll
            for (int d = 0; d < 3; ++d) D.face_force[3 * (size_t)ff + d] = s[d];
            D.fh_count[ff] = 0;
        }
    }
}

// External (contact) force of touched node i, in contact order: its own
// penalty records, less the reactions of the faces that contain it.
__device__ __forceinline__ Double3 contact_apply_sum(const InstDev& D, int i) {
#pragma clang fp contract(off)
    double s[3] = {0.0, 0.0, 0.0};
    const int kb = gld(D.tn_ptr + i), ke = gld(D.tn_ptr + i + 1);
    for (int k = kb; k < ke; ++k) {
        const int e = gld(D.tn_ent + k);
        if (e >= 0) {
            if (gld(D.rec_face + e) >= 0) {
                const Double3 r = ld_d3(D.rec_f, e);
                s[0] += r.x;
                s[1] += r.y;
                s[2] += r.z;
            }
        } else {
            const Double3 r = ld_d3(D.face_force, -e - 1);
            s[0] -= r.x;
            s[1] -= r.y;
            s[2] -= r.z;
        }
    }
    return Double3{s[0], s[1], s[2]};
}

// contact_apply_sum of touched node i into fext.
__device__ __forceinline__ void contact_apply_node(const InstDev& D, int i, double* fext) {
    st_d3(fext, D.tn_node[i], contact_apply_sum(D, i));
}

// First half step of a free node (src/solver.rs:313-323): v_1/2 = v + dt/2
// (f_ext - f_int) / m, u += dt v_1/2.  Returns false when the new
// displacement is not finite.
__device__ __forceinline__ bool half_update(const double* fi, const double* fe, double im, double dt, const double* v,
                                            const double* u, double* vh, double* un) {
#pragma clang fp contract(off)
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double a = (fe[d] - fi[d]) * im;
        vh[d] = v[d] + 0.5 * dt * a;
        un[d] = u[d] + dt * vh[d];
        ok = ok && isfinite(un[d]);
    }
    return ok;
}

// First half step of node n with contact forces fext; v holds v_1/2 after
// (fixed nodes: u = v = 0).  Returns false when the new displacement is not
// finite.
__device__ __forceinline__ bool half_node(const InstDev& D, int n, double dt, const double* fext) {
    const uint8_t fl = D.fixed[n];
    if (fl & NODE_FIXED) {
        st_d3(D.u, n, Double3{0.0, 0.0, 0.0});
        st_d3(D.v, n, Double3{0.0, 0.0, 0.0});
        return true;
    }
    const double im = D.inv_mass[n];
    // each nodal vector moves as one 24-byte access (16 + 8 bytes)
    const Double3 fi = ld_d3(D.f_int, n), v = ld_d3(D.v, n), u = ld_d3(D.u, n);
    const Double3 fe = (fl & NODE_CONTACT) ? ld_d3(fext, n) : Double3{0.0, 0.0, 0.0};
    const double fid[3] = {fi.x, fi.y, fi.z}, fed[3] = {fe.x, fe.y, fe.z};
    const double vd[3] = {v.x, v.y, v.z}, ud[3] = {u.x, u.y, u.z};
    double vh[3], un[3];
    const bool ok = half_update(fid, fed, im, dt, vd, ud, vh, un);
    st_d3(D.v, n, Double3{vh[0], vh[1], vh[2]});
    st_d3(D.u, n, Double3{un[0], un[1], un[2]});
    return ok;
}

// Pointers into the constant address space: loads through them at a
// wave-uniform address become scalar loads into SGPRs (the data is never
// written while a kernel runs).
typedef float f8v __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef const __attribute__((address_space(4))) f8v* cf8_ptr;
typedef const __attribute__((address_space(4))) f16v* cf16_ptr;
typedef const __attribute__((address_space(4))) int32_t* cint_ptr;
typedef const __attribute__((address_space(4))) MatParams* cmat_ptr;

// Constants of one geometry class (the class block of CPK floats).  With
// UNI the class is wave-uniform and each read is a scalar load of a node
// record (first 8 dwords of its 64-byte slot) or of a 64-byte line of the
// H part; otherwise each lane reads its own class's floats.
template <bool UNI>
struct ClassConsts {
    const float* k;
    // gradient g[3] and hourglass vectors ga[4] of node i
    __device__ __forceinline__ void node(int i, float* g, float* ga) const {
        if (UNI) {
            const f8v v = *(cf8_ptr)(k + CPK_NODE * i);
            g[0] = v[0], g[1] = v[1], g[2] = v[2];
            ga[0] = v[3], ga[1] = v[4], ga[2] = v[5], ga[3] = v[6];
        } else {
            const float* p = k + CPK_NODE * i;
#pragma unroll
            for (int b = 0; b < 3; ++b) g[b] = p[CPK_GRAD + b];
#pragma unroll
            for (int a = 0; a < 4; ++a) ga[a] = p[CPK_GAM + a];
        }
    }
    // 16 floats of the H part and volume from CPK_H + 16 j
    __device__ __forceinline__ f16v h16(int j) const {
        if (UNI) return ((cf16_ptr)(k + CPK_H))[j];
        f16v v;
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = k[CPK_H + 16 * j + e];
        return v;
    }
};

// Element force of slot s (f32) of geometry class c.  With UNI, c is the
// same for every lane that calls it, and the class's constants and material
// come in through scalar loads; without, each lane loads its own class's.
// Streamed so that few values are live at once:
// (1) one pass over the nodes accumulates the displacement gradient
// H = sum_i (u_i - u_0) (x) grad_i and the hourglass coordinates
// q = sum_i (u_i - u_0) gamma_i; (2) the material update gives tau;
// (3) r = H_hg q; (4) each node's force tau F^-T grad_i V + r gamma_i^T is
// written as it is formed.  Returns 1 when the element inverts now.
template <int KIND, bool UNI, int LS>
__device__ __forceinline__ int force_element_class(const InstDev& D, int s, int c, float* lq) {
#pragma clang fp contract(off)
    const unsigned E = (unsigned)D.E, nh = (unsigned)D.n_hc;
    // corner forces: fe[fe_index(corner, slot) + axis], each corner's xyz
    // in one 12-byte store (and one load in the gather)
    const rsrc_t rfe = make_rsrc(D.fe, fe_floats(D.E) * 4);
    const unsigned vs = (unsigned)s * 4u, vs3 = (unsigned)fe_index(0, s) * 4u;
    if (D.eroded[s]) {
#pragma unroll
        for (int i = 0; i < 8; ++i) st_f32x3(rfe, 0.0f, 0.0f, 0.0f, vs3, i * 768u);
        return 0;
    }
    const rsrc_t rconn = make_rsrc(D.conn, (size_t)8 * E * 4);
    const rsrc_t ru = make_rsrc(D.u, (size_t)D.N * 24);
    const ClassConsts<UNI> K{D.cpk + (size_t)c * CPK};
    // the force assembly rereads the node records with per-lane loads, and
    // the elastic kernel (whose Hencky law holds more SGPRs) reads H so
    // too: holding those in SGPRs as well would exceed the SGPR file
    const ClassConsts<false> Kv{D.cpk + (size_t)c * CPK};
    const ClassConsts<UNI && KIND == CRUSH_MAT_HONEYCOMB> Kh{D.cpk + (size_t)c * CPK};
    // the honeycomb state is loaded first so that its latency overlaps the
    // gradient pass
    const rsrc_t rst = make_rsrc(D.state, (size_t)25 * nh * 4);
    float st[KIND == CRUSH_MAT_HONEYCOMB ? 25 : 1];
    if (KIND == CRUSH_MAT_HONEYCOMB) {
#pragma unroll
        for (int k = 0; k < 25; ++k) st[k] = ld_f32(rst, vs, k * nh * 4u);
    }
    const unsigned vn0 = (unsigned)ld_i32(rconn, vs, 0) * 24u;
    double u0[3];
    ld_f64x3(ru, vn0, u0);
    const double u0x = u0[0], u0y = u0[1], u0z = u0[2];
    float hd[9], q[12];
#pragma unroll
    for (int k = 0; k < 9; ++k) hd[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 12; ++k) q[k] = 0.0f;
#pragma unroll
    for (int i = 1; i < 8; ++i) {
        const unsigned vn = (unsigned)ld_i32(rconn, vs, i * E * 4u) * 24u;
        double ui[3];
        ld_f64x3(ru, vn, ui);
        const float d0 = (float)(ui[0] - u0x);
        const float d1 = (float)(ui[1] - u0y);
        const float d2 = (float)(ui[2] - u0z);
        float gi[3], gai[4];
        K.node(i, gi, gai);
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float g = gi[b];
            hd[b] = fmf(d0, g, hd[b]);
            hd[3 + b] = fmf(d1, g, hd[3 + b]);
            hd[6 + b] = fmf(d2, g, hd[6 + b]);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const float ga = gai[a];
            q[3 * a] = fmf(d0, ga, q[3 * a]);
            q[3 * a + 1] = fmf(d1, ga, q[3 * a + 1]);
            q[3 * a + 2] = fmf(d2, ga, q[3 * a + 2]);
        }
    }
    float F[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = hd[k] + ((k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f);
    float J = det3f(F);
    if (!(J > 0.0f)) {
        D.eroded[s] = 1;
#pragma unroll
        for (int i = 0; i < 8; ++i) st_f32x3(rfe, 0.0f, 0.0f, 0.0f, vs3, i * 768u);
        return 1;
    }
    // with LS > 0 the hourglass coordinates wait in LDS (this thread's
    // slots, stride LS) while the material update holds its temporaries, and
    // F is formed again from the unchanged gradient afterwards: fewer values
    // live through the update, the same arithmetic (the elastic kernels,
    // whose Hencky law would otherwise spill; the honeycomb kernel fits
    // 128 VGPRs without it and runs faster so)
    if (LS > 0) {
#pragma unroll
        for (int k = 0; k < 12; ++k) lq[k * LS] = q[k];
    }
    const MatParams& mp = UNI ? *(const MatParams*)((cmat_ptr)D.mats + ((cint_ptr)D.cls_mat)[c])
                              : D.mats[D.cls_mat[c]];
    float tau[9];
    if (KIND == CRUSH_MAT_HONEYCOMB) {
        honeycomb_update(mp, hd, st, tau);
#pragma unroll
        for (int k = 0; k < 25; ++k) st_f32(rst, st[k], vs, k * nh * 4u);
    } else {
        hencky_elastic(mp, hd, tau);
    }
    if (LS > 0) {
#pragma unroll
        for (int k = 0; k < 9; ++k) F[k] = hd[k] + ((k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f);
        J = det3f(F);
#pragma unroll
        for (int k = 0; k < 12; ++k) q[k] = lq[k * LS];
    }
    // M = tau F^-T V
    float finv[9], M[9];
    inv3f(F, J, finv);
    const float vol = Kh.h16((CPK_VOL - CPK_H) / 16)[(CPK_VOL - CPK_H) % 16];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            M[3 * i + j] = vol * dot3f(tau[3 * i], finv[3 * j], tau[3 * i + 1], finv[3 * j + 1], tau[3 * i + 2], finv[3 * j + 2]);
    // r = H q with H packed as its upper triangle
    float rr[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) rr[k] = 0.0f;
    {
        int kk = 0;
#pragma unroll
        for (int r = 0; r < 12; ++r)
#pragma unroll
            for (int cc = r; cc < 12; ++cc) {
                const float h = Kh.h16(kk / 16)[kk % 16];
                ++kk;
                rr[r] = fmf(h, q[cc], rr[r]);
                if (cc != r) rr[cc] = fmf(h, q[r], rr[cc]);
            }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        float gi[3], ga[4];
        Kv.node(i, gi, ga);
        const float g0 = gi[0], g1 = gi[1], g2 = gi[2];
        float f[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            f[d] = dot3f(M[3 * d], g0, M[3 * d + 1], g1, M[3 * d + 2], g2);
#pragma unroll
            for (int a = 0; a < 4; ++a) f[d] = fmf(rr[3 * a + d], ga[a], f[d]);
        }
        st_f32x3(rfe, f[0], f[1], f[2], vs3, i * 768u);
    }
    return 0;
}

// Element force of slot s.  With SCALAR, when the calling lanes' slots
// share one geometry class (as runs of slots of one generated block do),
// the class is taken as wave-uniform and its constants come in through
// scalar loads.  A kernel that inlines this must need no SGPR spills (the
// build checks): a build whose force kernel spilled SGPRs to VGPR lanes
// gave results that changed from run to run.  With LS > 0, lq is the
// calling thread's first float of a 12 x LS workgroup array in LDS (slots
// lq[k * LS]); with LS = 0 it is unused.
template <int KIND, bool SCALAR, int LS>
__device__ __forceinline__ int force_element(const InstDev& D, int s, float* lq) {
    const int c = D.ecls[s];
    if (SCALAR) {
        const int c0 = __builtin_amdgcn_readfirstlane(c);
        if (__ballot(c != c0) == 0) return force_element_class<KIND, true, LS>(D, s, c0, lq);
    }
    return force_element_class<KIND, false, LS>(D, s, c, lq);
}

// Node-centric gather of the element forces into f (f64 sums in element
// order; src/solver.rs:327-332).
__device__ __forceinline__ void gather_sum(const InstDev& D, int n, double* f) {
#pragma clang fp contract(off)
    f[0] = f[1] = f[2] = 0.0;
    const int kb = gld(D.n2e_ptr + n), ke = gld(D.n2e_ptr + n + 1);
    // up to 8 incidences (every node of a hex mesh without merged nodes):
    // all codes, then all element forces, are loaded before the sums so
    // that the loads overlap, without branches: a missing incidence loads
    // the node's last code (or the array's first) and adds +0.0, which
    // leaves the sum unchanged (the sum starts at +0.0 and so is never
    // -0.0); the sums keep element order
    float g[8][3];
    int code[8];
    const int last = ke > 0 ? ke - 1 : 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int c = gld(D.n2e_idx + min(kb + j, last));
        code[j] = kb + j < ke ? c : -1;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        // one 12-byte load per incidence (element 0 for a missing one)
        const size_t base = code[j] >= 0 ? fe_index(code[j] & 7, code[j] >> 3) : 0;
        const Float3 w = ld_f3(D.fe + base);
        g[j][0] = code[j] >= 0 ? w.x : 0.0f;
        g[j][1] = code[j] >= 0 ? w.y : 0.0f;
        g[j][2] = code[j] >= 0 ? w.z : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
        for (int d = 0; d < 3; ++d) f[d] += (double)g[j][d];
    for (int k = kb + 8; k < ke; ++k) {
        const int cd = gld(D.n2e_idx + k);
        const Float3 w = ld_f3(D.fe + fe_index(cd & 7, cd >> 3));
        f[0] += (double)w.x;
        f[1] += (double)w.y;
        f[2] += (double)w.z;
    }
}

// Second half step (src/solver.rs:333-342): a = (f_ext - f_int) / m, v =
// (v_1/2 + dt/2 a) damp, 0 at a fixed node.  Returns false when v is not
// finite.
__device__ __forceinline__ bool gather_update(const double* f, const double* fe, double im, double dt, double damp,
                                              bool fixed, const double* v, double* vn, double* a) {
#pragma clang fp contract(off)
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        a[d] = (fe[d] - f[d]) * im;
        vn[d] = (v[d] + 0.5 * dt * a[d]) * damp;
        if (fixed) vn[d] = 0.0;
        ok = ok && isfinite(vn[d]);
    }
    return ok;
}

// Gather and second half step of node n with contact forces fext (f_int
// and v written).  Returns false when the node's state is not finite.
template <bool CHECK_U = true>
__device__ __forceinline__ bool gather_node(const InstDev& D, int n, double dt, double damp, const double* fext) {
    double f[3];
    gather_sum(D, n, f);
    const double im = D.inv_mass[n];
    const uint8_t fl = D.fixed[n];
    const Double3 v = ld_d3(D.v, n);
    const Double3 fe = (fl & NODE_CONTACT) ? ld_d3(fext, n) : Double3{0.0, 0.0, 0.0};
    const double vd[3] = {v.x, v.y, v.z}, fed[3] = {fe.x, fe.y, fe.z};
    double vn[3], a[3];
    bool ok = gather_update(f, fed, im, dt, damp, (fl & NODE_FIXED) != 0, vd, vn, a);
    st_d3(D.f_int, n, Double3{f[0], f[1], f[2]});
    st_d3(D.v, n, Double3{vn[0], vn[1], vn[2]});
    if (CHECK_U) {
        const Double3 u = ld_d3(D.u, n);
        ok = ok && isfinite(u.x) && isfinite(u.y) && isfinite(u.z);
    }
    return ok;
}

}  // namespace crush

namespace crush {

// ---------------------------------------------------------- block helpers

template <int B>
__device__ __forceinline__ double wave_reduce(double v, int op /* 0 sum, 1 min, 2 max */) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = op == 0 ? v + o : (op == 1 ? fmin(v, o) : fmax(v, o));
    }
    return v;
}

// Block-wide reduction with a fixed order (wave trees, then waves in
// order); every thread gets the result.  s_w holds B/64 doubles.
template <int B>
__device__ __forceinline__ double block_reduce(double v, int op, double* s_w) {
    v = wave_reduce<B>(v, op);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) s_w[wave] = v;
    __syncthreads();
    double r = s_w[0];
    for (int w = 1; w < B / 64; ++w) r = op == 0 ? r + s_w[w] : (op == 1 ? fmin(r, s_w[w]) : fmax(r, s_w[w]));
    __syncthreads();
    return r;
}

// block_reduce of the first n <= K values of v together (value k with
// op[k]), each in block_reduce's order, with one pair of barriers for all.
// s_w holds K * B/64 doubles.
template <int B, int K>
__device__ __forceinline__ void block_reduce_n(double* v, const int* op, int n, double* s_w) {
    constexpr int W = B / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (k < n) v[k] = wave_reduce<B>(v[k], op[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (k < n) s_w[k * W + wave] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (k < n) {
            double r = s_w[k * W];
            for (int w = 1; w < W; ++w)
                r = op[k] == 0 ? r + s_w[k * W + w] : (op[k] == 1 ? fmin(r, s_w[k * W + w]) : fmax(r, s_w[k * W + w]));
            v[k] = r;
        }
    __syncthreads();
}

// s plus u[3 g_front[i] + ax] for i = i0, i0 + B, ... < e, added in that
// order, with four terms' loads in flight together.
template <int B>
__device__ __forceinline__ double front_sum(const InstDev& D, int i0, int e, int ax, double s) {
#pragma clang fp contract(off)
    for (int i = i0; i < e; i += 4 * B) {
        int idx[4];
        double val[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) idx[j] = i + j * B < e ? D.g_front[i + j * B] : -1;
#pragma unroll
        for (int j = 0; j < 4; ++j) val[j] = idx[j] >= 0 ? D.u[3 * (size_t)idx[j] + ax] : 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (idx[j] >= 0) s += val[j];
    }
    return s;
}

// Block-wide exclusive scan of s[0..n) in place (B threads): runs of
// consecutive entries per thread, shuffled scan across each wave, waves
// in order through s_tmp (B/64 ints).
template <int B>
__device__ void block_scan(int* s, int n, int* s_tmp) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int per = (n + B - 1) / B;
    const int b = min(t * per, n), e = min(b + per, n);
    int sum = 0;
    for (int i = b; i < e; ++i) sum += s[i];
    int incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) s_tmp[wave] = incl;
    __syncthreads();
    int acc = incl - sum;
    for (int w = 0; w < wave; ++w) acc += s_tmp[w];
    for (int i = b; i < e; ++i) {
        const int v = s[i];
        s[i] = acc;
        acc += v;
    }
    __syncthreads();
}

// Uniform grid over one contact's face candidate boxes, built by the whole
// block: cell size = cell_scale x the largest face extent (grown until the
// grid has at most MAXC cells); face lists per cell in s_entries (order
// within a cell is irrelevant: the search keeps the lowest index).  The
// arrays are in LDS, or in device memory for grids larger than LDS holds;
// MAXC and MAXE are their capacities.
//
// margin_frac > 0 (the batched path, which keeps a grid over steps): every
// face is binned with its candidate box grown by its margin, margin_frac x
// the largest face extent x 2^level (cf_level), and that box is recorded
// (FG_BLO, FG_BHI); the grid serves later steps as long as every face's box
// stays inside its recorded box grown by its margin (kept_box_holds).  In
// large contacts a face that leaves it has its level raised (up to 3), so a
// face whose candidate box swings as it crushes stops forcing rebuilds while
// the others keep tight bins.
template <int B>
__device__ FaceGrid build_face_grid(const InstDev& D, const ContactDesc& cd, double cell_scale, const int MAXC,
                                    const int MAXE, int* s_start, int* s_cursor, int* s_entries, double* s_w,
                                    int* s_tmp, double margin_frac = 0.0) {
    const int t = threadIdx.x, F = D.F, fb = cd.face_begin, nf = cd.n_faces;
    const bool kept = margin_frac > 0.0;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY}, sz = 0.0, grow = 1.0;
    for (int f = t; f < nf; f += B) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], D.fgeo[(size_t)(FG_CLO + d) * F + fb + f]);
            hi[d] = fmax(hi[d], D.fgeo[(size_t)(FG_CHI + d) * F + fb + f]);
        }
        sz = fmax(sz, D.fgeo[(size_t)FG_SZ * F + fb + f]);
        if (kept) grow = fmax(grow, (double)(1 << D.cf_level[fb + f]));
    }
    FaceGrid g;
    double ghi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        g.glo[d] = block_reduce<B>(lo[d], 1, s_w);
        ghi[d] = block_reduce<B>(hi[d], 2, s_w);
    }
    sz = block_reduce<B>(sz, 2, s_w);
    grow = block_reduce<B>(grow, 2, s_w);
    const double m = margin_frac * sz;
    // per-face margin: m x 2^level on the kept grids, 0 otherwise
    auto face_margin = [&](int f) { return kept ? m * (double)(1 << D.cf_level[fb + f]) : 0.0; };
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        g.glo[d] -= m * grow;
        ghi[d] += m * grow;
    }
    double h = fmax(sz * cell_scale, 1e-9);
    // cell size: grown until the grid has at most MAXC cells; when the
    // faces' candidate boxes then make more than MAXE entries, doubled and
    // counted again (at most 4 times)
    int ncell = 1;
    for (int attempt = 0;; ++attempt) {
        long long cells = 1;
        for (int iter = 0; iter < 64; ++iter) {
            cells = 1;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double ext = ghi[d] - g.glo[d];
                g.dim[d] = ext > 0.0 ? max((int)fmin(ceil(ext / h), 1e6), 1) : 1;
                cells *= g.dim[d];
            }
            if (cells <= MAXC) break;
            h *= 1.25;
        }
        g.h = h;
        g.stale = 0;
        g.built = 1;
        g.margin = m;
        g.overflow = (nf == 0 || cells > MAXC || !(isfinite(g.glo[0]) && isfinite(ghi[0]))) ? 1 : 0;
        const int dx = g.dim[0], dy = g.dim[1];
        ncell = g.overflow ? 1 : (int)cells;
        __syncthreads();
        for (int i = t; i < ncell + 1; i += B) s_start[i] = 0;
        __syncthreads();
        if (!g.overflow) {
            for (int f = t; f < nf; f += B) {
                double flo[3], fhi[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    flo[d] = D.fgeo[(size_t)(FG_CLO + d) * F + fb + f];
                    fhi[d] = D.fgeo[(size_t)(FG_CHI + d) * F + fb + f];
                }
                if (!(flo[0] <= fhi[0] && flo[1] <= fhi[1] && flo[2] <= fhi[2])) continue;
                const double mf = face_margin(f);
                const int a0 = grid_cell(g, flo[0] - mf, 0), a1 = grid_cell(g, fhi[0] + mf, 0);
                const int b0 = grid_cell(g, flo[1] - mf, 1), b1 = grid_cell(g, fhi[1] + mf, 1);
                const int c0 = grid_cell(g, flo[2] - mf, 2), c1 = grid_cell(g, fhi[2] + mf, 2);
                for (int kz = c0; kz <= c1; ++kz)
                    for (int ky = b0; ky <= b1; ++ky)
                        for (int kx = a0; kx <= a1; ++kx) atomicAdd(&s_start[(kz * dy + ky) * dx + kx], 1);
            }
        }
        __syncthreads();
        block_scan<B>(s_start, ncell + 1, s_tmp);
        if (g.overflow || s_start[ncell] <= MAXE || attempt == 4) break;
        h *= 2.0;
    }
    if (s_start[ncell] > MAXE) g.overflow = 1;
    const int dx = g.dim[0], dy = g.dim[1];
    for (int i = t; i < ncell; i += B) s_cursor[i] = s_start[i];
    __syncthreads();
    if (!g.overflow) {
        for (int f = t; f < nf; f += B) {
            double flo[3], fhi[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                flo[d] = D.fgeo[(size_t)(FG_CLO + d) * F + fb + f];
                fhi[d] = D.fgeo[(size_t)(FG_CHI + d) * F + fb + f];
            }
            if (margin_frac > 0.0)
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    D.fgeo[(size_t)(FG_BLO + d) * F + fb + f] = flo[d];
                    D.fgeo[(size_t)(FG_BHI + d) * F + fb + f] = fhi[d];
                }
            if (!(flo[0] <= fhi[0] && flo[1] <= fhi[1] && flo[2] <= fhi[2])) continue;
            const double mf = face_margin(f);
            const int a0 = grid_cell(g, flo[0] - mf, 0), a1 = grid_cell(g, fhi[0] + mf, 0);
            const int b0 = grid_cell(g, flo[1] - mf, 1), b1 = grid_cell(g, fhi[1] + mf, 1);
            const int c0 = grid_cell(g, flo[2] - mf, 2), c1 = grid_cell(g, fhi[2] + mf, 2);
            for (int kz = c0; kz <= c1; ++kz)
                for (int ky = b0; ky <= b1; ++ky)
                    for (int kx = a0; kx <= a1; ++kx) s_entries[atomicAdd(&s_cursor[(kz * dy + ky) * dx + kx], 1)] = f;
        }
    }
    __syncthreads();
    return g;
}

// Contacts of more than LARGE_CONTACT_FACES faces (the batched path): a
// rebuild costs several times the per-step face pass, so their faces get
// margin levels, and the face pass runs in tiles over the whole GPU.
constexpr int LARGE_CONTACT_FACES = 256;

__host__ __device__ __forceinline__ bool large_moving_contact(const ContactDesc& cd) {
    return cd.n_faces > LARGE_CONTACT_FACES && !cd.all_static;
}

// Whether face f (global index) of a contact with a kept grid of margin m
// is still served by it: its candidate box, when valid, lies inside the box
// recorded at the build grown by its margin, m x 2^level (then every cell a
// node in that box maps to lists the face).  In large contacts a face
// outside has its level raised for the rebuild that follows.
__device__ __forceinline__ bool kept_box_holds(const InstDev& D, const ContactDesc& cd, int f, double m) {
    const int F = D.F;
    const int level = D.cf_level[f];
    const double mf = m * (double)(1 << level);
    bool valid = true, inside = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double lo = D.fgeo[(size_t)(FG_CLO + d) * F + f], hi = D.fgeo[(size_t)(FG_CHI + d) * F + f];
        valid = valid && lo <= hi;
        inside = inside && lo >= D.fgeo[(size_t)(FG_BLO + d) * F + f] - mf &&
                 hi <= D.fgeo[(size_t)(FG_BHI + d) * F + f] + mf;
    }
    if (valid && !inside) {
        if (level < 3 && cd.n_faces > LARGE_CONTACT_FACES) D.cf_level[f] = (uint8_t)(level + 1);
        return false;
    }
    return true;
}

// Whether the grid built with margin `m` still serves this step: every
// face that is not static passes kept_box_holds.  Every thread of the
// block calls it.
template <int B>
__device__ bool grid_still_valid(const InstDev& D, const ContactDesc& cd, double m) {
    const int fb = cd.face_begin;
    int stale = 0;
    for (int f = threadIdx.x; f < cd.n_faces; f += B) {
        if (D.cf_static[fb + f]) continue;
        if (!kept_box_holds(D, cd, fb + f, m)) stale = 1;
    }
    return !__syncthreads_or(stale);
}

// Crush of group g, |u_axis(accel) - mean u_axis(front)| (src/bin/crushrs.rs:265-271).
template <int B>
__device__ double group_crush(const InstDev& D, int g, double* s_w) {
    const int b = D.g_front_ptr[g], e = D.g_front_ptr[g + 1], ax = D.g_axis[g];
    double s = front_sum<B>(D, b + (int)threadIdx.x, e, ax, 0.0);
    s = block_reduce<B>(s, 0, s_w);
    if (D.g_accel[g] < 0 || e <= b) return 0.0;
    return fabs(D.u[3 * (size_t)D.g_accel[g] + ax] - s / (double)(e - b));
}

// Mass-weighted mean position and velocity of group g -> out[0..6]
// = {mass, x, y, z, vx, vy, vz}.
template <int B>
__device__ void group_means(const InstDev& D, int g, double* s_w, double* out) {
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = D.g_nodes_ptr[g] + (int)threadIdx.x; i < D.g_nodes_ptr[g + 1]; i += B) {
        const int n = D.g_nodes[i];
        const double m = D.mass[n];
        s[0] += m;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            s[1 + d] += m * (D.X[3 * (size_t)n + d] + D.u[3 * (size_t)n + d]);
            s[4 + d] += m * D.v[3 * (size_t)n + d];
        }
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) out[k] = block_reduce<B>(s[k], 0, s_w);
}

// Local-x component of acceleration a (src/model.rs:45-61): a projected on
// the unit vector from the frame origin's position xo to its x node's
// position xx (a[0] when the frame is degenerate).
__device__ __forceinline__ double local_x(const double* a, const double* xo, const double* xx) {
#pragma clang fp contract(off)
    double e[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) e[d] = xx[d] - xo[d];
    const double en = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    if (en < 1e-300) return a[0];
    return (e[0] / en) * a[0] + (e[1] / en) * a[1] + (e[2] / en) * a[2];
}

// Local-x acceleration of group g's accelerometer node after a step
// (src/solver.rs:150-187 with the acceleration of :333-342): a = (f_ext -
// f_int) / m at the node (0 for fixed nodes), projected on the unit vector
// from the frame's origin node to its x node in the current configuration
// (src/model.rs:45-61).  fext: the step's contact forces.
__device__ __forceinline__ double accel_local_x(const InstDev& D, int g, const double* fext) {
#pragma clang fp contract(off)
    const int n = D.g_accel[g];
    if (n < 0 || (D.fixed[n] & NODE_FIXED)) return 0.0;
    const double im = D.inv_mass[n];
    double a[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) a[d] = (fext[3 * (size_t)n + d] - D.f_int[3 * (size_t)n + d]) * im;
    const int o = D.g_frame[3 * g], x = D.g_frame[3 * g + 1];
    if (o < 0) return a[0];
    double xo[3], xx[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        xo[d] = D.X[3 * (size_t)o + d] + D.u[3 * (size_t)o + d];
        xx[d] = D.X[3 * (size_t)x + d] + D.u[3 * (size_t)x + d];
    }
    return local_x(a, xo, xx);
}

// Record slot of probe node n (batched path; -1 when n is not a probe).
__device__ __forceinline__ int probe_slot(const InstDev& D, int n) {
    for (int k = 0; k < D.n_probe; ++k)
        if (D.probe_node[k] == n) return k;
    return -1;
}

// accel_local_x from the probe records of the last gather (batched path).
__device__ __forceinline__ double accel_local_x_probed(const InstDev& D, int g) {
    const int sa = D.g_probe[3 * g], so = D.g_probe[3 * g + 1], sx = D.g_probe[3 * g + 2];
    if (sa < 0) return 0.0;
    const double* a = D.probe + 6 * (size_t)sa;
    if (so < 0) return a[0];
    return local_x(a, D.probe + 6 * (size_t)so + 3, D.probe + 6 * (size_t)sx + 3);
}

// Streams the sample (tn, a[g]) of every group g < 8 into the pulses: each
// grid time k pulse_dt in (t_prev, tn] gets the linear interpolation
// between the previous sample and this one (the resampling of
// src/signal.rs:43-64 done as the samples arrive).  Thread 0 of the
// instance's block only.
__device__ __forceinline__ void pulse_sample(const InstDev& D, StepCtl* c, double tn, const double* a) {
#pragma clang fp contract(off)
    const int ng = D.n_groups < 8 ? D.n_groups : 8;
    const double tp = c->pulse_t;
    int k = c->pulse_k;
    for (; k < D.n_pulse; ++k) {
        const double tk = (double)k * D.pulse_dt;
        if (!(tk <= tn)) break;
        const double w = (tk - tp) / (tn - tp);
        for (int g = 0; g < ng; ++g) D.pulse[(size_t)g * D.n_pulse + k] = c->pulse_a[g] + w * (a[g] - c->pulse_a[g]);
    }
    c->pulse_k = k;
    c->pulse_t = tn;
    for (int g = 0; g < ng; ++g) c->pulse_a[g] = a[g];
}

// Output state of element slot s for a frame (src/kernel.rs:269-320):
// the Cauchy stress in Voigt order xx yy zz xy yz xz (0 for an eroded or
// inverted element) and the compaction (honeycomb; 0 for elastic).  The
// honeycomb stress is R sigma R^T of its corotational state; the elastic
// stress is the Hencky tau / J at the current displacement (f32, as the
// force pass computes it).
__device__ __forceinline__ void element_output(const InstDev& D, int s, float* sig, float* plastic) {
    const int nh = D.n_hc;
    *plastic = s < nh ? D.state[(size_t)24 * nh + s] : 0.0f;
    for (int k = 0; k < 6; ++k) sig[k] = 0.0f;
    if (D.eroded[s]) return;
    float c[9];
    if (s < nh) {
        float R[9], sh[9];
        for (int k = 0; k < 9; ++k) R[k] = D.state[(size_t)(9 + k) * nh + s];
        const float* st = D.state;
        const float s0 = st[(size_t)18 * nh + s], s1 = st[(size_t)19 * nh + s], s2 = st[(size_t)20 * nh + s];
        const float s3 = st[(size_t)21 * nh + s], s4 = st[(size_t)22 * nh + s], s5 = st[(size_t)23 * nh + s];
        sh[0] = s0; sh[1] = s3; sh[2] = s5;
        sh[3] = s3; sh[4] = s1; sh[5] = s4;
        sh[6] = s5; sh[7] = s4; sh[8] = s2;
        float t[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) t[3 * i + j] = R[3 * i] * sh[j] + R[3 * i + 1] * sh[3 + j] + R[3 * i + 2] * sh[6 + j];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) c[3 * i + j] = t[3 * i] * R[3 * j] + t[3 * i + 1] * R[3 * j + 1] + t[3 * i + 2] * R[3 * j + 2];
    } else {
        const int cl = D.ecls[s];
        const int n0 = D.conn[s];
        float hd[9];
        for (int k = 0; k < 9; ++k) hd[k] = 0.0f;
        for (int i = 1; i < 8; ++i) {
            const int n = D.conn[(size_t)i * D.E + s];
            for (int a = 0; a < 3; ++a) {
                const float du = (float)(D.u[3 * (size_t)n + a] - D.u[3 * (size_t)n0 + a]);
                for (int b = 0; b < 3; ++b) hd[3 * a + b] = fmf(du, D.cpk[(size_t)cl * CPK + CPK_NODE * i + CPK_GRAD + b], hd[3 * a + b]);
            }
        }
        float F[9];
        for (int k = 0; k < 9; ++k) F[k] = hd[k] + ((k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f);
        const float J = det3f(F);
        if (!(J > 0.0f)) return;
        float tau[9];
        hencky_elastic(D.mats[D.cls_mat[cl]], hd, tau);
        for (int k = 0; k < 9; ++k) c[k] = tau[k] / J;
    }
    sig[0] = c[0];
    sig[1] = c[4];
    sig[2] = c[8];
    sig[3] = c[1];
    sig[4] = c[5];
    sig[5] = c[2];
}

// Doubles per history record: time, step, CRUSH_PART_FIELDS per part and
// CRUSH_ACCEL_FIELDS per accelerometer (groups < 8).
__host__ __device__ __forceinline__ size_t hist_record_len(int n_parts, int n_groups) {
    return 2 + (size_t)CRUSH_PART_FIELDS * n_parts + (size_t)CRUSH_ACCEL_FIELDS * (n_groups < 8 ? n_groups : 8);
}

// Current position of node n.
__device__ __forceinline__ void node_position(const InstDev& D, int n, double* x) {
#pragma clang fp contract(off)
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = D.X[3 * (size_t)n + d] + D.u[3 * (size_t)n + d];
}

// Acceleration (f_ext - f_int) / m of node n after a step, 0 at fixed nodes
// (src/solver.rs:333-342); fext: the step's contact forces.
__device__ __forceinline__ void node_accel(const InstDev& D, int n, double* a, const double* fext) {
#pragma clang fp contract(off)
    const bool fx = (D.fixed[n] & NODE_FIXED) != 0;
    const double im = D.inv_mass[n];
#pragma unroll
    for (int d = 0; d < 3; ++d) a[d] = fx ? 0.0 : (fext[3 * (size_t)n + d] - D.f_int[3 * (size_t)n + d]) * im;
}

// Unit axes (rows e1, e2, e3) of a body-fixed frame in the current
// configuration (src/model.rs:30-58): e1 from the origin to the x node, e3
// normal to the plane of the three nodes (e1 x p, with p towards the plane
// node, or towards global X, else Y, when p is parallel to e1), e2 = e3 x e1.
__device__ __forceinline__ void frame_axes(const InstDev& D, int o, int x, int pl, double* ax) {
#pragma clang fp contract(off)
    double xo[3], xx[3], xp[3];
    node_position(D, o, xo);
    node_position(D, x, xx);
    double e1[3], p[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < 3; ++d) e1[d] = xx[d] - xo[d];
    if (pl >= 0) {
        node_position(D, pl, xp);
        for (int d = 0; d < 3; ++d) p[d] = xp[d] - xo[d];
    }
    double n1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    if (n1 < 1e-300) {
        e1[0] = 1.0;
        e1[1] = e1[2] = 0.0;
        n1 = 1.0;
    }
    for (int d = 0; d < 3; ++d) e1[d] /= n1;
    const double pn = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    double c[3] = {p[1] * e1[2] - p[2] * e1[1], p[2] * e1[0] - p[0] * e1[2], p[0] * e1[1] - p[1] * e1[0]};
    if (sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) < 1e-12 * fmax(pn, 1e-300)) {
        const bool xa = fabs(e1[0]) < 0.9;
        p[0] = xa ? 1.0 : 0.0;
        p[1] = xa ? 0.0 : 1.0;
        p[2] = 0.0;
    }
    double e3[3] = {e1[1] * p[2] - e1[2] * p[1], e1[2] * p[0] - e1[0] * p[2], e1[0] * p[1] - e1[1] * p[0]};
    const double n3 = sqrt(e3[0] * e3[0] + e3[1] * e3[1] + e3[2] * e3[2]);
    for (int d = 0; d < 3; ++d) e3[d] /= n3;
    const double e2[3] = {e3[1] * e1[2] - e3[2] * e1[1], e3[2] * e1[0] - e3[0] * e1[2], e3[0] * e1[1] - e3[1] * e1[0]};
    for (int d = 0; d < 3; ++d) {
        ax[d] = e1[d];
        ax[3 + d] = e2[d];
        ax[6 + d] = e3[d];
    }
}

// One history record after a step (src/solver.rs:114-187): each part's
// mass, mass-weighted mean displacement, velocity and acceleration and
// kinetic energy (block sums in a fixed order), and each accelerometer
// node's position, displacement, velocity and acceleration with its
// frame's origin and axes.  Every thread of the block calls it; fext: the
// step's contact forces.
template <int B>
__device__ void history_sample(const InstDev& D, int k, long long step, double tn, double* s_w, const double* fext) {
#pragma clang fp contract(off)
    const int na = D.n_groups < 8 ? D.n_groups : 8;
    double* rec = D.hist + (size_t)k * hist_record_len(D.n_parts, D.n_groups);
    for (int p = 0; p < D.n_parts; ++p) {
        double sm[CRUSH_PART_FIELDS];
#pragma unroll
        for (int f = 0; f < CRUSH_PART_FIELDS; ++f) sm[f] = 0.0;
        for (int i = D.p_nodes_ptr[p] + (int)threadIdx.x; i < D.p_nodes_ptr[p + 1]; i += B) {
            const int n = D.p_nodes[i];
            const double m = D.mass[n];
            double a[3];
            node_accel(D, n, a, fext);
            sm[0] += m;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double vd = D.v[3 * (size_t)n + d];
                sm[1 + d] += m * D.u[3 * (size_t)n + d];
                sm[4 + d] += m * vd;
                sm[7 + d] += m * a[d];
                sm[10] += 0.5 * m * vd * vd;
            }
        }
#pragma unroll
        for (int f = 0; f < CRUSH_PART_FIELDS; ++f) sm[f] = block_reduce<B>(sm[f], 0, s_w);
        if (threadIdx.x == 0) {
            double* r = rec + 2 + (size_t)CRUSH_PART_FIELDS * p;
            const double inv = sm[0] > 0.0 ? 1.0 / sm[0] : 0.0;
            r[0] = sm[0];
            for (int f = 1; f < 10; ++f) r[f] = sm[f] * inv;
            r[10] = sm[10];
        }
    }
    if (threadIdx.x == 0) {
        rec[0] = tn;
        rec[1] = (double)step;
        for (int g = 0; g < na; ++g) {
            double* r = rec + 2 + (size_t)CRUSH_PART_FIELDS * D.n_parts + (size_t)CRUSH_ACCEL_FIELDS * g;
            const int n = D.g_accel[g];
            if (n < 0) {
                for (int f = 0; f < CRUSH_ACCEL_FIELDS; ++f) r[f] = 0.0;
                r[15] = r[19] = r[23] = 1.0;
                continue;
            }
            node_position(D, n, r);
            for (int d = 0; d < 3; ++d) {
                r[3 + d] = D.u[3 * (size_t)n + d];
                r[6 + d] = D.v[3 * (size_t)n + d];
            }
            node_accel(D, n, r + 9, fext);
            const int o = D.g_frame[3 * g], x = D.g_frame[3 * g + 1], pl = D.g_frame[3 * g + 2];
            if (o >= 0) {
                node_position(D, o, r + 12);
                frame_axes(D, o, x, pl, r + 15);
            } else {
                for (int d = 0; d < 3; ++d) r[12 + d] = r[d];
                for (int f = 0; f < 9; ++f) r[15 + f] = (f == 0 || f == 4 || f == 8) ? 1.0 : 0.0;
            }
        }
    }
}

// End of a step: time, step count, crush samples, separation and the stop
// decision (the loop condition of src/solver.rs:287 plus the optional
// retirement at separation).  c is the block's control block in LDS; only
// thread 0 writes it, and every thread returns the same stop decision.
// fext: the step's contact forces.
template <int B>
__device__ bool step_end(const InstDev& D, StepCtl* c, double dt, int pair_hits, bool bad, double* s_w,
                         const double* fext) {
    const long long step = c->step + 1;
    double cr[8];
    const int ng = D.n_groups < 8 ? D.n_groups : 8;
    for (int g = 0; g < ng; ++g) cr[g] = group_crush<B>(D, g, s_w);
    const int had_pair = pair_hits > 0 ? 1 : c->had_pair;
    const int quiet = pair_hits > 0 ? 0 : c->quiet + 1;
    bool sep = false;
    if (c->sep_steps > 0 && D.n_groups >= 2 && had_pair && quiet >= c->sep_steps && step % c->chk == 0) {
        double a[7], b[7];
        group_means<B>(D, 0, s_w, a);
        group_means<B>(D, 1, s_w, b);
        double dot = 0.0;
        for (int d = 0; d < 3; ++d) dot += (b[4 + d] / b[0] - a[4 + d] / a[0]) * (b[1 + d] / b[0] - a[1 + d] / a[0]);
        sep = dot > 0.0;
    }
    const double tn = c->t + dt;
    const bool done = sep || bad || !(tn < c->end_time);
    const int hk = c->hist_k;
    const bool hist = D.n_hist > 0 && (step % D.hist_steps == 0 || done);
    if (hist && hk < D.n_hist) history_sample<B>(D, hk, step, tn, s_w, fext);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (hist) c->hist_k = hk + 1;
        if (D.n_pulse > 0) {
            double a[8];
            for (int g = 0; g < ng; ++g) a[g] = accel_local_x(D, g, fext);
            pulse_sample(D, c, tn, a);
        }
        for (int g = 0; g < ng; ++g) c->crush_max[g] = fmax(c->crush_max[g], cr[g]);
        c->had_pair = had_pair;
        c->quiet = quiet;
        c->t = tn;
        c->dt = dt;
        c->step = step;
        if (sep) c->separated = 1;
        if (bad) c->error = 1;
        if (done) c->done = 1;
    }
    __syncthreads();
    return done;
}

}  // namespace crush
