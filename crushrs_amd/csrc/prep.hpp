// Host-side preparation of a model for the device: validation, element
// geometry classes, node/element incidence (CSR), contact tables and the
// initial time step.  Shared by the one-solve path, the batched path and
// the server.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/crushrs.h"

namespace crush {

constexpr int HSYM = 78;        // packed upper triangle of the 12x12 H
// Geometry constants of one class, contiguous (class-major [C][CPK] f32,
// 832 bytes, each class block 64-byte aligned): per node i a record in the
// 64-byte slot at CPK_NODE i (gradients dN_i/dX_b at +b, hourglass vectors
// gamma_ia at +3+a, the rest padding), then the packed H at CPK_H and the
// volume at CPK_VOL.  The force kernels read a wave-uniform class with
// scalar loads of whole records; with one record per 64-byte line and the
// records apart, no scalar load spans two lines.
constexpr int CPK_NODE = 16, CPK_GRAD = 0, CPK_GAM = 3, CPK_H = 8 * CPK_NODE, CPK_VOL = CPK_H + HSYM, CPK = 208;
static_assert(CPK_VOL < CPK && CPK % 16 == 0 && CPK_H % 16 == 0, "class block layout");
constexpr int SEARCH_BLOCK = 256;
constexpr int FACE_HITS_CAP = 16;  // recorded hits per face before the scan fallback
// Node flags (Prepared::fixed): a fixed node; a node contact forces reach
// (its f_ext is written every step; every other node's stays 0); on the
// batched path, a node whose gathered acceleration or position the pulses
// read (set in the instance image only).
constexpr uint8_t NODE_FIXED = 1, NODE_CONTACT = 2, NODE_PROBE = 4;

// Per-material parameters as the force kernels read them (f32).
struct MatParams {
    int32_t kind;        // CRUSH_MAT_ELASTIC or CRUSH_MAT_HONEYCOMB
    float youngs;
    float bulk;          // elastic: K = E / (3 (1 - 2 nu))
    float shear;         // elastic: G = E / (2 (1 + nu))
    float trans;         // honeycomb: transverse cap factor
    float knot_c[8];     // honeycomb knot table (unused knots at +inf)
    float knot_s[8];
    float knot_h[8];
};

struct ContactDesc {
    int32_t node_begin, n_nodes;   // into the flattened contact-node arrays
    int32_t face_begin, n_faces;   // into the flattened contact-face arrays
    double stiffness, max_distance, depth_cap;
    int32_t pair;
    // batched path: this contact's face-grid region in the instance's grid
    // array (offset in ints) and its capacities in cells and entries
    int32_t grid_cells;
    int32_t grid_entries, grid_off;
    int32_t all_static;  // every face of the contact is of fixed nodes
};

struct Prepared {
    int n_nodes = 0, n_elems = 0, n_hc = 0, n_classes = 0, n_parts = 0;
    // Element slots: honeycomb elements first, then elastic, each in
    // original order.
    std::vector<int32_t> slot_elem, elem_slot;
    std::vector<int32_t> conn;        // [8][E] by slot
    std::vector<int32_t> ecls;        // [E] geometry class per slot
    std::vector<float> cls_pack;      // [C][CPK] geometry constants
    std::vector<int32_t> cls_mat;     // [C]
    std::vector<double> cls_dtf;      // [C] (1 - 0.36 nu) / c_d
    std::vector<double> elem_mass;    // [E] by slot: V rho
    std::vector<MatParams> mats;      // [parts]
    std::vector<double> X, v0, mass;  // [3N], [3N], [N]
    std::vector<uint8_t> fixed;       // [N] NODE_FIXED | NODE_CONTACT
    // node -> (slot << 3 | corner), incidences in original element order
    std::vector<int32_t> n2e_ptr, n2e_idx;
    // contacts
    std::vector<ContactDesc> cdesc;
    std::vector<int32_t> cn_node;     // [K] kept contacting nodes
    std::vector<int32_t> cn_contact;  // [K]
    std::vector<double> cn_scale;     // [K]
    std::vector<int32_t> cf_nodes;    // [4][F] node ids of each contact face (SoA)
    std::vector<int32_t> cf_contact;  // [F]
    std::vector<uint8_t> cf_static;   // [F] 1: every node of the face is fixed
    std::vector<int32_t> tn_node, tn_ptr, tn_ent;  // contact apply lists
    std::vector<int32_t> tn_of;                    // [N] apply list of each node, -1 if none
    std::vector<int32_t> sblk_contact, sblk_start;  // search blocks
    int n_cnodes = 0, n_cfaces = 0;
    // time step
    double contact_dt = 0, stable = 0, dt_floor = 0, dt0 = 0;
    bool adaptive = true;
    crush_settings settings{};
    // groups (outputs)
    int n_groups = 0;
    std::vector<int32_t> g_nodes_ptr, g_nodes, g_front_ptr, g_front, g_accel, g_axis;
    // accelerometer frames [3*G] (origin, x-axis node, plane node; -1 =
    // none) and directions of travel [G]
    std::vector<int32_t> g_frame, g_heading;
    // accelerometer pulse samples per group (0 = no pulse): t = k pulse_dt
    // for k < n_pulse covers [0, end_time]
    int n_pulse = 0;
    // part histories: nodes of each part (CSR, ascending) and the device
    // samples one solve keeps (0 = no history)
    std::vector<int32_t> p_nodes_ptr, p_nodes;
    int n_hist = 0;
    // total mass and lumped masses of each group (for outputs)
    std::vector<double> g_mass, g_v0;  // g_v0: [3*G] initial mean velocity
};

// Validate and prepare; throws crush::Error.
void prepare(const crush_model& m, Prepared& p);

// Upper bound of the history samples recorded on the device after the
// initial one (see prepare).
constexpr int HISTORY_MAX_SAMPLES = 1 << 16;
int history_capacity(const Prepared& p, int every);

// Element geometry of one hexahedron (product implementation): centre
// gradient [24], volume, lumped masses [8] (unit density), hourglass base
// [32] ([4*i+a]) and the packed 12x12 stiffness [78] for E = 1.
void hex_geometry(const double* x, double poisson, double hourglass, double* grad, double* vol,
                  double* lumped, double* gamma, double* hsym);

// Characteristic length V / A_max and volume of a hexahedron (f64).
void hex_char_length(const double* p, double* length, double* volume);

}  // namespace crush
