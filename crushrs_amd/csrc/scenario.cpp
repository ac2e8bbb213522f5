// Scenario builders: restatement of the reference's model construction
// (src/mesh.rs, src/model.rs:167-265, src/vehicle.rs, src/headon.rs:40-66,
// src/bin/crushrs.rs:403-432,522-573) in C++.
#include "scenario.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <fstream>
#include <set>
#include <sstream>
#include <unordered_map>

#include "common.hpp"
#include "toml_lite.hpp"

namespace crush {

// ------------------------------------------------------------------ mesh

int Mesh::add_part(const std::string& name) {
    parts.push_back(name);
    return (int)parts.size() - 1;
}

int Mesh::part_index(const std::string& name) const {
    for (size_t p = 0; p < parts.size(); ++p)
        if (parts[p] == name) return (int)p;
    return -1;
}

std::vector<int> Mesh::part_hexes(int part) const {
    std::vector<int> r;
    for (size_t e = 0; e < hexes.size(); ++e)
        if (hex_part[e] == part) r.push_back((int)e);
    return r;
}

// src/mesh.rs:70-78: sorted unique nodes of the part's hexahedra.
std::vector<int> Mesh::part_nodes(int part) const {
    std::set<int> s;
    for (size_t e = 0; e < hexes.size(); ++e)
        if (hex_part[e] == part) s.insert(hexes[e].begin(), hexes[e].end());
    return std::vector<int>(s.begin(), s.end());
}

// src/mesh.rs:81-86.
bool Mesh::nodes_of(const std::string& name, std::vector<int>& out) const {
    auto it = node_sets.find(name);
    if (it != node_sets.end()) {
        out = it->second;
        return true;
    }
    int p = part_index(name);
    if (p < 0) return false;
    out = part_nodes(p);
    return true;
}

// src/mesh.rs:93-100.
bool Mesh::face_set_nodes(const std::string& name, std::vector<int>& out) const {
    auto it = face_sets.find(name);
    if (it == face_sets.end()) return false;
    std::set<int> s;
    for (int f : it->second) s.insert(faces[f].begin(), faces[f].end());
    out.assign(s.begin(), s.end());
    return true;
}

// src/mesh.rs:107-110.
int Mesh::add_hex_block(const std::string& name, Vec3 origin, Vec3 size, double h,
                        const std::vector<std::pair<BlockFace, std::string>>& surfaces) {
    std::array<int, 3> n;
    for (int a = 0; a < 3; ++a) n[a] = (int)std::max(std::round(size[a] / h), 1.0);
    return add_hex_block_n(name, origin, size, n, surfaces);
}

// src/mesh.rs:113-171: node (i, j, k) of the block has id
// node0 + i + (n0+1)(j + (n1+1)k); hexahedra in C3D8 order; face quads
// counter-clockwise about the outward normal.
int Mesh::add_hex_block_n(const std::string& name, Vec3 origin, Vec3 size, std::array<int, 3> n,
                          const std::vector<std::pair<BlockFace, std::string>>& surfaces) {
    for (int a = 0; a < 3; ++a) n[a] = std::max(n[a], 1);
    Vec3 d;
    for (int a = 0; a < 3; ++a) d[a] = size[a] / (double)n[a];
    const int node0 = (int)nodes.size();
    auto id = [&](int i, int j, int k) { return node0 + i + (n[0] + 1) * (j + (n[1] + 1) * k); };
    for (int k = 0; k <= n[2]; ++k)
        for (int j = 0; j <= n[1]; ++j)
            for (int i = 0; i <= n[0]; ++i)
                nodes.push_back({origin[0] + (double)i * d[0], origin[1] + (double)j * d[1],
                                 origin[2] + (double)k * d[2]});
    int part = add_part(name);
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                hexes.push_back({id(i, j, k), id(i + 1, j, k), id(i + 1, j + 1, k), id(i, j + 1, k),
                                 id(i, j, k + 1), id(i + 1, j, k + 1), id(i + 1, j + 1, k + 1),
                                 id(i, j + 1, k + 1)});
                hex_part.push_back(part);
            }
    for (const auto& sf : surfaces) {
        int a = 0;
        bool max_side = false;
        switch (sf.first) {
            case BlockFace::XMin: a = 0; max_side = false; break;
            case BlockFace::XMax: a = 0; max_side = true; break;
            case BlockFace::YMin: a = 1; max_side = false; break;
            case BlockFace::YMax: a = 1; max_side = true; break;
            case BlockFace::ZMin: a = 2; max_side = false; break;
            case BlockFace::ZMax: a = 2; max_side = true; break;
        }
        int b = (a + 1) % 3, c = (a + 2) % 3;
        int fixed = max_side ? n[a] : 0;
        auto at = [&](int ib, int ic) {
            int ijk[3];
            ijk[a] = fixed;
            ijk[b] = ib;
            ijk[c] = ic;
            return id(ijk[0], ijk[1], ijk[2]);
        };
        std::vector<int> set;
        for (int ic = 0; ic < n[c]; ++ic)
            for (int ib = 0; ib < n[b]; ++ib) {
                int p0 = at(ib, ic), p1 = at(ib + 1, ic), p2 = at(ib + 1, ic + 1), p3 = at(ib, ic + 1);
                if (max_side)
                    faces.push_back({p0, p1, p2, p3});
                else
                    faces.push_back({p0, p3, p2, p1});
                set.push_back((int)faces.size() - 1);
            }
        face_sets[sf.second] = set;
    }
    std::vector<int> all;
    for (int i = node0; i < (int)nodes.size(); ++i) all.push_back(i);
    node_sets[name] = all;
    return part;
}

// src/mesh.rs:176-234: keep the lowest index of coincident nodes.
int Mesh::merge_coincident_nodes(double tol) {
    struct KeyHash {
        size_t operator()(const std::array<long long, 3>& k) const {
            size_t h = 1469598103934665603ull;
            for (long long v : k) h = (h ^ (size_t)v) * 1099511628211ull;
            return h;
        }
    };
    std::unordered_map<std::array<long long, 3>, int, KeyHash> first;
    std::vector<int> map(nodes.size());
    std::vector<Vec3> kept;
    kept.reserve(nodes.size());
    int merged = 0;
    for (size_t i = 0; i < nodes.size(); ++i) {
        const Vec3& p = nodes[i];
        std::array<long long, 3> k;
        for (int d = 0; d < 3; ++d) k[d] = std::llround(p[d] / tol);
        int found = -1;
        for (int dx = -1; dx <= 1 && found < 0; ++dx)
            for (int dy = -1; dy <= 1 && found < 0; ++dy)
                for (int dz = -1; dz <= 1 && found < 0; ++dz) {
                    auto it = first.find({k[0] + dx, k[1] + dy, k[2] + dz});
                    if (it == first.end()) continue;
                    const Vec3& q = kept[it->second];
                    bool close = true;
                    for (int d = 0; d < 3; ++d) close = close && std::fabs(q[d] - p[d]) <= tol;
                    if (close) found = it->second;
                }
        if (found >= 0) {
            map[i] = found;
            ++merged;
        } else {
            map[i] = (int)kept.size();
            first[k] = (int)kept.size();
            kept.push_back(p);
        }
    }
    if (merged == 0) return 0;
    nodes = kept;
    for (auto& h : hexes)
        for (auto& n : h) n = map[n];
    for (auto& f : faces)
        for (auto& n : f) n = map[n];
    for (auto& kv : node_sets) {
        std::vector<int> v;
        for (int n : kv.second) v.push_back(map[n]);
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        kv.second = v;
    }
    return merged;
}

// Mesh::min_edge_at (src/mesh.rs:237-250): shortest hexahedron edge at a
// node.
double Mesh::min_edge_at(int node) const {
    static const int EDGES[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                     {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    double mn = INFINITY;
    for (const auto& h : hexes)
        for (const auto& e : EDGES)
            if (h[e[0]] == node || h[e[1]] == node) {
                const Vec3& p = nodes[h[e[0]]];
                const Vec3& q = nodes[h[e[1]]];
                mn = std::fmin(mn, std::sqrt((p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) +
                                             (p[2] - q[2]) * (p[2] - q[2])));
            }
    return mn;
}

// Mesh::best_aligned_neighbour (src/mesh.rs:262-288): among the part's
// nodes within `radius` of `node` (excluding `exclude`), the first whose
// direction from `node` has the largest cosine with `dir`.
std::pair<int, double> Mesh::best_aligned_neighbour(int node, Vec3 dir, double radius, int part,
                                                    const std::vector<int>& exclude) const {
    std::vector<int> cand;
    if (part >= 0)
        cand = part_nodes(part);
    else
        for (int i = 0; i < (int)nodes.size(); ++i) cand.push_back(i);
    const Vec3& o = nodes[node];
    const double dn = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    int best = -1;
    double bc = 0.0;
    for (int n : cand) {
        if (n == node || std::find(exclude.begin(), exclude.end(), n) != exclude.end()) continue;
        const double r[3] = {nodes[n][0] - o[0], nodes[n][1] - o[1], nodes[n][2] - o[2]};
        const double rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        if (rn > radius || rn == 0.0) continue;
        const double c = (r[0] * dir[0] + r[1] * dir[1] + r[2] * dir[2]) / (rn * dn);
        if (best < 0 || c > bc) {
            best = n;
            bc = c;
        }
    }
    return {best, bc};
}

// src/mesh.rs:549-556: first node of minimum distance among the part's
// sorted nodes.
int Mesh::nearest_node(Vec3 at, int part) const {
    std::vector<int> cand;
    if (part >= 0)
        cand = part_nodes(part);
    else
        for (int i = 0; i < (int)nodes.size(); ++i) cand.push_back(i);
    if (cand.empty()) fail(CRUSH_E_INPUT, "mesh has no nodes");
    int best = cand[0];
    double bd = INFINITY;
    for (int n : cand) {
        double d2 = 0.0;
        for (int d = 0; d < 3; ++d) d2 += (nodes[n][d] - at[d]) * (nodes[n][d] - at[d]);
        if (d2 < bd) {
            bd = d2;
            best = n;
        }
    }
    return best;
}

// ------------------------------------------------------------- materials

Material Material::elastic(double e, double nu, double rho) {
    Material r;
    r.m.youngs_modulus = e;
    r.m.poisson_ratio = nu;
    r.m.density = rho;
    r.m.model = CRUSH_MAT_ELASTIC;
    r.m.transverse_factor = 1.0;
    return r;
}

// Material::honeycomb (src/material.rs:179-181).
Material Material::honeycomb(double e, double rho, double sy, double h) {
    Material r = elastic(e, 0.0, rho);
    r.m.model = CRUSH_MAT_HONEYCOMB;
    r.m.yield_stress = sy;
    r.m.hardening = h;
    return r;
}

// Material::honeycomb_curve (src/material.rs:186-190) with Plasticity::check
// (src/material.rs:116-151).
Material Material::honeycomb_curve(double e, double rho, const std::vector<std::array<double, 2>>& curve,
                                   bool has_dens, std::array<double, 2> dens) {
    if (curve.size() < 2) fail(CRUSH_E_INPUT, "yield curve needs at least 2 knots");
    Material r = honeycomb(e, rho, curve[0][1], (curve[1][1] - curve[0][1]) / (curve[1][0] - curve[0][0]));
    int n = (int)curve.size();
    if (n + (has_dens ? 1 : 0) > CRUSH_MAX_KNOTS)
        fail(CRUSH_E_INPUT, "yield curve needs 2..=%d knots (one fewer with densification)", CRUSH_MAX_KNOTS);
    if (curve[0][0] != 0.0) fail(CRUSH_E_INPUT, "yield curve must start at compaction 0");
    for (int i = 1; i < n; ++i)
        if (curve[i][0] <= curve[i - 1][0]) fail(CRUSH_E_INPUT, "yield curve compaction must be strictly increasing");
    for (int i = 0; i < n; ++i)
        if (curve[i][1] <= 0.0) fail(CRUSH_E_INPUT, "yield curve stresses must be positive");
    if (has_dens) {
        if (dens[0] <= 0.0 || dens[1] < 0.0) fail(CRUSH_E_INPUT, "densification needs c_lock > 0 and k_lock >= 0");
        if (curve.back()[0] >= dens[0]) fail(CRUSH_E_INPUT, "densification c_lock must lie beyond the last curve knot");
    }
    r.m.n_curve = n;
    for (int i = 0; i < n; ++i) {
        r.m.curve[i][0] = curve[i][0];
        r.m.curve[i][1] = curve[i][1];
    }
    r.m.has_densification = has_dens ? 1 : 0;
    r.m.densification[0] = dens[0];
    r.m.densification[1] = dens[1];
    return r;
}

// Material::scaled (src/material.rs:200-215).
Material Material::scaled(double f) const {
    Material r = *this;
    r.m.youngs_modulus *= f;
    r.m.density *= f;
    if (r.m.model != CRUSH_MAT_ELASTIC) {
        r.m.yield_stress *= f;
        r.m.hardening *= f;
        for (int i = 0; i < r.m.n_curve; ++i) r.m.curve[i][1] *= f;
        if (r.m.has_densification) r.m.densification[1] *= f;
    }
    return r;
}

// ---------------------------------------------------------------- model

crush_settings default_settings() {
    crush_settings s{};
    s.end_time = 0.15;
    s.dt_scale = 0.85;
    s.time_step = 0.0;
    s.adaptive_check_steps = 5;
    s.hourglass = 0.1;
    s.velocity_damping = 1.0;
    s.mass_scaling = 0.25;
    s.separation_steps = 0;
    // the accelerometer pulse at PULSE_DT (src/vehicle.rs:1102), as the
    // reference's head-on and barrier runs record it (history_steps = 1)
    s.pulse_dt = 1e-4;
    return s;
}

// Model::new (src/model.rs:168-172): every part starts as elastic(1, 0, 1).
Model::Model(Mesh m) : mesh(std::move(m)) {
    materials.assign(mesh.parts.size(), Material::elastic(1.0, 0.0, 1.0));
    initial_velocity.assign(mesh.nodes.size(), Vec3{0, 0, 0});
    settings = default_settings();
}

void Model::set_material(const std::string& part, const Material& mat) {
    int p = mesh.part_index(part);
    if (p < 0) fail(CRUSH_E_INPUT, "no part '%s'", part.c_str());
    materials[p] = mat;
}

void Model::set_initial_velocity(const std::string& set, Vec3 v) {
    std::vector<int> nodes;
    if (!mesh.nodes_of(set, nodes)) fail(CRUSH_E_INPUT, "no node set or part '%s'", set.c_str());
    for (int n : nodes) initial_velocity[n] = v;
}

void Model::fix(const std::string& set) {
    std::vector<int> nodes;
    if (!mesh.nodes_of(set, nodes)) fail(CRUSH_E_INPUT, "no node set or part '%s'", set.c_str());
    fixed_nodes.insert(fixed_nodes.end(), nodes.begin(), nodes.end());
}

// src/model.rs:240-245.
void Model::add_contact_one_way(const std::string& nodes, const std::string& faces, double k, double max_d) {
    Contact c;
    if (!mesh.face_set_nodes(nodes, c.nodes)) fail(CRUSH_E_INPUT, "no face set '%s'", nodes.c_str());
    auto it = mesh.face_sets.find(faces);
    if (it == mesh.face_sets.end()) fail(CRUSH_E_INPUT, "no face set '%s'", faces.c_str());
    c.faces = it->second;
    c.stiffness = k;
    c.max_distance = max_d;
    contacts.push_back(c);
}

// src/model.rs:257-265: each side's nodes against the other's faces at
// half the stiffness.
void Model::add_contact_pair_scaled(const std::string& a, const std::string& b, double k, double max_d,
                                    const std::vector<double>& sa, const std::vector<double>& sb) {
    const std::string* names[2][2] = {{&a, &b}, {&b, &a}};
    const std::vector<double>* scales[2] = {&sa, &sb};
    for (int side = 0; side < 2; ++side) {
        const std::string& p = *names[side][0];
        const std::string& s = *names[side][1];
        Contact c;
        if (!mesh.face_set_nodes(p, c.nodes)) fail(CRUSH_E_INPUT, "no face set '%s'", p.c_str());
        auto it = mesh.face_sets.find(s);
        if (it == mesh.face_sets.end()) fail(CRUSH_E_INPUT, "no face set '%s'", s.c_str());
        c.faces = it->second;
        if (!scales[side]->empty() && scales[side]->size() != c.nodes.size())
            fail(CRUSH_E_INPUT, "contact node scale length %zu vs %zu nodes of '%s'", scales[side]->size(),
                 c.nodes.size(), p.c_str());
        c.node_scale = *scales[side];
        c.stiffness = 0.5 * k;
        c.max_distance = max_d;
        c.pair = 1;
        contacts.push_back(c);
    }
}

ModelHolder::ModelHolder(Model m) : model(std::move(m)) { flatten(); }

void ModelHolder::flatten() {
    const Mesh& mesh = model.mesh;
    nodes.resize(3 * mesh.nodes.size());
    v0.resize(3 * mesh.nodes.size());
    for (size_t i = 0; i < mesh.nodes.size(); ++i)
        for (int d = 0; d < 3; ++d) {
            nodes[3 * i + d] = mesh.nodes[i][d];
            v0[3 * i + d] = model.initial_velocity[i][d];
        }
    hexes.resize(8 * mesh.hexes.size());
    for (size_t e = 0; e < mesh.hexes.size(); ++e)
        for (int k = 0; k < 8; ++k) hexes[8 * e + k] = mesh.hexes[e][k];
    hex_part.assign(mesh.hex_part.begin(), mesh.hex_part.end());
    faces.resize(4 * mesh.faces.size());
    for (size_t f = 0; f < mesh.faces.size(); ++f)
        for (int k = 0; k < 4; ++k) faces[4 * f + k] = mesh.faces[f][k];
    fixed.assign(model.fixed_nodes.begin(), model.fixed_nodes.end());
    materials.clear();
    for (const auto& m : model.materials) materials.push_back(m.m);
    contacts.clear();
    for (const auto& c : model.contacts) {
        crush_contact cc{};
        cc.n_nodes = (int32_t)c.nodes.size();
        cc.n_faces = (int32_t)c.faces.size();
        cc.nodes = c.nodes.data();
        cc.node_scale = c.node_scale.empty() ? nullptr : c.node_scale.data();
        cc.faces = c.faces.data();
        cc.stiffness = c.stiffness;
        cc.max_distance = c.max_distance;
        cc.pair = c.pair;
        contacts.push_back(cc);
    }
    groups.clear();
    for (const auto& g : model.groups) {
        crush_group cg{};
        cg.n_nodes = (int32_t)g.nodes.size();
        cg.n_front = (int32_t)g.front.size();
        cg.nodes = g.nodes.data();
        cg.front = g.front.data();
        cg.accel_node = g.accel_node;
        cg.axis = g.axis;
        cg.frame_origin = g.frame_origin;
        cg.frame_x = g.frame_x;
        cg.frame_plane = g.frame_plane;
        cg.heading = g.heading;
        groups.push_back(cg);
    }
    view = crush_model{};
    view.n_nodes = (int32_t)mesh.nodes.size();
    view.n_hexes = (int32_t)mesh.hexes.size();
    view.n_parts = (int32_t)mesh.parts.size();
    view.n_faces = (int32_t)mesh.faces.size();
    view.n_fixed = (int32_t)fixed.size();
    view.n_contacts = (int32_t)contacts.size();
    view.n_groups = (int32_t)groups.size();
    view.nodes = nodes.data();
    view.hexes = hexes.data();
    view.hex_part = hex_part.data();
    view.materials = materials.data();
    view.initial_velocity = v0.data();
    view.fixed_nodes = fixed.data();
    view.faces = faces.data();
    view.contacts = contacts.data();
    view.groups = groups.data();
    view.settings = model.settings;
}

// -------------------------------------------------------------- vehicles

namespace vehicle {

static const double TRANSVERSE_FACTOR = 4.0;  // src/vehicle.rs:493
static const double LOCK_COMPACTION = 1.4;    // src/vehicle.rs:504
static const double LOCK_SLOPE_FACTOR = 1.0;  // src/vehicle.rs:505

static void set_name(crush_vehicle& v, const std::string& n) {
    if (n.size() >= sizeof v.name) fail(CRUSH_E_INPUT, "vehicle name too long: '%s'", n.c_str());
    std::memset(v.name, 0, sizeof v.name);
    std::memcpy(v.name, n.data(), n.size());
}

// Vehicle::blank (src/vehicle.rs:172-193).
crush_vehicle blank() {
    crush_vehicle v;
    std::memset(&v, 0, sizeof v);
    v.transverse_factor = TRANSVERSE_FACTOR;
    return v;
}

static double area(const crush_vehicle& v) { return v.size[1] * v.size[2]; }

bool has_crush_zone(const crush_vehicle& v) { return v.crush_zone_length < v.size[0] - 1e-9; }
static bool has_rail_box(const crush_vehicle& v) { return v.has_rail_box && has_crush_zone(v); }
static double honeycomb_length(const crush_vehicle& v) {
    return v.crush_zone_length - (v.has_bumper ? v.bumper[0] : 0.0);
}
static double density(const crush_vehicle& v) { return v.mass / (v.size[0] * v.size[1] * v.size[2]); }
static double crush_density(const crush_vehicle& v) {
    if (v.has_crush_zone_mass && has_crush_zone(v))
        return (v.crush_zone_mass - (v.has_bumper ? v.bumper[1] : 0.0)) / (honeycomb_length(v) * area(v));
    return density(v);
}
static double body_density(const crush_vehicle& v) {
    if (v.has_crush_zone_mass && has_crush_zone(v))
        return (v.mass - v.crush_zone_mass) / ((v.size[0] - v.crush_zone_length) * area(v));
    return density(v);
}
static bool in_rail_box(const crush_vehicle& v, double y, double z) {
    if (!v.has_rail_box) return false;
    return std::fabs(y) < 0.5 * v.rail_box_width && z > v.rail_box_z[0] && z < v.rail_box_z[1];
}
static std::pair<double, double> rail_box_factors(const crush_vehicle& v, double area_fraction) {
    double f = v.has_rail_box ? std::min(std::max(v.rail_box_force_fraction, 0.0), 1.0) : 0.0;
    double a = std::min(std::max(area_fraction, 1e-6), 1.0 - 1e-6);
    return {f / a, (1.0 - f) / (1.0 - a)};
}

// src/vehicle.rs:384-390.
double contact_stiffness_per_node(const crush_vehicle& v, int n_face, double h) {
    double hf = has_crush_zone(v) ? (v.has_crush_element_size ? v.crush_element_size : h) : h;
    return v.modulus * area(v) / ((double)n_face * hf);
}

// src/vehicle.rs:397-436.
static std::vector<double> front_contact_scales(const crush_vehicle& v, const Mesh& mesh, double h) {
    if (!has_rail_box(v)) return {};
    std::vector<int> nodes;
    if (!mesh.face_set_nodes(std::string(v.name) + "_front", nodes)) return {};
    int ny = (int)std::max(std::round(v.size[1] / h), 1.0);
    int nz = (int)std::max(std::round(v.size[2] / h), 1.0);
    double dy = v.size[1] / ny, dz = v.size[2] / nz;
    double lo[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
    double hi[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
    for (int n : nodes)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], mesh.nodes[n][d]);
            hi[d] = std::max(hi[d], mesh.nodes[n][d]);
        }
    double y_center = 0.5 * (lo[1] + hi[1]), z_bottom = lo[2];
    int cells = ny * nz, n_rails = 0;
    for (int c = 0; c < cells; ++c)
        if (in_rail_box(v, (double)(c % ny) * dy + 0.5 * dy - 0.5 * v.size[1], (double)(c / ny) * dz + 0.5 * dz))
            ++n_rails;
    if (n_rails == 0 || n_rails == cells) return {};
    auto fr = rail_box_factors(v, (double)n_rails / (double)cells);
    std::vector<double> out;
    const double sy[4] = {-0.5, 0.5, -0.5, 0.5}, sz[4] = {-0.5, -0.5, 0.5, 0.5};
    for (int n : nodes) {
        const Vec3& p = mesh.nodes[n];
        double sum = 0.0;
        int count = 0;
        for (int q = 0; q < 4; ++q) {
            double yc = p[1] - y_center + sy[q] * dy, zc = p[2] - z_bottom + sz[q] * dz;
            if (std::fabs(yc) < 0.5 * v.size[1] && zc > 0.0 && zc < v.size[2]) {
                sum += in_rail_box(v, yc, zc) ? fr.first : fr.second;
                ++count;
            }
        }
        out.push_back(count > 0 ? sum / count : 1.0);
    }
    return out;
}

// Vehicle::yield_curve (src/vehicle.rs:509-511).
static std::vector<std::array<double, 2>> yield_curve(const crush_vehicle& v, double l, double a) {
    std::vector<std::array<double, 2>> r;
    for (int i = 0; i < v.n_force_table; ++i) {
        double x = v.force_table[i][0], f = v.force_table[i][1];
        r.push_back({-std::log(1.0 - std::min(x / l, 0.95)), f / a});
    }
    return r;
}

// Vehicle::crush_material / crush_material_axial (src/vehicle.rs:470-490).
Material crush_material(const crush_vehicle& v) {
    double a = area(v);
    Material m;
    if (v.n_force_table > 0) {
        std::vector<std::array<double, 2>> curve;
        if (v.n_compaction_map > 0 && v.n_compaction_map == v.n_force_table) {
            for (int i = 0; i < v.n_force_table; ++i) curve.push_back({v.compaction_map[i], v.force_table[i][1] / a});
        } else {
            curve = yield_curve(v, honeycomb_length(v), a);
        }
        m = Material::honeycomb_curve(v.modulus, crush_density(v), curve, true,
                                      {LOCK_COMPACTION, LOCK_SLOPE_FACTOR * v.modulus});
    } else {
        m = Material::honeycomb(v.modulus, crush_density(v), v.curve[0] / a, v.curve[1] * v.crush_zone_length / a);
    }
    m.m.transverse_factor = v.transverse_factor;
    return m;
}

static Material body_material(const crush_vehicle& v) {
    return Material::elastic(v.body_modulus, 0.3, body_density(v));
}

// CrushCurve::from_kw400 (src/vehicle.rs:49-53) and Vehicle::from_ncap
// (src/vehicle.rs:230-258).
static crush_vehicle from_ncap(const std::string& name, double mass, Vec3 size, double l_c, double kw400,
                               double a_over_b, double restitution, double test_speed) {
    double c = 2.0 * (0.4 - 0.025) / (0.4 * 0.4 - 0.025 * 0.025);
    double k = kw400 / (1.0 + c * a_over_b);
    double fy = k * a_over_b;
    double kinetic = 0.5 * mass * test_speed * test_speed;
    double qa = 0.5 * k, qb = fy;
    double c_max = (-qb + std::sqrt(qb * qb + 4.0 * qa * kinetic)) / (2.0 * qa);
    double f_max = fy + k * c_max;
    double ar = size[1] * size[2];
    crush_vehicle v = blank();
    set_name(v, name);
    v.mass = mass;
    for (int d = 0; d < 3; ++d) v.size[d] = size[d];
    v.crush_zone_length = l_c;
    v.curve[0] = fy;
    v.curve[1] = k;
    v.modulus = f_max * f_max * size[0] / (2.0 * ar * restitution * restitution * kinetic);
    v.body_modulus = v.modulus;
    return v;
}

static double crash3_ratio(double a, double b) { return a / b * 0.01; }

static void set_force_table(crush_vehicle& v, const double* knots, const double* force_kn, int n) {
    v.n_force_table = n;
    for (int i = 0; i < n; ++i) {
        v.force_table[i][0] = knots[i];
        v.force_table[i][1] = force_kn[i] * 1e3;
    }
}

// Presets: src/vehicle.rs:527-646.
crush_vehicle preset(const std::string& name, bool* found) {
    *found = true;
    auto neon = [] {
        crush_vehicle v = from_ncap("neon", 1354.0, {4.36, 1.71, 1.35}, 4.36, 1.251e6, crash3_ratio(75.65, 11.07), 0.12,
                                    35.0 * 0.44704);
        v.has_side_curve = 1;
        v.side_curve[0] = 230.0 * 4.44822 / 0.0254 * v.size[0];
        v.side_curve[1] = 187.0 * 6894.76 * v.size[0];
        return v;
    };
    auto neon_tuned = [&] {
        crush_vehicle v = neon();
        v.curve[0] = 16.73e3;
        v.curve[1] = 2.930e5;
        v.modulus = 23.1e6;
        v.body_modulus = v.modulus;
        v.side_curve[0] = 2.75e3;
        v.side_curve[1] = 4.657e6;
        v.has_side_modulus = 1;
        v.side_modulus = 125.8e6;
        return v;
    };
    auto silverado = [] {
        return from_ncap("silverado", 2622.0, {5.85, 2.03, 1.87}, 5.85, 2.550e6, crash3_ratio(86.36, 11.72), 0.12,
                         35.0 * 0.44704);
    };
    auto expedition = [] {
        crush_vehicle v = from_ncap("expedition", 2460.0, {5.2, 2.0, 1.9}, 5.2, 2.5e6, crash3_ratio(86.36, 11.72),
                                    0.12, 30.0 * 0.44704);
        v.crush_zone_length = 1.7;
        v.has_crush_element_size = 1;
        v.crush_element_size = 0.1;
        v.has_crush_zone_mass = 1;
        v.crush_zone_mass = 320.0;
        v.has_bumper = 1;
        v.bumper[0] = 0.1;
        v.bumper[1] = 35.0;
        v.bumper[2] = 3.0e8;
        v.has_rail_box = 1;
        v.rail_box_width = 1.25;
        v.rail_box_z[0] = 0.15;
        v.rail_box_z[1] = 0.72;
        v.rail_box_force_fraction = 0.75;
        v.modulus = 7.3e6;
        v.body_modulus = 3.93e8;
        const double knots[7] = {0.0, 0.1064, 0.2128, 0.3192, 0.4256, 0.5320, 0.6384};
        const double force[7] = {29.0, 183.0, 259.0, 417.0, 457.0, 457.0, 457.0};
        set_force_table(v, knots, force, 7);
        return v;
    };
    if (name == "silverado") {
        crush_vehicle v = silverado();
        v.curve[0] = 51.21e3;
        v.curve[1] = 5.592e5;
        v.modulus = 27.8e6;
        v.body_modulus = v.modulus;
        return v;
    }
    if (name == "silverado-raw") return silverado();
    if (name == "neon") return neon_tuned();
    if (name == "neon-raw") return neon();
    if (name == "neon-pulse") {
        crush_vehicle v = neon_tuned();
        v.crush_zone_length = 1.5;
        v.has_crush_element_size = 1;
        v.crush_element_size = 0.1;
        v.has_crush_zone_mass = 1;
        v.crush_zone_mass = 190.0;
        v.has_bumper = 1;
        v.bumper[0] = 0.1;
        v.bumper[1] = 25.0;
        v.bumper[2] = 3.0e8;
        v.has_rail_box = 1;
        v.rail_box_width = 0.45;
        v.rail_box_z[0] = 0.06;
        v.rail_box_z[1] = 0.41;
        v.rail_box_force_fraction = 0.8;
        v.modulus = 4e6;
        v.body_modulus = 2.7e8;
        const double knots[7] = {0.0, 0.1165, 0.2330, 0.3496, 0.4661, 0.5826, 0.6991};
        const double force[7] = {13.0, 135.0, 135.0, 146.0, 361.0, 364.0, 364.0};
        set_force_table(v, knots, force, 7);
        return v;
    }
    if (name == "expedition") return expedition();
    if (name == "navigator") {
        crush_vehicle v = expedition();
        set_name(v, "navigator");
        v.mass = 2873.0;
        v.size[0] = 5.175;
        v.size[1] = 2.049;
        v.size[2] = 1.9;
        return v;
    }
    *found = false;
    return blank();
}

static double num(const toml::Value& v, const std::string& key) {
    if (v.kind != toml::Value::Number) fail(CRUSH_E_INPUT, "vehicle toml: '%s' must be a number", key.c_str());
    return v.num;
}

static void num_array(const toml::Value& v, const std::string& key, double* out, int n) {
    if (v.kind != toml::Value::Array || (int)v.items.size() != n)
        fail(CRUSH_E_INPUT, "vehicle toml: '%s' must be an array of %d numbers", key.c_str(), n);
    for (int i = 0; i < n; ++i) out[i] = num(v.items[i], key);
}

// Vehicle::load (src/vehicle.rs:212-219) over the serde schema of
// src/vehicle.rs:105-168 with Vehicle::blank defaults.
crush_vehicle load_toml(const std::string& path) {
    std::ifstream f(path);
    if (!f) fail(CRUSH_E_NOTFOUND, "%s: cannot open", path.c_str());
    std::stringstream ss;
    ss << f.rdbuf();
    auto kv = toml::parse(ss.str());
    crush_vehicle v = blank();
    for (const auto& e : kv) {
        const std::string& k = e.first;
        const toml::Value& val = e.second;
        if (k == "name") {
            if (val.kind != toml::Value::String) fail(CRUSH_E_INPUT, "%s: name must be a string", path.c_str());
            set_name(v, val.str);
        } else if (k == "mass") v.mass = num(val, k);
        else if (k == "size") num_array(val, k, v.size, 3);
        else if (k == "crush_zone_length") v.crush_zone_length = num(val, k);
        else if (k == "curve.yield_force") v.curve[0] = num(val, k);
        else if (k == "curve.stiffness") v.curve[1] = num(val, k);
        else if (k == "modulus") v.modulus = num(val, k);
        else if (k == "body_modulus") v.body_modulus = num(val, k);
        else if (k == "side_curve.yield_force") { v.has_side_curve = 1; v.side_curve[0] = num(val, k); }
        else if (k == "side_curve.stiffness") { v.has_side_curve = 1; v.side_curve[1] = num(val, k); }
        else if (k == "side_modulus") { v.has_side_modulus = 1; v.side_modulus = num(val, k); }
        else if (k == "rear_curve.yield_force") { v.has_rear_curve = 1; v.rear_curve[0] = num(val, k); }
        else if (k == "rear_curve.stiffness") { v.has_rear_curve = 1; v.rear_curve[1] = num(val, k); }
        else if (k == "rear_modulus") { v.has_rear_modulus = 1; v.rear_modulus = num(val, k); }
        else if (k == "force_table") {
            if (val.kind != toml::Value::Array || val.items.size() > CRUSH_MAX_TABLE)
                fail(CRUSH_E_INPUT, "%s: force_table must hold at most %d knots", path.c_str(), CRUSH_MAX_TABLE);
            v.n_force_table = (int)val.items.size();
            for (size_t i = 0; i < val.items.size(); ++i) num_array(val.items[i], k, v.force_table[i], 2);
        } else if (k == "compaction_map") {
            if (val.kind != toml::Value::Array || val.items.size() > CRUSH_MAX_TABLE)
                fail(CRUSH_E_INPUT, "%s: compaction_map must hold at most %d values", path.c_str(), CRUSH_MAX_TABLE);
            v.n_compaction_map = (int)val.items.size();
            for (size_t i = 0; i < val.items.size(); ++i) v.compaction_map[i] = num(val.items[i], k);
        } else if (k == "crush_element_size") { v.has_crush_element_size = 1; v.crush_element_size = num(val, k); }
        else if (k == "transverse_factor") v.transverse_factor = num(val, k);
        else if (k == "bumper") { v.has_bumper = 1; num_array(val, k, v.bumper, 3); }
        else if (k == "crush_zone_mass") { v.has_crush_zone_mass = 1; v.crush_zone_mass = num(val, k); }
        else if (k == "rail_box.width") { v.has_rail_box = 1; v.rail_box_width = num(val, k); }
        else if (k == "rail_box.z_range") { v.has_rail_box = 1; num_array(val, k, v.rail_box_z, 2); }
        else if (k == "rail_box.force_fraction") { v.has_rail_box = 1; v.rail_box_force_fraction = num(val, k); }
        else fail(CRUSH_E_INPUT, "%s: unknown vehicle field '%s'", path.c_str(), k.c_str());
    }
    bool bad_size = v.size[0] <= 0.0 || v.size[1] <= 0.0 || v.size[2] <= 0.0;
    if (v.name[0] == '\0' || v.mass <= 0.0 || bad_size || v.modulus <= 0.0)
        fail(CRUSH_E_INPUT, "%s: a vehicle needs name, mass, size and modulus", path.c_str());
    return v;
}

// Vehicle::side_profile (src/vehicle.rs:271-284).
crush_vehicle side_profile(const crush_vehicle& s) {
    crush_vehicle v = blank();
    set_name(v, std::string(s.name) + "_side");
    v.mass = s.mass;
    v.size[0] = s.size[1];
    v.size[1] = s.size[0];
    v.size[2] = s.size[2];
    v.crush_zone_length = s.size[1];
    v.curve[0] = s.has_side_curve ? s.side_curve[0] : s.curve[0];
    v.curve[1] = s.has_side_curve ? s.side_curve[1] : s.curve[1];
    v.modulus = s.has_side_modulus ? s.side_modulus : s.modulus;
    v.body_modulus = v.modulus;
    return v;
}

// Vehicle::rear_profile (src/vehicle.rs:297-316).
bool rear_profile(const crush_vehicle& s, crush_vehicle& v) {
    if (!s.has_rear_curve) return false;
    v = s;
    set_name(v, std::string(s.name) + "_rear");
    v.crush_zone_length = s.size[0];
    v.curve[0] = s.rear_curve[0];
    v.curve[1] = s.rear_curve[1];
    v.modulus = s.has_rear_modulus ? s.rear_modulus : s.modulus;
    v.body_modulus = v.modulus;
    v.has_side_curve = v.has_side_modulus = v.has_rear_curve = v.has_rear_modulus = 0;
    v.n_force_table = 0;
    v.has_crush_element_size = 0;
    v.n_compaction_map = 0;
    v.has_crush_zone_mass = 0;
    v.has_bumper = 0;
    v.has_rail_box = 0;
    return true;
}

}  // namespace vehicle

// ---------------------------------------------------------- builders

// add_vehicle: src/vehicle.rs:663-744.
int add_vehicle(Mesh& mesh, const crush_vehicle& v, double front_x, double y_center, double z_bottom,
                Heading heading, double h) {
    const std::string name = v.name;
    double x0 = heading == Heading::PlusX ? front_x - v.size[0] : front_x;
    BlockFace face = heading == Heading::PlusX ? BlockFace::XMax : BlockFace::XMin;
    std::string front = name + "_front", bottom = name + "_bottom";
    Vec3 origin = {x0, y_center - v.size[1] / 2.0, z_bottom};
    Vec3 size = {v.size[0], v.size[1], v.size[2]};
    if (!vehicle::has_crush_zone(v))
        return mesh.add_hex_block(name, origin, size, h, {{face, front}, {BlockFace::ZMin, bottom}});
    double t_b = v.has_bumper ? v.bumper[0] : 0.0;
    double l_c = v.crush_zone_length - t_b;
    double l_b = v.size[0] - v.crush_zone_length;
    int ny = (int)std::max(std::round(v.size[1] / h), 1.0);
    int nz = (int)std::max(std::round(v.size[2] / h), 1.0);
    int nxb = (int)std::max(std::round(l_b / h), 1.0);
    int nxc = (int)std::max(std::round(l_c / (v.has_crush_element_size ? v.crush_element_size : h)), 1.0);
    double body_x0, crush_x0, bumper_x0;
    if (heading == Heading::PlusX) {
        body_x0 = x0;
        crush_x0 = x0 + l_b;
        bumper_x0 = x0 + l_b + l_c;
    } else {
        body_x0 = x0 + l_c + t_b;
        crush_x0 = x0 + t_b;
        bumper_x0 = x0;
    }
    auto sub = [&](const char* n) { return name + "_bottom_" + n; };
    int part = mesh.add_hex_block_n(name, {body_x0, origin[1], origin[2]}, {l_b, v.size[1], v.size[2]}, {nxb, ny, nz},
                                    {{BlockFace::ZMin, sub("body")}});
    std::string crush_name = name + "_crush";
    std::vector<std::string> names = {crush_name};
    std::vector<std::string> bottoms = {sub("body"), sub("crush")};
    int crush_part;
    if (t_b > 0.0) {
        crush_part = mesh.add_hex_block_n(crush_name, {crush_x0, origin[1], origin[2]}, {l_c, v.size[1], v.size[2]},
                                          {nxc, ny, nz}, {{BlockFace::ZMin, sub("crush")}});
        std::string bumper_name = name + "_bumper";
        mesh.add_hex_block_n(bumper_name, {bumper_x0, origin[1], origin[2]}, {t_b, v.size[1], v.size[2]}, {1, ny, nz},
                             {{face, front}, {BlockFace::ZMin, sub("bumper")}});
        names.push_back(bumper_name);
        bottoms.push_back(sub("bumper"));
    } else {
        crush_part = mesh.add_hex_block_n(crush_name, {crush_x0, origin[1], origin[2]}, {l_c, v.size[1], v.size[2]},
                                          {nxc, ny, nz}, {{face, front}, {BlockFace::ZMin, sub("crush")}});
    }
    std::vector<int> bottom_faces;
    for (const auto& b : bottoms) {
        auto it = mesh.face_sets.find(b);
        if (it != mesh.face_sets.end()) {
            bottom_faces.insert(bottom_faces.end(), it->second.begin(), it->second.end());
            mesh.face_sets.erase(it);
        }
    }
    mesh.face_sets[bottom] = bottom_faces;
    if (v.has_rail_box && vehicle::has_crush_zone(v)) {
        int bumper_part = mesh.part_index(name + "_bumper");
        std::vector<std::pair<int, std::string>> carve = {{crush_part, name + "_rails"}};
        if (bumper_part >= 0) carve.push_back({bumper_part, name + "_bumper_rails"});
        for (const auto& cv : carve) {
            int to = mesh.add_part(cv.second);
            for (size_t e = 0; e < mesh.hexes.size(); ++e) {
                if (mesh.hex_part[e] != cv.first) continue;
                double ys = 0.0, zs = 0.0;
                for (int k = 0; k < 8; ++k) ys += mesh.nodes[mesh.hexes[e][k]][1];
                for (int k = 0; k < 8; ++k) zs += mesh.nodes[mesh.hexes[e][k]][2];
                double yc = ys / 8.0 - y_center, zc = zs / 8.0 - z_bottom;
                if (vehicle::in_rail_box(v, yc, zc)) mesh.hex_part[e] = to;
            }
        }
    }
    mesh.merge_coincident_nodes(1e-6 * h);
    std::vector<int> all = mesh.node_sets[name];
    for (const auto& n : names) {
        const auto& s = mesh.node_sets[n];
        all.insert(all.end(), s.begin(), s.end());
    }
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    mesh.node_sets[name] = all;
    return part;
}

// assign_vehicle: src/vehicle.rs:750-787.
void assign_vehicle(Model& model, const crush_vehicle& v, Heading heading, double speed) {
    const std::string name = v.name;
    std::string crush_name = name + "_crush";
    if (vehicle::has_crush_zone(v) && model.mesh.part_index(crush_name) >= 0) {
        model.set_material(name, vehicle::body_material(v));
        Material crush = vehicle::crush_material(v);
        bool has_plate = v.has_bumper != 0;
        Material plate;
        if (has_plate) plate = Material::elastic(v.bumper[2], 0.3, v.bumper[1] / (v.bumper[0] * vehicle::area(v)));
        int rails = vehicle::has_rail_box(v) ? model.mesh.part_index(name + "_rails") : -1;
        double f_rails = 1.0, f_skin = 1.0;
        if (rails >= 0) {
            int skin = model.mesh.part_index(crush_name);
            double n_rails = (double)model.mesh.part_hexes(rails).size();
            double n_skin = (double)model.mesh.part_hexes(skin).size();
            if (n_rails > 0.0 && n_skin > 0.0) {
                auto fr = vehicle::rail_box_factors(v, n_rails / (n_rails + n_skin));
                f_rails = fr.first;
                f_skin = fr.second;
            }
            model.set_material(name + "_rails", crush.scaled(f_rails));
        }
        model.set_material(crush_name, crush.scaled(f_skin));
        if (has_plate) {
            if (model.mesh.part_index(name + "_bumper") >= 0) model.set_material(name + "_bumper", plate.scaled(f_skin));
            if (model.mesh.part_index(name + "_bumper_rails") >= 0)
                model.set_material(name + "_bumper_rails", plate.scaled(f_rails));
        }
    } else {
        model.set_material(name, vehicle::crush_material(v));
    }
    double vx = heading == Heading::PlusX ? speed : -speed;
    model.set_initial_velocity(name, {vx, 0.0, 0.0});
}

// add_ground: src/vehicle.rs:829-833.
int add_ground(Mesh& mesh, const std::string& name, double z, std::array<double, 2> x, std::array<double, 2> y,
               double thickness) {
    double h = std::max(std::min(x[1] - x[0], y[1] - y[0]) / 8.0, thickness);
    return mesh.add_hex_block(name, {x[0], y[0], z - thickness}, {x[1] - x[0], y[1] - y[0], thickness}, h,
                              {{BlockFace::ZMax, name + "_face"}});
}

// add_ground_contact: src/vehicle.rs:838-847.
void add_ground_contact(Model& model, const crush_vehicle& v, const std::string& ground, double clearance, double h) {
    if (model.mesh.part_index(ground) < 0) return;
    const std::string name = v.name;
    std::vector<int> bottom_nodes, front_nodes;
    int n_bottom = model.mesh.face_set_nodes(name + "_bottom", bottom_nodes) ? (int)bottom_nodes.size() : 1;
    int n_face = model.mesh.face_set_nodes(name + "_front", front_nodes) ? (int)front_nodes.size() : n_bottom;
    model.set_material(ground, Material::elastic(1e9, 0.3, 1000.0));
    model.fix(ground);
    model.add_contact_one_way(name + "_bottom", ground + "_face", vehicle::contact_stiffness_per_node(v, n_face, h),
                              clearance + 0.3);
}

// add_rigid_wall: src/vehicle.rs:851-859.
int add_rigid_wall(Mesh& mesh, const std::string& name, double x, std::array<double, 2> y, std::array<double, 2> z,
                   Heading facing, double thickness) {
    double x0 = facing == Heading::PlusX ? x - thickness : x;
    BlockFace face = facing == Heading::PlusX ? BlockFace::XMax : BlockFace::XMin;
    double h = std::max(std::min(y[1] - y[0], z[1] - z[0]) / 4.0, thickness);
    return mesh.add_hex_block(name, {x0, y[0], z[0]}, {thickness, y[1] - y[0], z[1] - z[0]}, h,
                              {{face, name + "_face"}});
}

// add_vehicle_accelerometer (src/vehicle.rs:792-822) -> the accelerometer
// node chosen by add_accelerometer_at (src/model.rs:185-187).
int vehicle_accel_node(const Model& model, const std::string& part, const std::string& front_face) {
    int p = model.mesh.part_index(part);
    if (p < 0) fail(CRUSH_E_INPUT, "no part '%s'", part.c_str());
    std::vector<int> nodes = model.mesh.part_nodes(p);
    double lo[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
    double hi[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
    for (int n : nodes)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], model.mesh.nodes[n][d]);
            hi[d] = std::max(hi[d], model.mesh.nodes[n][d]);
        }
    std::vector<int> front;
    if (!model.mesh.face_set_nodes(front_face, front)) fail(CRUSH_E_INPUT, "no face set '%s'", front_face.c_str());
    double c[3] = {0, 0, 0};
    for (int n : front)
        for (int d = 0; d < 3; ++d) c[d] += model.mesh.nodes[n][d] / (double)front.size();
    Vec3 at = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), lo[2] + 0.35 * (hi[2] - lo[2])};
    for (int d = 0; d < 2; ++d) {
        double len = hi[d] - lo[d];
        if (c[d] <= lo[d] + 1e-6 * len)
            at[d] = hi[d] - 0.25 * len;
        else if (c[d] >= hi[d] - 1e-6 * len)
            at[d] = lo[d] + 0.25 * len;
    }
    return model.mesh.nearest_node(at, p);
}

// The frame of add_accelerometer_at (src/model.rs:185-213) at node rec of
// part p: local x from the origin to the x node, which is rec's neighbour
// best aligned with +X (within 1.75 x the shortest edge at rec, cosine >
// 0.5), or else rec itself with its best -X neighbour as the origin; the
// plane node is the origin's neighbour best aligned with +Y (cosine > 0.5),
// else with -Y, other than the x node.  Returns (origin, x node, plane).
static std::array<int, 3> accelerometer_frame(const Mesh& mesh, int rec, int p) {
    const double radius = 1.75 * mesh.min_edge_at(rec);
    int origin = rec, x_axis = rec;
    auto fwd = mesh.best_aligned_neighbour(rec, {1.0, 0.0, 0.0}, radius, p, {});
    if (fwd.first >= 0 && fwd.second > 0.5) {
        x_axis = fwd.first;
    } else {
        auto back = mesh.best_aligned_neighbour(rec, {-1.0, 0.0, 0.0}, radius, p, {});
        if (back.first >= 0) origin = back.first;
    }
    if (origin == x_axis) fail(CRUSH_E_INPUT, "accelerometer: part '%s' has no neighbouring nodes along X", mesh.parts[p].c_str());
    int plane = -1;
    auto up = mesh.best_aligned_neighbour(origin, {0.0, 1.0, 0.0}, radius, p, {x_axis});
    if (up.first >= 0 && up.second > 0.5) {
        plane = up.first;
    } else {
        auto down = mesh.best_aligned_neighbour(origin, {0.0, -1.0, 0.0}, radius, p, {x_axis});
        if (down.first < 0)
            fail(CRUSH_E_INPUT, "accelerometer: part '%s' has no neighbouring nodes to define a frame", mesh.parts[p].c_str());
        plane = down.first;
    }
    return {origin, x_axis, plane};
}

Group make_group(const Model& model, const std::string& name, const std::string& front_face, int axis,
                 int heading) {
    Group g;
    if (!model.mesh.nodes_of(name, g.nodes)) fail(CRUSH_E_INPUT, "no node set or part '%s'", name.c_str());
    if (!model.mesh.face_set_nodes(front_face, g.front)) fail(CRUSH_E_INPUT, "no face set '%s'", front_face.c_str());
    g.accel_node = vehicle_accel_node(model, name, front_face);
    auto fr = accelerometer_frame(model.mesh, g.accel_node, model.mesh.part_index(name));
    g.frame_origin = fr[0];
    g.frame_x = fr[1];
    g.frame_plane = fr[2];
    g.axis = axis;
    g.heading = heading;
    return g;
}

// HeadOn::build: src/headon.rs:40-66.
Model build_headon(const crush_vehicle& a, const crush_vehicle& b, double speed_a, double speed_b,
                   const crush_headon_options& opt) {
    double h = opt.element_size;
    double gap = 0.005;
    Mesh mesh;
    add_vehicle(mesh, a, -gap, 0.0, 0.0, Heading::PlusX, h);
    add_vehicle(mesh, b, gap, opt.offset, 0.0, Heading::MinusX, h);
    bool road = opt.clearance >= 0.0;
    if (road) {
        double w = std::max(a.size[1], b.size[1]) + 2.0 * std::fabs(opt.offset);
        add_ground(mesh, "ground", -opt.clearance, {-a.size[0] - 3.0, b.size[0] + 3.0}, {-w, w}, 0.3);
    }
    Model model(std::move(mesh));
    assign_vehicle(model, a, Heading::PlusX, speed_a);
    assign_vehicle(model, b, Heading::MinusX, speed_b);
    if (road) {
        add_ground_contact(model, a, "ground", opt.clearance, h);
        add_ground_contact(model, b, "ground", opt.clearance, h);
    }
    std::string fa = std::string(a.name) + "_front", fb = std::string(b.name) + "_front";
    std::vector<int> na, nb;
    if (!model.mesh.face_set_nodes(fa, na) || !model.mesh.face_set_nodes(fb, nb))
        fail(CRUSH_E_INPUT, "vehicle front face sets missing");
    int n_face = (int)std::min(na.size(), nb.size());
    auto sa = vehicle::front_contact_scales(a, model.mesh, h);
    auto sb = vehicle::front_contact_scales(b, model.mesh, h);
    double k = std::min(vehicle::contact_stiffness_per_node(a, n_face, h), vehicle::contact_stiffness_per_node(b, n_face, h));
    model.add_contact_pair_scaled(fa, fb, k, 0.3, sa, sb);
    model.settings.end_time = opt.end_time;
    model.groups.push_back(make_group(model, a.name, fa, 0, 1));
    model.groups.push_back(make_group(model, b.name, fb, 0, -1));
    return model;
}

// `crushrs crash`: src/bin/crushrs.rs:403-432.
Model build_crash(double truck_mph, double car_mph) {
    bool found;
    crush_vehicle truck = vehicle::preset("silverado", &found);
    crush_vehicle car = vehicle::preset("neon", &found);
    double h = TUNED_ELEMENT_SIZE, gap = 0.005;
    Mesh mesh;
    add_vehicle(mesh, truck, -gap, 0.0, 0.0, Heading::PlusX, h);
    add_vehicle(mesh, car, gap, 0.0, 0.0, Heading::MinusX, h);
    Model model(std::move(mesh));
    assign_vehicle(model, truck, Heading::PlusX, truck_mph * MPH);
    assign_vehicle(model, car, Heading::MinusX, car_mph * MPH);
    std::vector<int> nf;
    model.mesh.face_set_nodes("neon_front", nf);
    int n_face = (int)nf.size();
    double k = std::min(vehicle::contact_stiffness_per_node(car, n_face, h),
                        vehicle::contact_stiffness_per_node(truck, n_face, h));
    model.add_contact_pair_scaled("silverado_front", "neon_front", k, 0.3, {}, {});
    model.settings.end_time = 0.15;
    model.groups.push_back(make_group(model, "silverado", "silverado_front", 0, 1));
    model.groups.push_back(make_group(model, "neon", "neon_front", 0, -1));
    return model;
}

// `crushrs tbone`: src/bin/crushrs.rs:550-573 (any striking vehicle, any
// struck vehicle's side profile).
Model build_tbone(const crush_vehicle& truck, const crush_vehicle& car, double speed_a, double speed_b,
                  double offset, double clearance, double h) {
    crush_vehicle side = vehicle::side_profile(car);
    double gap = 0.005;
    Mesh mesh;
    add_vehicle(mesh, truck, -gap, 0.0, 0.0, Heading::PlusX, h);
    std::string side_face = std::string(car.name) + "_side";
    Vec3 car_origin = {gap, -offset - side.size[1] / 2.0, 0.0};
    mesh.add_hex_block(side.name, car_origin, {side.size[0], side.size[1], side.size[2]}, h,
                       {{BlockFace::XMin, side_face}, {BlockFace::ZMin, side_face + "_bottom"}});
    bool road = clearance >= 0.0;
    if (road) {
        double w = side.size[1] + truck.size[1];
        add_ground(mesh, "ground", -clearance, {-truck.size[0] - 3.0, side.size[0] + 3.0}, {-w, w}, 0.3);
    }
    Model model(std::move(mesh));
    assign_vehicle(model, truck, Heading::PlusX, speed_a);
    model.set_material(side.name, vehicle::crush_material(side));
    model.set_initial_velocity(side.name, {0.0, speed_b, 0.0});
    if (road) {
        add_ground_contact(model, truck, "ground", clearance, h);
        // The side block's underbody set is "<car>_side_bottom", which is
        // the "<side name>_bottom" set add_ground_contact looks up.
        add_ground_contact(model, side, "ground", clearance, h);
    }
    std::vector<int> nf;
    model.mesh.face_set_nodes(side_face, nf);
    int n_face = (int)nf.size();
    double k = std::min(vehicle::contact_stiffness_per_node(side, n_face, h),
                        vehicle::contact_stiffness_per_node(truck, n_face, h));
    std::string tf = std::string(truck.name) + "_front";
    model.add_contact_pair_scaled(tf, side_face, k, 0.3, {}, {});
    model.settings.end_time = 0.15;
    model.groups.push_back(make_group(model, truck.name, tf, 0, 1));
    model.groups.push_back(make_group(model, side.name, side_face, 0, 0));
    return model;
}

// barrier_model: src/vehicle.rs:951-973.
Model build_barrier(const crush_vehicle& v, double h, double speed, double end_time) {
    double gap = 0.005;
    Mesh mesh;
    add_vehicle(mesh, v, -gap, 0.0, 0.0, Heading::PlusX, h);
    double w = v.size[1], z = v.size[2];
    add_rigid_wall(mesh, "wall", 0.0, {-w * 0.75, w * 0.75}, {-0.2 * z, 1.2 * z}, Heading::MinusX, 0.2);
    add_ground(mesh, "ground", -GROUND_CLEARANCE, {-v.size[0] - 1.0, -0.3}, {-w, w}, 0.3);
    Model model(std::move(mesh));
    assign_vehicle(model, v, Heading::PlusX, speed);
    model.set_material("wall", Material::elastic(1e9, 0.3, 1000.0));
    model.fix("wall");
    add_ground_contact(model, v, "ground", GROUND_CLEARANCE, h);
    std::string front = std::string(v.name) + "_front";
    std::vector<int> nf;
    model.mesh.face_set_nodes(front, nf);
    auto scales = vehicle::front_contact_scales(v, model.mesh, h);
    model.add_contact_pair_scaled(front, "wall_face", vehicle::contact_stiffness_per_node(v, (int)nf.size(), h), 0.3,
                                  scales, {});
    model.settings.end_time = end_time;
    model.groups.push_back(make_group(model, v.name, front, 0, 1));
    return model;
}

}  // namespace crush

// ------------------------------------------------------------------- ABI

using namespace crush;

extern "C" {

void crush_headon_defaults(crush_headon_options* opt) {
    if (!opt) return;
    opt->offset = 0.0;
    opt->clearance = GROUND_CLEARANCE;
    opt->element_size = TUNED_ELEMENT_SIZE;
    opt->end_time = PULSE_END;
}

int32_t crush_vehicle_load(const char* name_or_path, crush_vehicle* out) {
    return guarded([&] {
        if (!name_or_path || !out) fail(CRUSH_E_INPUT, "null argument");
        std::string s = name_or_path;
        if (s.size() > 5 && s.compare(s.size() - 5, 5, ".toml") == 0) {
            *out = vehicle::load_toml(s);
            return;
        }
        bool found = false;
        crush_vehicle v = vehicle::preset(s, &found);
        if (!found)
            fail(CRUSH_E_NOTFOUND,
                 "unknown vehicle '%s' (neon | neon-raw | neon-pulse | silverado | silverado-raw | expedition | "
                 "navigator, or a vehicle .toml)",
                 s.c_str());
        *out = v;
    });
}

int32_t crush_vehicle_side_profile(const crush_vehicle* v, crush_vehicle* out) {
    return guarded([&] {
        if (!v || !out) fail(CRUSH_E_INPUT, "null argument");
        *out = vehicle::side_profile(*v);
    });
}

int32_t crush_vehicle_rear_profile(const crush_vehicle* v, crush_vehicle* out) {
    return guarded([&] {
        if (!v || !out) fail(CRUSH_E_INPUT, "null argument");
        if (!vehicle::rear_profile(*v, *out))
            fail(CRUSH_E_NOTFOUND, "%s has no rear curve: tune it with a [rear] section", v->name);
    });
}

int32_t crush_scenario_headon(const crush_vehicle* a, const crush_vehicle* b, double speed_a, double speed_b,
                              const crush_headon_options* opt, crush_model_handle** out) {
    return guarded([&] {
        if (!a || !b || !out) fail(CRUSH_E_INPUT, "null argument");
        crush_headon_options o;
        crush_headon_defaults(&o);
        if (opt) o = *opt;
        if (!(o.element_size > 0.0)) fail(CRUSH_E_INPUT, "element_size must be positive");
        if (std::strcmp(a->name, b->name) == 0)
            fail(CRUSH_E_INPUT, "vehicles need distinct names (the reference suffixes _a / _b)");
        *out = new crush_model_handle(build_headon(*a, *b, speed_a, speed_b, o));
    });
}

int32_t crush_scenario_crash(double truck_mph, double car_mph, crush_model_handle** out) {
    return guarded([&] {
        if (!out) fail(CRUSH_E_INPUT, "null argument");
        *out = new crush_model_handle(build_crash(truck_mph, car_mph));
    });
}

int32_t crush_scenario_tbone(const crush_vehicle* striking, const crush_vehicle* struck, double speed_a,
                             double speed_b, double offset, double clearance, double element_size,
                             crush_model_handle** out) {
    return guarded([&] {
        if (!striking || !struck || !out) fail(CRUSH_E_INPUT, "null argument");
        if (!(element_size > 0.0)) fail(CRUSH_E_INPUT, "element_size must be positive");
        *out = new crush_model_handle(build_tbone(*striking, *struck, speed_a, speed_b, offset, clearance, element_size));
    });
}

int32_t crush_scenario_barrier(const crush_vehicle* v, double element_size, double speed, double end_time,
                               crush_model_handle** out) {
    return guarded([&] {
        if (!v || !out) fail(CRUSH_E_INPUT, "null argument");
        if (!(element_size > 0.0)) fail(CRUSH_E_INPUT, "element_size must be positive");
        *out = new crush_model_handle(build_barrier(*v, element_size, speed, end_time));
    });
}

const crush_model* crush_model_view(const crush_model_handle* h) { return h ? &h->holder.view : nullptr; }

int32_t crush_model_set_settings(crush_model_handle* h, const crush_settings* s) {
    return guarded([&] {
        if (!h || !s) fail(CRUSH_E_INPUT, "null argument");
        h->holder.model.settings = *s;
        h->holder.view.settings = *s;
    });
}

int32_t crush_model_node_set(const crush_model_handle* h, const char* name, int32_t face_set, const int32_t** nodes,
                             int32_t* n) {
    return guarded([&] {
        if (!h || !name || !nodes || !n) fail(CRUSH_E_INPUT, "null argument");
        // Kept alive in a per-handle cache so the pointer stays valid.
        static thread_local std::map<std::pair<const void*, std::string>, std::vector<int>> cache;
        std::vector<int> out;
        bool ok = face_set ? h->holder.model.mesh.face_set_nodes(name, out) : h->holder.model.mesh.nodes_of(name, out);
        if (!ok) fail(CRUSH_E_NOTFOUND, "no %s '%s'", face_set ? "face set" : "node set or part", name);
        auto& slot = cache[{(const void*)h, std::string(name) + (face_set ? "#f" : "#n")}];
        slot = out;
        *nodes = slot.data();
        *n = (int32_t)slot.size();
    });
}

int32_t crush_model_part_names(const crush_model_handle* h, char* buf, int32_t cap) {
    return guarded([&] {
        if (!h || !buf || cap <= 0) fail(CRUSH_E_INPUT, "null argument");
        std::string s;
        for (size_t p = 0; p < h->holder.model.mesh.parts.size(); ++p) {
            if (p) s += '\n';
            s += h->holder.model.mesh.parts[p];
        }
        std::strncpy(buf, s.c_str(), (size_t)cap - 1);
        buf[cap - 1] = '\0';
    });
}

void crush_model_free(crush_model_handle* h) { delete h; }

}  // extern "C"
