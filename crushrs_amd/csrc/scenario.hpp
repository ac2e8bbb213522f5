// Host-side model construction: mesh generator, materials, vehicles and the
// scenario builders, owning the arrays a flattened crush_model points into.
#pragma once

#include <array>
#include <map>
#include <string>
#include <vector>

#include "../../include/crushrs.h"
#include "../../include/crushrs_scenario.h"

namespace crush {

using Vec3 = std::array<double, 3>;

enum class BlockFace { XMin, XMax, YMin, YMax, ZMin, ZMax };

// src/mesh.rs:37-52 (Mesh).
struct Mesh {
    std::vector<Vec3> nodes;
    std::vector<std::array<int, 8>> hexes;
    std::vector<int> hex_part;
    std::vector<std::string> parts;
    std::vector<std::array<int, 4>> faces;
    std::map<std::string, std::vector<int>> node_sets;
    std::map<std::string, std::vector<int>> face_sets;

    int add_part(const std::string& name);
    int part_index(const std::string& name) const;  // -1 when absent
    std::vector<int> part_hexes(int part) const;
    std::vector<int> part_nodes(int part) const;
    bool nodes_of(const std::string& name, std::vector<int>& out) const;
    bool face_set_nodes(const std::string& name, std::vector<int>& out) const;
    int add_hex_block(const std::string& name, Vec3 origin, Vec3 size, double element_size,
                      const std::vector<std::pair<BlockFace, std::string>>& surfaces);
    int add_hex_block_n(const std::string& name, Vec3 origin, Vec3 size, std::array<int, 3> n,
                        const std::vector<std::pair<BlockFace, std::string>>& surfaces);
    int merge_coincident_nodes(double tol);
    int nearest_node(Vec3 at, int part) const;
    double min_edge_at(int node) const;
    // (node, cosine), node -1 when none
    std::pair<int, double> best_aligned_neighbour(int node, Vec3 dir, double radius, int part,
                                                  const std::vector<int>& exclude) const;
};

// src/material.rs:41-71,154-237 (Material + Plasticity), flattened.
struct Material {
    crush_material m{};
    static Material elastic(double e, double nu, double rho);
    static Material honeycomb(double e, double rho, double sy, double h);
    static Material honeycomb_curve(double e, double rho, const std::vector<std::array<double, 2>>& curve,
                                    bool has_dens, std::array<double, 2> dens);
    Material scaled(double factor) const;
};

struct Contact {
    std::vector<int> nodes;
    std::vector<int> faces;
    std::vector<double> node_scale;
    double stiffness = 0.0;
    double max_distance = 0.0;
    int pair = 0;
};

struct Group {
    std::vector<int> nodes;
    std::vector<int> front;
    int accel_node = -1;
    int axis = 0;
    int frame_origin = -1, frame_x = -1;  // accelerometer frame (local x from origin to x)
    int frame_plane = -1;                 // the frame's x-y plane node
    int heading = 0;                      // +1 / -1 along global x, 0 unknown
};

// src/model.rs:152-165 (Model) plus the vehicle groups of the outputs.
struct Model {
    Mesh mesh;
    std::vector<Material> materials;
    std::vector<Vec3> initial_velocity;
    std::vector<int> fixed_nodes;
    std::vector<Contact> contacts;
    std::vector<Group> groups;
    crush_settings settings{};

    explicit Model(Mesh m);
    void set_material(const std::string& part, const Material& mat);
    void set_initial_velocity(const std::string& set, Vec3 v);
    void fix(const std::string& set);
    void add_contact_one_way(const std::string& nodes, const std::string& faces, double k, double max_d);
    void add_contact_pair_scaled(const std::string& a, const std::string& b, double k, double max_d,
                                 const std::vector<double>& sa, const std::vector<double>& sb);
};

// Reference Settings defaults (src/model.rs:134-149).
crush_settings default_settings();

// Owner of a model plus the flat view handed to the solver.
struct ModelHolder {
    Model model;
    std::vector<double> nodes, v0;
    std::vector<int32_t> hexes, hex_part, faces, fixed;
    std::vector<crush_material> materials;
    std::vector<crush_contact> contacts;
    std::vector<crush_group> groups;
    crush_model view{};
    explicit ModelHolder(Model m);
    void flatten();
};

// Vehicle (src/vehicle.rs:105-168) as the plain C struct plus the methods
// the builders need.
namespace vehicle {
crush_vehicle blank();
crush_vehicle preset(const std::string& name, bool* found);
crush_vehicle load_toml(const std::string& path);
crush_vehicle side_profile(const crush_vehicle& v);
bool rear_profile(const crush_vehicle& v, crush_vehicle& out);
bool has_crush_zone(const crush_vehicle& v);
double contact_stiffness_per_node(const crush_vehicle& v, int n_face, double h);
Material crush_material(const crush_vehicle& v);
}  // namespace vehicle

enum class Heading { PlusX, MinusX };

constexpr double GROUND_CLEARANCE = 0.15;    // src/vehicle.rs:499
constexpr double TUNED_ELEMENT_SIZE = 0.3;   // src/vehicle.rs:647
constexpr double PULSE_END = 0.15;           // src/vehicle.rs:1104
constexpr double MPH = 0.44704;              // src/lib.rs:28

int add_vehicle(Mesh& mesh, const crush_vehicle& v, double front_x, double y_center, double z_bottom,
                Heading heading, double element_size);
void assign_vehicle(Model& model, const crush_vehicle& v, Heading heading, double speed);
int add_ground(Mesh& mesh, const std::string& name, double z, std::array<double, 2> x,
               std::array<double, 2> y, double thickness);
void add_ground_contact(Model& model, const crush_vehicle& v, const std::string& ground, double clearance,
                        double element_size);
int add_rigid_wall(Mesh& mesh, const std::string& name, double x, std::array<double, 2> y,
                   std::array<double, 2> z, Heading facing, double thickness);
// add_vehicle_accelerometer's node (src/vehicle.rs:792-822, model.rs:185-187).
int vehicle_accel_node(const Model& model, const std::string& part, const std::string& front_face);
Group make_group(const Model& model, const std::string& name, const std::string& front_face, int axis,
                 int heading);

Model build_headon(const crush_vehicle& a, const crush_vehicle& b, double speed_a, double speed_b,
                   const crush_headon_options& opt);
Model build_crash(double truck_mph, double car_mph);
Model build_tbone(const crush_vehicle& striking, const crush_vehicle& struck, double speed_a, double speed_b,
                  double offset, double clearance, double element_size);
Model build_barrier(const crush_vehicle& v, double element_size, double speed, double end_time);

}  // namespace crush

struct crush_model_handle {
    crush::ModelHolder holder;
    explicit crush_model_handle(crush::Model m) : holder(std::move(m)) {}
};
