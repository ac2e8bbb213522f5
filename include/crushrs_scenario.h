/*
 * crushrs_scenario.h — host-side scenario builders of the crash-solver core.
 *
 * These produce the crush_model the solver entry points consume, the way
 * the reference's builders produce its `Model`:
 *   Vehicle presets / TOML       src/vehicle.rs:105-226, :523-647
 *   add_vehicle / assign_vehicle src/vehicle.rs:663-787
 *   road, wall, ground contact   src/vehicle.rs:829-859
 *   barrier_model                src/vehicle.rs:951-973
 *   HeadOn::build                src/headon.rs:40-66
 *   `crushrs crash`              src/bin/crushrs.rs:403-432
 *   `crushrs tbone`              src/bin/crushrs.rs:550-573
 *   `crushrs rearend`            src/bin/crushrs.rs:522-545
 * A model handle owns every array its crush_model view points into.
 */
#ifndef CRUSHRS_SCENARIO_H
#define CRUSHRS_SCENARIO_H

#include "crushrs.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CRUSH_MAX_TABLE 8

/* A homogenised vehicle: src/vehicle.rs:105-168 (Vehicle).  Optional
 * fields carry a has_* flag (Rust Option). */
typedef struct crush_vehicle {
    char name[64];
    double mass;                 /* kg                                         */
    double size[3];              /* length (x), width (y), height (z), m       */
    double crush_zone_length;    /* m                                          */
    double curve[2];             /* bilinear [yield_force N, stiffness N/m]    */
    double modulus;              /* crush zone Young's modulus, Pa             */
    double body_modulus;         /* body Young's modulus, Pa                   */
    int32_t has_side_curve;
    int32_t has_side_modulus;
    double side_curve[2];
    double side_modulus;
    int32_t has_rear_curve;
    int32_t has_rear_modulus;
    double rear_curve[2];
    double rear_modulus;
    int32_t n_force_table;       /* 0 = none, else knots of [crush m, force N] */
    int32_t n_compaction_map;    /* 0 = none                                   */
    double force_table[CRUSH_MAX_TABLE][2];
    double compaction_map[CRUSH_MAX_TABLE];
    int32_t has_crush_element_size;
    int32_t has_bumper;
    double crush_element_size;
    double bumper[3];            /* [thickness m, mass kg, modulus Pa]        */
    double transverse_factor;
    int32_t has_crush_zone_mass;
    int32_t has_rail_box;
    double crush_zone_mass;
    double rail_box_width;
    double rail_box_z[2];
    double rail_box_force_fraction;
} crush_vehicle;

typedef struct crush_model_handle crush_model_handle;

/* Options of HeadOn (src/headon.rs:16-35); defaults via
 * crush_headon_defaults: offset 0, road 0.15 m below, 0.3 m elements,
 * end_time 0.15 s. */
typedef struct crush_headon_options {
    double offset;        /* lateral offset of B (m)                         */
    double clearance;     /* road below the underbodies (m); < 0 = no road   */
    double element_size;  /* transverse element size (m)                    */
    double end_time;      /* s                                               */
} crush_headon_options;

void crush_headon_defaults(crush_headon_options* opt);

/* A built-in preset by name (silverado, silverado-raw, neon, neon-raw,
 * neon-pulse, expedition, navigator; src/bin/crushrs.rs:246-260) or a
 * vehicle TOML file (Vehicle::load, src/vehicle.rs:212-219). */
int32_t crush_vehicle_load(const char* name_or_path, crush_vehicle* out);

/* Vehicle::side_profile (src/vehicle.rs:271-284). */
int32_t crush_vehicle_side_profile(const crush_vehicle* v, crush_vehicle* out);

/* Vehicle::rear_profile (src/vehicle.rs:297-316); CRUSH_E_NOTFOUND when the
 * vehicle has no rear curve (the reference returns None). */
int32_t crush_vehicle_rear_profile(const crush_vehicle* v, crush_vehicle* out);

/* HeadOn::build: A heads +x, B heads -x (src/headon.rs:40-66).  Speeds in
 * m/s; a negative speed_b makes B head +x (rear-end use).  Groups 0 and 1
 * are the two vehicles. */
int32_t crush_scenario_headon(const crush_vehicle* a, const crush_vehicle* b, double speed_a,
                              double speed_b, const crush_headon_options* opt,
                              crush_model_handle** out);

/* `crushrs crash truck_mph car_mph`: tuned Silverado vs tuned Neon, no
 * road (src/bin/crushrs.rs:403-432). */
int32_t crush_scenario_crash(double truck_mph, double car_mph, crush_model_handle** out);

/* `crushrs tbone`: striking vehicle's front into the struck vehicle's side
 * profile moving at speed_b along +y (src/bin/crushrs.rs:550-573).
 * clearance < 0 = no road. */
int32_t crush_scenario_tbone(const crush_vehicle* striking, const crush_vehicle* struck,
                             double speed_a, double speed_b, double offset, double clearance,
                             double element_size, crush_model_handle** out);

/* barrier_model: rigid wall at x = 0 (src/vehicle.rs:951-973). */
int32_t crush_scenario_barrier(const crush_vehicle* v, double element_size, double speed,
                               double end_time, crush_model_handle** out);

/* The flattened model a handle owns (valid until crush_model_free). */
const crush_model* crush_model_view(const crush_model_handle* h);

/* Replace the settings of a handle's model. */
int32_t crush_model_set_settings(crush_model_handle* h, const crush_settings* s);

/* Nodes of a named node set / part, or of a face set when face_set != 0
 * (Mesh::nodes_of, Mesh::face_set_nodes: src/mesh.rs:81-100). */
int32_t crush_model_node_set(const crush_model_handle* h, const char* name, int32_t face_set,
                             const int32_t** nodes, int32_t* n);

/* Part names in part order, '\n'-separated, into buf (truncated). */
int32_t crush_model_part_names(const crush_model_handle* h, char* buf, int32_t cap);

void crush_model_free(crush_model_handle* h);

#ifdef __cplusplus
}
#endif
#endif
