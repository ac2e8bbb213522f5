/*
 * crushrs.h — C ABI of the MI355X crash-solver core.
 *
 * The reference crate (osullivryan/crushrs) exposes one solver entry point,
 * `pub fn run(model: &Model) -> Results` (src/solver.rs:190, re-exported as
 * `crushrs::run` in src/lib.rs:26).  This header is the drop-in boundary for
 * that call: plain C structs of caller-owned arrays in, caller-owned arrays
 * out, integer status codes instead of panics.  Everything behind it (mesh
 * classes, HIP kernels, device memory) is private to libcrushrs_amd.so.
 *
 * Conventions
 *   - Node, element and face indices are 0-based int32.
 *   - Vectors are interleaved xyz: v[3*node + axis].
 *   - Every pointer in an input struct is borrowed for the duration of the
 *     call only; the library copies what it needs to the device.
 *   - Functions return CRUSH_OK (0) or a CRUSH_E_* code; the message of the
 *     last failure on the calling thread is crush_last_error().
 */
#ifndef CRUSHRS_H
#define CRUSHRS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CRUSH_ABI_VERSION 3

/* Status codes.  The reference panics in all of these cases
 * (src/solver.rs:302 non-finite displacement, src/model.rs:186-242 unknown
 * sets, src/material.rs:188,432 invalid curves). */
#define CRUSH_OK            0
#define CRUSH_E_INPUT       1  /* malformed model, settings or arguments       */
#define CRUSH_E_NONFINITE   2  /* displacement became non-finite               */
#define CRUSH_E_OOM         3  /* device or host allocation failed             */
#define CRUSH_E_DEVICE      4  /* HIP runtime error / no device                */
#define CRUSH_E_UNSUPPORTED 5  /* material model not on the GPU path           */
#define CRUSH_E_BUSY        6  /* server full (admission refused)              */
#define CRUSH_E_NOTFOUND    7  /* unknown vehicle / file / ticket              */

/* Material models: src/material.rs:20-36 (PlasticModel) plus plain elastic
 * (Material with plasticity = None, src/material.rs:165).  The GPU path runs
 * ELASTIC (finite-strain Hencky, src/material.rs:708-771, elastic branch) and
 * HONEYCOMB (corotational rate form, src/material.rs:472-535). */
#define CRUSH_MAT_ELASTIC   0
#define CRUSH_MAT_HONEYCOMB 1
#define CRUSH_MAT_J2        2
#define CRUSH_MAT_FOAM      3

#define CRUSH_MAX_KNOTS 8  /* src/material.rs:39 MAX_KNOTS */

/* One part's material: src/material.rs:41-71 (Plasticity) and :154-162. */
typedef struct crush_material {
    double youngs_modulus;
    double poisson_ratio;
    double density;
    int32_t model;            /* CRUSH_MAT_*                                   */
    int32_t n_curve;          /* 0 (bilinear) or 2..8 tabulated knots          */
    double yield_stress;      /* bilinear: initial yield stress (Pa)           */
    double hardening;         /* bilinear: hardening modulus (Pa)              */
    double curve[CRUSH_MAX_KNOTS][2]; /* [compaction, yield stress] knots      */
    int32_t has_densification;
    int32_t reserved0;
    double densification[2];  /* [c_lock, k_lock] lock-up                      */
    double transverse_factor; /* cap factor on local y, z normal stresses      */
} crush_material;

/* One-way node-to-face penalty contact: src/model.rs:12-28 (Contact).
 * add_contact_pair (src/model.rs:250-265) is two of these at half stiffness. */
typedef struct crush_contact {
    int32_t n_nodes;
    int32_t n_faces;
    const int32_t* nodes;      /* contacting nodes                             */
    const double* node_scale;  /* per-node stiffness factor, NULL = all 1      */
    const int32_t* faces;      /* indices into crush_model.faces               */
    double stiffness;          /* N/m per node                                 */
    double max_distance;       /* m                                            */
    int32_t pair;              /* 1: vehicle-to-vehicle (used by separation)   */
    int32_t reserved0;
} crush_contact;

/* src/model.rs:75-150 (Settings).  simd and threads select CPU code paths
 * of the reference and have no counterpart here. */
typedef struct crush_settings {
    double end_time;             /* s (reference default 0.15)                  */
    double dt_scale;             /* safety factor on the stable step (0.85)     */
    double time_step;            /* fixed step if > 0, else adaptive            */
    double hourglass;            /* stabilisation stiffness / E (0.1)           */
    double velocity_damping;     /* per-step velocity factor (1.0)              */
    double mass_scaling;         /* dt floor as fraction of initial step (0.25) */
    int32_t adaptive_check_steps;/* re-evaluate dt every N steps (5)            */
    /* Retire at separation (no counterpart in the reference, which always
     * runs to end_time): 0 = off.  When > 0 the solve stops after a check
     * step (every adaptive_check_steps) at which (a) pair contacts have had
     * no penetrating node for this many consecutive steps, (b) they had
     * one before, and (c) groups 0 and 1 move apart:
     * (v1 - v0) . (x1 - x0) > 0 with mass-weighted group means. */
    int32_t separation_steps;
    /* Accelerometer pulse (src/solver.rs:150-187 sampling, src/signal.rs:
     * 43-64 resampling): when > 0, the local-x acceleration of every
     * group's accelerometer node is sampled after each step and resampled
     * linearly onto t = k * pulse_dt as the solve runs (the reference's
     * history_steps = 1 with Pulse::from_history at PULSE_DT = 1e-4 s).
     * 0 = off. */
    double pulse_dt;
    /* Part and accelerometer histories (src/solver.rs:114-187, 276-282,
     * 349-357): a sample of the initial state, one after every
     * history_steps-th step and one after the last step.  0 = off. */
    int32_t history_steps;
    /* Deformed-mesh frames (src/solver.rs:265-277, 346-355): the initial
     * state, after every frame_steps-th step and after the last step.
     * crush_run only (batches and servers ignore it).  0 = off. */
    int32_t frame_steps;
} crush_settings;

/* A vehicle, for the delta-v / crush outputs of src/bin/crushrs.rs:237-242
 * (part_delta_v) and :265-271 (crush_rear). */
typedef struct crush_group {
    int32_t n_nodes;
    int32_t n_front;
    const int32_t* nodes;      /* node set <name> (all parts of the vehicle)   */
    const int32_t* front;      /* nodes of face set <name>_front               */
    int32_t accel_node;        /* rear-seat accelerometer node, -1 = none      */
    int32_t axis;              /* crush axis, 0 = x                            */
    /* accelerometer frame (src/model.rs:185-213): local x runs from node
     * frame_origin to node frame_x in the current configuration; -1 = the
     * global x axis */
    int32_t frame_origin;
    int32_t frame_x;
    /* direction of travel along global x (+1 / -1; 0 = unknown): the pulse
     * delta-v is reported in that direction (src/headon.rs:70-78) */
    int32_t heading;
    /* the frame's plane node (local x-y plane, +y side), -1 = none: local
     * y and z of the accelerometer history (src/model.rs:30-58) */
    int32_t frame_plane;
} crush_group;

/* src/model.rs:152-165 (Model) + src/mesh.rs:37-52 (Mesh), flattened. */
typedef struct crush_model {
    int32_t n_nodes;
    int32_t n_hexes;
    int32_t n_parts;
    int32_t n_faces;
    int32_t n_fixed;
    int32_t n_contacts;
    int32_t n_groups;
    int32_t reserved0;
    const double* nodes;             /* [3*n_nodes] reference positions       */
    const int32_t* hexes;            /* [8*n_hexes] C3D8 ordering              */
    const int32_t* hex_part;         /* [n_hexes]                              */
    const crush_material* materials; /* [n_parts]                              */
    const double* initial_velocity;  /* [3*n_nodes]                            */
    const int32_t* fixed_nodes;      /* [n_fixed]                              */
    const int32_t* faces;            /* [4*n_faces] quads, CCW about normal    */
    const crush_contact* contacts;   /* [n_contacts]                           */
    const crush_group* groups;       /* [n_groups]                             */
    crush_settings settings;
} crush_model;

typedef struct crush_group_result {
    double mass;           /* lumped mass of the node set                     */
    double delta_v[3];     /* final minus initial mass-weighted mean velocity */
    double crush_end;      /* |u_axis(accel) - mean u_axis(front)| at the end */
    double crush_max;      /* max of the same over all steps                  */
    /* Delta-v from the accelerometer pulse (src/headon.rs:70-78,
     * tests/impact.rs:305-306): the CFC 60 filtered local-x pulse integrated
     * over the record, negated for heading +1 (speed at the start minus
     * speed at the end along the direction of travel).  0 without a pulse. */
    double pulse_delta_v;
} crush_group_result;

/* src/solver.rs:63-80 (Results).  Output arrays are caller-owned; pass NULL
 * for any the caller does not want. */
typedef struct crush_results {
    double time;                 /* simulated time reached (s)                 */
    int64_t steps;
    int64_t eroded_elements;     /* elements eroded (inverted) during the run  */
    double wall_ms;              /* host wall time of the solve                */
    double device_ms;            /* device time of the step loop               */
    double added_mass_max;       /* largest total mass added by mass scaling   */
    int32_t separated;           /* 1 if retired at separation                 */
    int32_t n_groups;            /* entries written to groups                  */
    double* displacement;        /* [3*n_nodes] or NULL                        */
    double* velocity;            /* [3*n_nodes] or NULL                        */
    double* masses;              /* [n_nodes] lumped masses, or NULL           */
    crush_group_result* groups;  /* [n_groups] or NULL                         */
    /* accelerometer pulses, unfiltered local-x acceleration (m/s^2) at
     * t = k * pulse_dt: pulse[g * pulse_capacity + k] for k < pulse_samples;
     * NULL for none.  pulse_samples is set whenever pulse_dt > 0 (the count
     * may exceed pulse_capacity, then only the first pulse_capacity are
     * written). */
    double* pulse;
    int32_t pulse_capacity;
    int32_t pulse_samples;
    /* Histories (settings.history_steps > 0; src/solver.rs:19-62, 114-187).
     * Sample k < min(history_samples, history_capacity) was taken after
     * step history_step[k] at time history_time[k]; sample 0 is the initial
     * state.  history[(k * n_parts + p) * CRUSH_PART_FIELDS + f] is part
     * p's state (PartState): mass, mass-weighted mean displacement,
     * velocity and acceleration (xyz each) and kinetic energy.
     * accel_history[(k * n_accel + g) * CRUSH_ACCEL_FIELDS + f], n_accel =
     * min(n_groups, 8), is group g's accelerometer node (NodeHistory and
     * FrameHistory): position, displacement, velocity, acceleration (xyz
     * each, global axes), the frame origin's position and the frame's
     * local x, y, z unit axes in global components (the identity when the
     * group has no frame).  Any pointer may be NULL. */
    double* history_time;
    int64_t* history_step;
    double* history;
    double* accel_history;
    int32_t history_capacity;
    int32_t history_samples;
    /* Frames (settings.frame_steps > 0, crush_run only; Frame,
     * src/solver.rs:8-19): frame k < min(frame_samples, frame_capacity) at
     * frame_time[k]: nodal displacement [k][3 * n_nodes], and per element
     * in model order the equivalent plastic strain / compaction [k][n_hexes],
     * the Cauchy stress in Voigt order xx yy zz xy yz xz [k][6 * n_hexes]
     * and the eroded flag [k][n_hexes].  Any pointer may be NULL. */
    double* frame_time;
    float* frame_displacement;
    float* frame_plastic_strain;
    float* frame_stress;
    uint8_t* frame_eroded;
    int32_t frame_capacity;
    int32_t frame_samples;
} crush_results;

#define CRUSH_PART_FIELDS 11
#define CRUSH_ACCEL_FIELDS 24

/* ---- core entry points ------------------------------------------------- */

/* Library/ABI version (CRUSH_ABI_VERSION). */
int32_t crush_abi_version(void);

/* Message of the last failure on this thread ("" if none). */
const char* crush_last_error(void);

/* Number of visible HIP devices (0 when none); never fails. */
int32_t crush_device_count(void);

/* The explicit solve on the GPU, one model at a time.
 * Replaces crushrs::run (src/solver.rs:190-381, src/lib.rs:26). */
int32_t crush_run(const crush_model* model, crush_results* out);

/* Size of the static device scratch of crush_run_noalloc: 64 MiB per GPU,
 * held from library load until the process ends. */
#define CRUSH_NOALLOC_SCRATCH_BYTES (64u << 20)

/* crush_run with no device memory allocated: every device buffer of the
 * solve (the prepared model, nodal and element state, contact tables) is
 * carved from a static device array of CRUSH_NOALLOC_SCRATCH_BYTES, and
 * the solve runs on the null stream with no streams or events created.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid hipMalloc).
 * Results are those of crush_run.  A model whose buffers do not fit
 * (about 150k hexes) is refused with CRUSH_E_INPUT before any kernel runs;
 * frames (settings.frame_steps > 0) are refused with CRUSH_E_UNSUPPORTED.
 * Every call on one GPU uses that GPU's one array: calls from several host
 * threads on one GPU take turns (a lock in the library), and the array
 * keeps the last call's contents. */
int32_t crush_run_noalloc(const crush_model* model, crush_results* out);

/* Parity taps: run the first n_steps steps of the solve on the GPU and
 * record, per step s (0-based, after the step), f_int[s*3N..] (internal
 * nodal forces at the new configuration, src/solver.rs:327-328),
 * f_ext[s*3N..] (contact forces used by that step, src/solver.rs:307-310),
 * u[s*3N..] (displacement after the step),
 * hit[s*n_contact_nodes..] (face index hit by each contacting node of each
 * contact, -1 = none, contacts concatenated in order, nodes in contact
 * order after the static broad phase of src/contact.rs:48-70, i.e. only
 * nodes kept by it), eroded[s*n_hexes..] (1 = eroded) and dt[s].
 * Any output pointer may be NULL.  Returns the number of steps taken in
 * *steps_done (fewer than n_steps if end_time was reached). */
int32_t crush_debug_run_steps(const crush_model* model, int32_t n_steps,
                              double* f_int, double* f_ext, double* u, int32_t* hit,
                              uint8_t* eroded, double* dt, int32_t* steps_done);

/* Number of contacting nodes the static broad phase keeps, summed over the
 * contacts (the row length of crush_debug_run_steps' hit). */
int32_t crush_contact_node_count(const crush_model* model, int32_t* count);

/* Sizes a caller allocates for the histories and frames of a solve of this
 * model: the largest history_samples and frame_samples it can report
 * (each 0 when the settings ask for none).  No counterpart in the
 * reference, whose Results grow as the solve runs. */
int32_t crush_output_capacity(const crush_model* model, int32_t* history_samples, int32_t* frame_samples);


/* ---- batched solves ----------------------------------------------------
 * Many independent models solved together on one GPU: the counterpart of
 * the reference's calibration sweeps, which run solver::run for each model
 * in a rayon par_iter (src/vehicle.rs:1227-1233, src/headon.rs:179-182).
 * Each instance is solved exactly as crush_run would (same kernels' per-item
 * arithmetic); results do not depend on the other instances in the batch. */

typedef struct crush_batch crush_batch;

typedef struct crush_batch_options {
    int32_t threads_per_instance; /* workgroup of the per-instance step kernel: 64, 128, 256 or 512 (0: 256) */
    int32_t force_tile;           /* element slots per force workgroup: 64, 128 or 256 (0: 256) */
    int32_t steps_per_graph;      /* explicit steps captured in one hipGraph (0: 32) */
    int32_t host_threads;         /* host threads preparing models (0: all cores, max 16) */
    double contact_cell_scale;    /* contact grid cell / largest face extent (0: 1.0) */
    int32_t autotune;             /* 1: choose the three knobs above that do not change
                                     results by timed trials at create time */
    int32_t reserved0;
} crush_batch_options;

typedef struct crush_batch_stats {
    double prepare_ms;            /* host preparation of all models           */
    double upload_ms;             /* host->device copy of the instance images */
    double device_ms;             /* device time of crush_batch_run           */
    double wall_ms;               /* host wall time of crush_batch_run        */
    int64_t instances;
    int64_t element_steps;        /* sum over instances of hexes x steps      */
    int64_t launches;             /* explicit steps launched (last run)       */
    int64_t graph_replays;        /* hipGraph replays (last run)              */
    double pool_bytes;            /* device memory of the instance images     */
    double latency_p50_ms;        /* median time from run start to retirement */
    double latency_p99_ms;        /* 99th percentile of the same              */
    int64_t hc_element_steps;     /* honeycomb part of element_steps          */
    /* With kernel timing on (crush_batch_time_kernels), summed durations of
     * the kernels of the run, in ms: */
    double k_end_ms;              /* per-instance control kernel (step end,
                                     stable step)                            */
    double k_contact_ms;          /* contact: faces and grids, search, face
                                     reactions                               */
    double k_node_ms;             /* node pass: gather, contact force and next
                                     half step                               */
    double k_force_hc_ms;         /* honeycomb element forces                 */
    double k_force_el_ms;         /* elastic element forces                   */
    double k_ms[12];              /* each kernel of a tick in launch order:
                                     control, faces + grid, search,
                                     reactions, node pass, honeycomb,
                                     elastic (the rest 0); a launch that
                                     runs the control pass with the face
                                     tiles or the grid pass counts as
                                     control                                 */
} crush_batch_stats;

/* Fill opt with the defaults. */
void crush_batch_defaults(crush_batch_options* opt);

/* Prepare n models (host threads) and upload them to the device. */
int32_t crush_batch_create(const crush_model* const* models, int32_t n, const crush_batch_options* opt,
                           crush_batch** out);

/* Solve every instance to its end (end_time or separation). */
int32_t crush_batch_run(crush_batch* b, crush_batch_stats* stats);

/* Options in effect (after autotuning). */
int32_t crush_batch_options_get(const crush_batch* b, crush_batch_options* opt);

/* Contact grid cell scales the autotuner chose per size class (no
 * counterpart in the reference, which fixes its own grouping:
 * src/kernel_simd.rs:562).  The classes hold equal counts of instances by
 * element count, ascending; up to cap of them are written to scale and
 * max_elems (the largest element count in the class), and their number to
 * n_classes (0 when the batch was not autotuned). */
int32_t crush_batch_class_scales(const crush_batch* b, int32_t cap, double* scale, int32_t* max_elems,
                                 int32_t* n_classes);

/* Kernel timing on (1) or off (0): with it on, crush_batch_run launches
 * the steps directly instead of replaying the hipGraph, records HIP events
 * around each kernel and reports the summed kernel durations in its stats. */
int32_t crush_batch_time_kernels(crush_batch* b, int32_t on);

/* Restore every instance to its initial state (so the batch can be run
 * again, e.g. for timing). */
int32_t crush_batch_reset(crush_batch* b);

/* Results of instance i (fields as crush_run; the displacement and
 * velocity arrays are copied only when the pointers are non-NULL).
 * device_ms is the time from the start of the last crush_batch_run to the
 * retirement of this instance; wall_ms is 0. */
int32_t crush_batch_result(crush_batch* b, int32_t i, crush_results* out);

void crush_batch_free(crush_batch* b);

/* ---- serving ------------------------------------------------------------
 * A request loop for delta-v queries: models are submitted one at a time
 * from any thread, prepared on the submitting thread, admitted to the GPU
 * as slots and device memory free up (admission by an HBM budget), stepped
 * together with every other resident request (continuous batching: new
 * requests join between graph replays of a few steps) and retired at their
 * end time or at separation (crush_settings.separation_steps).  Each
 * result is the one crush_run gives for that model alone.  There is no
 * counterpart in the reference, which solves one model per call. */

typedef struct crush_server crush_server;

typedef struct crush_server_options {
    double hbm_budget_bytes;      /* device memory for resident requests (0: 16 GiB)   */
    int32_t max_resident;         /* requests on the device at once (0: 4096)          */
    int32_t steps_per_graph;      /* steps between admissions (0: 16)                  */
    int32_t threads_per_instance; /* as crush_batch_options (0: 256)                   */
    int32_t force_tile;           /* as crush_batch_options (0: 256)                   */
    double contact_cell_scale;    /* as crush_batch_options (0: 1.0)                   */
    int32_t max_queued;           /* submissions waiting for admission (0: 1 << 20)    */
    int32_t reserved0;
} crush_server_options;

#define CRUSH_SUBMIT_FIELDS 1     /* keep the final u and v for crush_server_poll      */

typedef struct crush_server_stats {
    int64_t submitted, admitted, completed, failed;
    int32_t resident, queued;
    double device_bytes_in_use;
    double latency_p50_ms;        /* submit to result, over completed requests          */
    double latency_p99_ms;
    double queue_p50_ms;          /* submit to admission                                */
    double busy_ms;               /* driver time with requests resident                 */
    int64_t element_steps;        /* sum over completed requests                         */
    int64_t replays;
    int64_t captures;             /* step graphs captured (lists outgrew the grids)      */
    double admit_ms;              /* driver time uploading admitted requests             */
    double lists_ms;              /* driver time rebuilding lists and capturing graphs   */
    double replay_ms;             /* driver time in replays, through the done-flag read  */
    double retire_ms;             /* driver time reading the results of retired requests */
} crush_server_stats;

void crush_server_defaults(crush_server_options* opt);
int32_t crush_server_create(const crush_server_options* opt, crush_server** out);
/* Queue a model; the model may be freed when this returns.  flags:
 * CRUSH_SUBMIT_FIELDS.  CRUSH_E_BUSY when the queue is full or the model
 * can never fit the HBM budget. */
int32_t crush_server_submit(crush_server* s, const crush_model* model, int32_t flags, int64_t* ticket);
/* Non-blocking: *done = 1 and out filled (then the ticket is released) when
 * the request has finished, else *done = 0.  The displacement / velocity
 * arrays are filled only for CRUSH_SUBMIT_FIELDS requests; out->wall_ms is
 * the submit-to-result latency and out->device_ms the admit-to-result time. */
int32_t crush_server_poll(crush_server* s, int64_t ticket, crush_results* out, int32_t* done);
/* Blocking form of crush_server_poll (timeout_ms <= 0: no limit). */
int32_t crush_server_wait(crush_server* s, int64_t ticket, crush_results* out, double timeout_ms);
int32_t crush_server_stats_get(crush_server* s, crush_server_stats* st);
void crush_server_destroy(crush_server* s);

#ifdef __cplusplus
}
#endif

#endif /* CRUSHRS_H */
