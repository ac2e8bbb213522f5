// Shared state of the batched step engine (kernels in batch.hip): instance
// images in one device pool, the active lists, the captured step graph.
// Used by crush_batch_* (a fixed set of instances) and crush_server_*
// (slots that are filled and freed as requests come and go).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "batch.hpp"
#include "hip_util.hpp"
#include "phases.hpp"

namespace crush {

// Smallest face-grid region of a contact (cells, entries); larger contacts
// get 8 cells and 64 entries per face.
constexpr int BATCH_GRID_CELLS = 1024;
constexpr int BATCH_GRID_ENTRIES = 4096;
// Cells of a grid built in LDS (k_grid workgroups of 512 or more threads)
constexpr int LDS_GRID_CELLS = 4096;
// Below this many search tiles a batch searches with 4 lanes per node, and
// below SEARCH_WIDE_TILES with 16 (a tile's 64 nodes over 4 workgroups)
constexpr int SEARCH_SPLIT_TILES = 2048;
#ifndef CRUSH_SEARCH_WIDE
#define CRUSH_SEARCH_WIDE 256
#endif
constexpr int SEARCH_WIDE_TILES = CRUSH_SEARCH_WIDE;
constexpr int NODE_TILE = 256;
constexpr int CONTACT_TILE = 64;
// Contact grids are kept over steps: faces are binned grown by this
// fraction of the contact's largest face extent, and the grid is rebuilt
// when a face leaves that growth.
constexpr double GRID_MARGIN = 0.1;

// Offsets of every array of an instance image relative to its start; the
// first `init_bytes` must be uploaded, the rest is scratch.
struct ImageOffsets {
    size_t conn, ecls, cls_mat, cpk, cdtf, emass, mats, state, eroded, extra;
    size_t X, u, v, f_int, f_ext, f_ext2, inv_mass, mass, added, fixed, probe_node, g_probe, n2e_ptr, n2e_idx;
    size_t cdesc, cn_node, cn_scale, cf_nodes, cf_contact, cf_static, cf_level, tn_node, tn_ptr, tn_ent, tn_of, rec_face, fh_count;
    size_t g_nodes_ptr, g_nodes, g_front_ptr, g_front, g_accel, g_axis, g_mass, g_v0, g_dv, g_frame, ctl;
    size_t p_nodes_ptr, p_nodes;
    size_t init_bytes;
    size_t fe, fgeo, fbox, rec_f, fh_list, face_force, cgrid, cgp, dtp_min, dtp_sum, pulse, hist, probe;
    size_t total;
};

// Active lists: the instance list and the tile lists of the tile kernels.
enum { L_INST, L_HC, L_EL, L_NODE, L_FACE, L_GRID, L_SEARCH, L_FGEO, N_LISTS };
// Kernels of one tick, in launch order.
enum { K_END, K_GRID, K_SEARCH, K_FACEFORCE, K_NODE, K_FORCE_HC, K_FORCE_EL, N_KERNELS };

struct BatchImpl {
    crush_batch_options opt{};
    std::vector<Prepared> prep;
    std::vector<ImageOffsets> offs;
    std::vector<size_t> base_off;
    DevBuf<uint8_t> pool;
    DevBuf<InstDev> table;
    std::vector<InstDev> host_tab;  // host copy of the table
    // autotuned contact grid cell scale per size class (classes of equal
    // count by element count; class_elems = the largest in each)
    std::vector<double> class_scale;
    std::vector<int> class_elems;
    // active lists: the instances (d_list) and the tiles of each tile
    // kernel (d_tiles[L_*]); their counts are device ints (d_counts[L_*])
    // read by the kernels
    DevBuf<int32_t> d_list, d_counts, n_done;
    DevBuf<int2> d_tiles[N_LISTS];
    DevBuf<uint8_t> done_flag;
    std::vector<uint8_t> host_done;
    int cnt[N_LISTS] = {};  // entries in the current lists
    int cap[N_LISTS] = {};  // grids of the captured graph
    int chunk[N_LISTS] = {1, 1, 1, 1, 1, 1, 1, 1};  // tiles per XCD chunk (force and node lists)
    double headroom = 1.0;  // captured grids / current lists (server: > 1)
    bool need_capture = false;
    int grid_block = 64;    // k_grid workgroup: the largest moving contact's faces, in waves
    int search_lanes = 1;   // k_search_tiles lanes per contacting node (1, 4 or 16)
    bool fresh = true;  // no step taken since creation or reset
    DevBuf<BatchSummary> summary;
    std::vector<BatchSummary> host_summary;
    bool summary_valid = false;
    DevBuf<unsigned long long> t_start;
    unsigned long long host_t_start = 0;
    double clock_khz = 1.0;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    double prepare_ms = 0, upload_ms = 0;
    long long launches = 0, replays = 0;
    // kernel timing: with it on, each replay launches its steps directly,
    // with an event before each kernel of a step and one after the last
    bool time_kernels = false;
    std::vector<hipEvent_t> ev;
    double t_kernel[N_KERNELS] = {};

    ~BatchImpl() {
        if (graph) (void)hipGraphExecDestroy(graph);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Image of one prepared model (uploaded part in `out`) and its layout.
ImageOffsets build_image(const Prepared& P, std::vector<uint8_t>& out);
// Device pointers of an image placed at `base`.
InstDev make_inst(const Prepared& P, const ImageOffsets& o, uintptr_t base);
// Rebuild the active lists from the slots with host_done == 0; grows the
// list buffers (and marks a capture due) when needed.
void rebuild_lists(BatchImpl& bi);
// Capture steps_per_graph steps with grids of headroom x the current lists.
void capture(BatchImpl& bi);
// One replay of steps_per_graph steps (stable-step partials of fresh
// instances first).
void replay(BatchImpl& bi);
// Stable-step partials of every listed instance that has not stepped yet.
void dt_init(BatchImpl& bi);
// Summaries of slots [0, n) into bi.host_summary.
void collect(BatchImpl& bi, int n);
int steps_per_graph(const crush_batch_options& o);
// Contact grid cell scale of the options (1 when unset).
double cell_scale_of(const crush_batch_options& o);

}  // namespace crush
