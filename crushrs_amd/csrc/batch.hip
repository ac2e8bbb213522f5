// Batched solves: thousands of independent models on one GPU.
//
// Each instance is one contiguous image in a device pool (its arrays plus a
// control block) described by an InstDev pointer table.  One tick of every
// active instance is seven launches (eight with a large contact), each over
// all instances; a tick gathers one step and runs the contact, half step
// and forces of the next:
//   k_ctl              one workgroup per instance: end of the gathered step
//                      and the stable step of the next (at check steps);
//   k_facegeo_tiles    face geometry of contacts of more than 256 moving
//                      faces, in tiles of faces;
//   k_grid             one workgroup per contact of an instance: face
//                      geometry of the other contacts and the face grid,
//                      stored to device memory;
//   k_search_tiles     node search against the grid (tiles of contacting
//                      nodes); k_faceforce_tiles face reactions (tiles of
//                      faces);
//   k_node_tiles       node-centric gather and second half step of the
//                      step in flight, fused with the contact force and
//                      first half step of the next (tiles of 256 nodes);
//   k_force_hc,        element forces and the next check step's stable-step
//   k_force_el         partials (tiles of element slots), honeycomb and
//                      elastic tiles in separate launches so each fills the
//                      GPU at its own register budget.
// A step whose end reads the full-step velocity (the last step, history
// samples, separation checks) is gathered in a tick of its own.
// No arithmetic of an instance depends on the others, so its results are
// bitwise the same alone or in any mix.  A fixed number of steps is
// captured in one hipGraph and replayed until every instance has retired;
// retired instances leave the active lists between replays.
#include "engine.hpp"
#include "outputs.hpp"
#include "signal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

#include "hip_util.hpp"
#include "phases.hpp"

namespace crush {


// ------------------------------------------------------------- images

namespace {

struct ImageBuilder {
    std::vector<uint8_t> bytes;
    size_t put_raw(const void* p, size_t n) {
        size_t off = (bytes.size() + 255) & ~size_t(255);
        bytes.resize(off + std::max<size_t>(n, 1), 0);
        if (n && p) std::memcpy(bytes.data() + off, p, n);
        return off;
    }
    template <class T>
    size_t put(const std::vector<T>& v) {
        return put_raw(v.data(), v.size() * sizeof(T));
    }
    size_t zeros(size_t n) { return put_raw(nullptr, n); }
};

template <class T>
T* at(uintptr_t base, size_t off) {
    return reinterpret_cast<T*>(base + off);
}

}  // namespace

static StepCtl initial_ctl(const Prepared& P) {
    StepCtl c;
    std::memset(&c, 0, sizeof c);
    c.dt = P.dt0;
    c.dt_cfl = P.dt0;
    c.end_time = P.settings.end_time;
    c.contact_dt = P.contact_dt;
    c.dt_floor = P.dt_floor;
    c.dt_scale = P.settings.dt_scale;
    c.damping = P.settings.velocity_damping;
    c.chk = P.settings.adaptive_check_steps > 0 ? P.settings.adaptive_check_steps : 1;
    c.sep_steps = P.settings.separation_steps;
    c.adaptive = P.adaptive ? 1 : 0;
    return c;
}

// Probe nodes of an instance, ascending: each accelerometer node of groups
// < 8 and its frame's origin and x node.
static std::vector<int32_t> probe_nodes(const Prepared& P) {
    std::vector<int32_t> v;
    for (int g = 0; g < std::min(P.n_groups, 8); ++g)
        for (int n : {P.g_accel[g], P.g_frame[3 * g], P.g_frame[3 * g + 1]})
            if (n >= 0) v.push_back(n);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    return v;
}

// Stable-step partial slots of one instance (enough for the smallest tile).
// Stable-step partials of an instance: one per 64 element slots,
// honeycomb slots first, then elastic ones (dt_slots of them).
__host__ __device__ __forceinline__ int dt_slots(int n_hc, int E) { return (n_hc + 63) / 64 + (E - n_hc + 63) / 64; }
static size_t dtp_slots(const Prepared& P) { return (size_t)dt_slots(P.n_hc, P.n_elems) + 2; }

// Host image of one instance (initialized part only).
ImageOffsets build_image(const Prepared& P, std::vector<uint8_t>& out) {
    ImageBuilder b;
    ImageOffsets o{};
    const int N = P.n_nodes, E = P.n_elems;
    o.conn = b.put(P.conn);
    o.ecls = b.put(P.ecls);
    o.cls_mat = b.put(P.cls_mat);
    o.cpk = b.put(P.cls_pack);
    o.cdtf = b.put(P.cls_dtf);
    o.emass = b.put(P.elem_mass);
    o.mats = b.put(P.mats);
    {
        std::vector<float> st((size_t)25 * std::max(P.n_hc, 1), 0.0f);
        for (int k : {9, 13, 17})
            for (int i = 0; i < P.n_hc; ++i) st[(size_t)k * P.n_hc + i] = 1.0f;
        o.state = b.put(st);
    }
    o.eroded = b.zeros((size_t)E);
    o.extra = b.zeros(sizeof(double) * E);
    o.X = b.put(P.X);
    o.u = b.zeros(sizeof(double) * 3 * N);
    o.v = b.put(P.v0);
    o.f_int = b.zeros(sizeof(double) * 3 * N);
    o.f_ext = b.zeros(sizeof(double) * 3 * N);
    o.f_ext2 = b.zeros(sizeof(double) * 3 * N);
    {
        std::vector<double> im(N);
        for (int i = 0; i < N; ++i) im[i] = P.mass[i] > 0.0 ? 1.0 / P.mass[i] : 0.0;
        o.inv_mass = b.put(im);
    }
    o.mass = b.put(P.mass);
    o.added = b.zeros(sizeof(double) * N);
    const std::vector<int32_t> probes = probe_nodes(P);
    {
        std::vector<uint8_t> fl = P.fixed;
        for (int32_t n : probes) fl[n] |= NODE_PROBE;
        o.fixed = b.put(fl);
    }
    o.probe_node = b.put(probes);
    {
        auto slot = [&](int n) {
            return n < 0 ? -1 : (int32_t)(std::lower_bound(probes.begin(), probes.end(), n) - probes.begin());
        };
        std::vector<int32_t> gp;
        for (int g = 0; g < std::min(P.n_groups, 8); ++g) {
            const int na = P.g_accel[g];
            gp.push_back(na >= 0 && !(P.fixed[na] & NODE_FIXED) ? slot(na) : -1);
            gp.push_back(slot(P.g_frame[3 * g]));
            gp.push_back(slot(P.g_frame[3 * g + 1]));
        }
        o.g_probe = b.put(gp);
    }
    o.n2e_ptr = b.put(P.n2e_ptr);
    o.n2e_idx = b.put(P.n2e_idx);

    // face-grid region of each contact: [starts: cells + 1][entries], plus a
    // cursor array of `cells` for grids built in device memory (contacts
    // with more faces than the LDS grid is sized for)
    std::vector<ContactDesc> cdesc = P.cdesc;
    size_t grid_ints = 0;
    for (ContactDesc& cd : cdesc) {
        cd.grid_cells = std::min(std::max(8 * cd.n_faces, BATCH_GRID_CELLS), 1 << 16);
        cd.grid_entries = std::min(std::max(64 * cd.n_faces, BATCH_GRID_ENTRIES), 1 << 21);
        cd.grid_off = (int32_t)grid_ints;
        grid_ints += (size_t)cd.grid_cells + 1 + cd.grid_entries + cd.grid_cells;
    }
    o.cdesc = b.put(cdesc);
    o.cn_node = b.put(P.cn_node);
    o.cn_scale = b.put(P.cn_scale);
    o.cf_nodes = b.put(P.cf_nodes);
    o.cf_contact = b.put(P.cf_contact);
    o.cf_static = b.put(P.cf_static);
    o.cf_level = b.zeros((size_t)P.n_cfaces);
    o.tn_node = b.put(P.tn_node);
    o.tn_ptr = b.put(P.tn_ptr);
    o.tn_ent = b.put(P.tn_ent);
    o.tn_of = b.put(P.tn_of);
    o.rec_face = b.put(std::vector<int32_t>(std::max(P.n_cnodes, 1), -1));
    o.fh_count = b.zeros(sizeof(int32_t) * std::max(P.n_cfaces, 1));
    o.g_nodes_ptr = b.put(P.g_nodes_ptr);
    o.g_nodes = b.put(P.g_nodes);
    o.g_front_ptr = b.put(P.g_front_ptr);
    o.g_front = b.put(P.g_front);
    o.g_accel = b.put(P.g_accel);
    o.g_axis = b.put(P.g_axis);
    o.g_mass = b.put(P.g_mass);
    o.g_v0 = b.put(P.g_v0);
    o.g_dv = b.zeros(sizeof(double) * 3 * std::max(P.n_groups, 1));
    o.g_frame = b.put(P.g_frame);
    o.p_nodes_ptr = b.put(P.p_nodes_ptr);
    o.p_nodes = b.put(P.p_nodes);
    o.cgp = b.zeros(sizeof(FaceGrid) * std::max<size_t>(P.cdesc.size(), 1));
    {
        StepCtl c = initial_ctl(P);
        o.ctl = b.put_raw(&c, sizeof c);
    }
    o.init_bytes = (b.bytes.size() + 255) & ~size_t(255);
    b.bytes.resize(o.init_bytes, 0);
    // scratch (never uploaded)
    size_t off = o.init_bytes;
    auto scratch = [&](size_t n) {
        size_t r = off;
        off = (off + std::max<size_t>(n, 1) + 255) & ~size_t(255);
        return r;
    };
    o.fe = scratch(sizeof(float) * fe_floats(E));
    o.fgeo = scratch(sizeof(double) * FG * (size_t)P.n_cfaces);
    o.fbox = scratch(sizeof(float4) * 2 * (size_t)P.n_cfaces);
    o.rec_f = scratch(sizeof(double) * 3 * (size_t)P.n_cnodes);
    o.fh_list = scratch(sizeof(int32_t) * FACE_HITS_CAP * (size_t)P.n_cfaces);
    o.face_force = scratch(sizeof(double) * 3 * (size_t)P.n_cfaces);
    o.cgrid = scratch(sizeof(int32_t) * grid_ints);
    o.dtp_min = scratch(sizeof(double) * dtp_slots(P));
    o.dtp_sum = scratch(sizeof(double) * dtp_slots(P));
    // pulses: filled in order from sample 0 as the steps end, never read
    // before they are written
    o.pulse = scratch(sizeof(double) * (size_t)std::min(P.n_groups, 8) * (size_t)P.n_pulse);
    // history records, written in order as the samples are taken
    o.hist = scratch(sizeof(double) * hist_record_len(P.n_parts, P.n_groups) * (size_t)P.n_hist);
    o.probe = scratch(sizeof(double) * 6 * probes.size());
    o.total = off;
    out.swap(b.bytes);
    return o;
}

InstDev make_inst(const Prepared& P, const ImageOffsets& o, uintptr_t base) {
    InstDev I{};
    I.N = P.n_nodes;
    I.E = P.n_elems;
    I.n_hc = P.n_hc;
    I.C = P.n_classes;
    I.F = P.n_cfaces;
    I.K = P.n_cnodes;
    I.T = (int)P.tn_node.size();
    I.n_contacts = (int)P.cdesc.size();
    I.n_groups = P.n_groups;
    I.n_sblk = 0;
    I.conn = at<int32_t>(base, o.conn);
    I.ecls = at<int32_t>(base, o.ecls);
    I.cls_mat = at<int32_t>(base, o.cls_mat);
    I.cpk = at<float>(base, o.cpk);
    I.cdtf = at<double>(base, o.cdtf);
    I.emass = at<double>(base, o.emass);
    I.mats = at<MatParams>(base, o.mats);
    I.state = at<float>(base, o.state);
    I.eroded = at<uint8_t>(base, o.eroded);
    I.fe = at<float>(base, o.fe);
    I.extra = at<double>(base, o.extra);
    I.X = at<double>(base, o.X);
    I.u = at<double>(base, o.u);
    I.v = at<double>(base, o.v);
    I.f_int = at<double>(base, o.f_int);
    I.f_ext = at<double>(base, o.f_ext);
    I.f_ext2 = at<double>(base, o.f_ext2);
    I.inv_mass = at<double>(base, o.inv_mass);
    I.mass = at<double>(base, o.mass);
    I.added = at<double>(base, o.added);
    I.fixed = at<uint8_t>(base, o.fixed);
    I.n2e_ptr = at<int32_t>(base, o.n2e_ptr);
    I.n2e_idx = at<int32_t>(base, o.n2e_idx);
    I.cdesc = at<ContactDesc>(base, o.cdesc);
    I.cn_node = at<int32_t>(base, o.cn_node);
    I.cn_scale = at<double>(base, o.cn_scale);
    I.cf_nodes = at<int32_t>(base, o.cf_nodes);
    I.cf_contact = at<int32_t>(base, o.cf_contact);
    I.cf_static = at<uint8_t>(base, o.cf_static);
    I.cf_level = at<uint8_t>(base, o.cf_level);
    I.tn_node = at<int32_t>(base, o.tn_node);
    I.tn_ptr = at<int32_t>(base, o.tn_ptr);
    I.tn_ent = at<int32_t>(base, o.tn_ent);
    I.tn_of = at<int32_t>(base, o.tn_of);
    I.sblk_contact = nullptr;
    I.sblk_start = nullptr;
    I.fgeo = at<double>(base, o.fgeo);
    I.fbox = at<float4>(base, o.fbox);
    I.rec_face = at<int32_t>(base, o.rec_face);
    I.rec_f = at<double>(base, o.rec_f);
    I.fh_count = at<int32_t>(base, o.fh_count);
    I.fh_list = at<int32_t>(base, o.fh_list);
    I.face_force = at<double>(base, o.face_force);
    I.g_nodes_ptr = at<int32_t>(base, o.g_nodes_ptr);
    I.g_nodes = at<int32_t>(base, o.g_nodes);
    I.g_front_ptr = at<int32_t>(base, o.g_front_ptr);
    I.g_front = at<int32_t>(base, o.g_front);
    I.g_accel = at<int32_t>(base, o.g_accel);
    I.g_axis = at<int32_t>(base, o.g_axis);
    I.g_mass = at<double>(base, o.g_mass);
    I.g_v0 = at<double>(base, o.g_v0);
    I.g_dv = at<double>(base, o.g_dv);
    I.g_frame = at<int32_t>(base, o.g_frame);
    I.pulse = at<double>(base, o.pulse);
    I.n_pulse = P.n_pulse;
    I.pulse_dt = P.settings.pulse_dt;
    I.p_nodes_ptr = at<int32_t>(base, o.p_nodes_ptr);
    I.p_nodes = at<int32_t>(base, o.p_nodes);
    I.hist = at<double>(base, o.hist);
    I.n_parts = P.n_parts;
    I.n_hist = P.n_hist;
    I.hist_steps = std::max(P.settings.history_steps, 1);
    I.ctl = at<StepCtl>(base, o.ctl);
    I.cgrid = at<int32_t>(base, o.cgrid);
    I.cgp = at<FaceGrid>(base, o.cgp);
    I.dtp_min = at<double>(base, o.dtp_min);
    I.dtp_sum = at<double>(base, o.dtp_sum);
    I.probe_node = at<int32_t>(base, o.probe_node);
    I.probe = at<double>(base, o.probe);
    I.n_probe = (int)probe_nodes(P).size();
    I.g_probe = at<int32_t>(base, o.g_probe);
    I.cell_scale = 1.0;
    return I;
}

// ------------------------------------------------------------- kernels

// Outputs of a retired instance: per-group delta-v and end crush.
template <int B>
__device__ void finish_instance(const InstDev& D, StepCtl* c, double* s_w) {
    const int t = threadIdx.x;
    if (t == 0) c->t_retire = wall_clock64();
    for (int g = 0; g < D.n_groups; ++g) {
        double m[7];
        group_means<B>(D, g, s_w, m);
        if (t == 0)
            for (int d = 0; d < 3; ++d) D.g_dv[3 * g + d] = m[0] > 0.0 ? m[4 + d] / m[0] - D.g_v0[3 * g + d] : 0.0;
        if (g < 8) {
            const double cr = group_crush<B>(D, g, s_w);
            if (t == 0) {
                c->crush_end[g] = cr;
                c->crush_max[g] = fmax(c->crush_max[g], cr);
            }
        }
    }
}

// The partial of the 64 slots of its kind from `first` (first minus the
// kind's first slot is a multiple of 64).
template <int KIND>
__device__ __forceinline__ int dt_slot(const InstDev& D, int first) {
    return KIND == CRUSH_MAT_HONEYCOMB ? first / 64 : (D.n_hc + 63) / 64 + (first - D.n_hc) / 64;
}

// Stable-step partial of one wave's 64 slots [first, first + 64) of a kind
// that ends at `end` (src/kernel.rs:207-258): the min of the element steps
// and the sum of the scaling masses, formed by the wave alone (no barrier)
// in a fixed order.  Every lane of the wave calls it with its slot s.  The
// min sets the step; the sum is only reported (the largest added mass).
template <int KIND>
__device__ __forceinline__ void dt_partial(const InstDev& D, double dt_floor, int s, int first, int end) {
    double mn = INFINITY, ex = 0.0;
    if (s < end) dt_element(D, s, dt_floor, &mn, &ex);
    mn = wave_reduce<64>(mn, 1);
    ex = wave_reduce<64>(ex, 0);
    if ((threadIdx.x & 63) == 0 && first < end) {
        const int slot = dt_slot<KIND>(D, first);
        D.dtp_min[slot] = mn;
        D.dtp_sum[slot] = ex;
    }
}

// Tile of workgroup b of a grid of `grid` workgroups.  Workgroups are
// dispatched to the 8 XCDs in turn (workgroup b to XCD b mod 8).  Dealing
// the tile list to the XCDs in chunks of `chunk` consecutive tiles keeps
// neighbouring tiles of an instance, which share the nodes on their
// boundaries, under one L2.  A grid that is not a multiple of 8 chunks
// keeps the identity.
__device__ __forceinline__ int xcd_tile(int b, int grid, int chunk) {
    if (grid % (8 * chunk)) return b;
    const int j = b >> 3;  // the workgroup's turn on its XCD
    return ((j / chunk) * 8 + (b & 7)) * chunk + j % chunk;
}

// Crush of every group < 8 (group_crush's arithmetic, src/bin/crushrs.rs:
// 265-271) in two parts, so that the loads of all groups are in flight
// together: each thread's partial sums over the front nodes, the
// accelerometer displacements and the front counts (0 where the crush is
// 0); then, after crush_finish's block sums, the crush.
struct CrushPart {
    double s[8], ua[8];
    int cnt[8];
};

template <int B>
__device__ __forceinline__ void crush_partials(const InstDev& D, int ng, CrushPart& p) {
#pragma clang fp contract(off)
    const int t = threadIdx.x;
    int b[8], e[8], ax[8], na[8], fr[8];
#pragma unroll
    for (int g = 0; g < 8; ++g)
        if (g < ng) {
            b[g] = D.g_front_ptr[g];
            e[g] = D.g_front_ptr[g + 1];
            ax[g] = D.g_axis[g];
            na[g] = D.g_accel[g];
        }
#pragma unroll
    for (int g = 0; g < 8; ++g)
        if (g < ng) {
            fr[g] = b[g] + t < e[g] ? D.g_front[b[g] + t] : -1;
            p.ua[g] = na[g] >= 0 ? D.u[3 * (size_t)na[g] + ax[g]] : 0.0;
            p.cnt[g] = (na[g] < 0 || e[g] <= b[g]) ? 0 : e[g] - b[g];
        }
#pragma unroll
    for (int g = 0; g < 8; ++g)
        if (g < ng) {
            p.s[g] = 0.0;
            if (fr[g] >= 0) p.s[g] += D.u[3 * (size_t)fr[g] + ax[g]];
            p.s[g] = front_sum<B>(D, b[g] + t + B, e[g], ax[g], p.s[g]);
        }
}

template <int B>
__device__ __forceinline__ void crush_finish(int ng, CrushPart& p, double* s_wn, double* cr) {
#pragma clang fp contract(off)
    const int op[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    block_reduce_n<B, 8>(p.s, op, ng, s_wn);
#pragma unroll
    for (int g = 0; g < 8; ++g)
        if (g < ng) cr[g] = p.cnt[g] == 0 ? 0.0 : fabs(p.ua[g] - p.s[g] / (double)p.cnt[g]);
}

// The control block moves between device memory and LDS as 8-byte words,
// one per thread.
constexpr int CTL_WORDS = (int)(sizeof(StepCtl) / 8);
static_assert(sizeof(StepCtl) % 8 == 0, "StepCtl is copied as 8-byte words");
template <int B>
__device__ __forceinline__ void ctl_load(const StepCtl* src, StepCtl* c) {
    static_assert(CTL_WORDS <= B, "one word per thread");
    if ((int)threadIdx.x < CTL_WORDS)
        reinterpret_cast<unsigned long long*>(c)[threadIdx.x] =
            reinterpret_cast<const unsigned long long*>(src)[threadIdx.x];
    __syncthreads();
}
// After the block's last change to c.
template <int B>
__device__ __forceinline__ void ctl_store(StepCtl* dst, const StepCtl* c) {
    __syncthreads();
    if ((int)threadIdx.x < CTL_WORDS)
        reinterpret_cast<unsigned long long*>(dst)[threadIdx.x] =
            reinterpret_cast<const unsigned long long*>(c)[threadIdx.x];
}

// Control pass of a tick (src/solver.rs:287-352), one workgroup per active
// instance.  A tick runs, per instance, the contact, first half step and
// element forces of one step (PH_HALF) and the gather of the step before it
// (PH_GATHER); a gather and the next half step share one node pass, which
// keeps the gathered forces and the full-step velocity in registers.  The
// pass
//  (a) after a gather alone, closes that step in full (step_end); after a
//      gather fused with the next half step, streams the step's pulse
//      sample from the probe records, and stops the instance if that gather
//      left a non-finite velocity;
//  (b) holds the step in flight (half step and forces done) for a gather
//      alone when its end reads the full-step state: the solve may end
//      there, it takes a history sample or it checks separation.  Any other
//      step is closed now as far as its displacement allows (time, step
//      count, crush, contact and erosion counts);
//  (c) opens the next step with its stable step (from the force tiles'
//      partials, at check steps).
// Every step thus runs the same arithmetic, in the same order per node, as
// the one-solve path.  One difference: when a fused gather leaves a
// non-finite velocity, the solve stops at that step with the error flag,
// but the velocities and displacements left behind (and the reported
// delta-v, crush and last history record) are those after the next half
// step.
template <int B>
__device__ __forceinline__ void ctl_block(const InstDev* __restrict__ table, const int32_t* __restrict__ list,
                          const int32_t* __restrict__ n_list, uint8_t* __restrict__ done_flag,
                          int* __restrict__ n_done, int bid) {
    __shared__ double s_w[B / 64];
    __shared__ double s_wn[8 * (B / 64)];
    __shared__ double s_a[8];
    __shared__ StepCtl s_ctl;
    if (bid >= *n_list) return;
    const int inst = list[bid];
    const InstDev& D = table[inst];
    StepCtl* c = &s_ctl;
    const int t = threadIdx.x;
    ctl_load<B>(D.ctl, c);
    if (c->done) return;
    const int ph = c->phase;
    const int ng = D.n_groups < 8 ? D.n_groups : 8;
    // loads that no decision below changes, issued together: the probe
    // samples (a), the crush inputs of the step in flight (b) and the force
    // waves' stable-step partials (c), reduced per thread in slot order
    if ((ph & PH_GATHER) && (ph & PH_HALF) && D.n_pulse > 0 && t < ng) s_a[t] = accel_local_x_probed(D, t);
    CrushPart cp;
    if (ph & PH_HALF) crush_partials<B>(D, ng, cp);
    double mn = INFINITY, sm = 0.0;
    if (c->adaptive) {
        const int slots = dt_slots(D.n_hc, D.E);
        // four slots' loads in flight together, combined in slot order
        for (int j0 = t; j0 < slots; j0 += 4 * B) {
            double a[4], m[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const bool on = j0 + k * B < slots;
                a[k] = on ? D.dtp_min[j0 + k * B] : INFINITY;
                m[k] = on ? D.dtp_sum[j0 + k * B] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (j0 + k * B < slots) {
                    mn = fmin(mn, a[k]);
                    sm += m[k];
                }
        }
    }
    // (a) the last tick's gather
    if (ph & PH_GATHER) {
        bool stop;
        if (ph & PH_HALF) {
            stop = (c->bad_step & BAD_GATHER) != 0;
            if (stop) {
                const int hk = c->hist_k;
                if (D.n_hist > 0 && hk < D.n_hist) history_sample<B>(D, hk, c->step, c->t, s_w, fext_step(D, c->step));
            }
            __syncthreads();  // thread g has formed group g's sample in s_a
            if (t == 0) {
                if (D.n_pulse > 0) pulse_sample(D, c, c->t, s_a);
                if (stop) {
                    if (D.n_hist > 0) c->hist_k += 1;
                    c->error = 1;
                    c->done = 1;
                }
            }
            __syncthreads();
        } else {
            const double dt = step_dt(*c);
            const int hits = c->pair_hits;
            const bool bad = c->bad_step != 0;
            double* fext = fext_step(D, c->step + 1);
            __syncthreads();
            if (t == 0) {
                c->eroded_total += (unsigned long long)c->eroded_step;
                c->eroded_step = 0;
                c->bad_step = 0;
                c->pair_hits = 0;
            }
            __syncthreads();
            stop = step_end<B>(D, c, dt, hits, bad, s_w, fext);
        }
        if (stop) {
            finish_instance<B>(D, c, s_w);
            __syncthreads();
            if (t == 0) c->phase = 0;
            ctl_store<B>(D.ctl, c);
            if (t == 0) {
                done_flag[inst] = 1;
                atomicAdd(n_done, 1);
            }
            return;
        }
    }
    // (b) the step in flight
    bool closed = false;
    if (ph & PH_HALF) {
        const long long q = c->step + 1;
        const double dt = step_dt(*c);
        const int hits = c->pair_hits;
        const int had_pair = hits > 0 ? 1 : c->had_pair;
        const int quiet = hits > 0 ? 0 : c->quiet + 1;
        const double tn = c->t + dt;
        const bool hold = (c->bad_step & BAD_HALF) != 0 || !(tn < c->end_time) ||
                          (D.n_hist > 0 && q % D.hist_steps == 0) ||
                          (c->sep_steps > 0 && D.n_groups >= 2 && had_pair && quiet >= c->sep_steps && q % c->chk == 0);
        if (hold) {
            __syncthreads();
            if (t == 0) c->phase = PH_GATHER;
            ctl_store<B>(D.ctl, c);
            return;
        }
        double cr[8];
        crush_finish<B>(ng, cp, s_wn, cr);
        __syncthreads();
        if (t == 0) {
            for (int g = 0; g < ng; ++g) c->crush_max[g] = fmax(c->crush_max[g], cr[g]);
            c->eroded_total += (unsigned long long)c->eroded_step;
            c->eroded_step = 0;
            c->bad_step = 0;
            c->pair_hits = 0;
            c->had_pair = had_pair;
            c->quiet = quiet;
            c->t = tn;
            c->dt = dt;
            c->step = q;
        }
        __syncthreads();
        closed = true;
    }
    // (c) stable step and selective mass scaling of the next step
    // (src/solver.rs:288-299) from the force tiles' partials, reduced by the
    // whole block
    const bool check = c->adaptive && c->step % c->chk == 0;
    if (check) {
        double v[2] = {mn, sm};
        const int op[2] = {1, 0};
        block_reduce_n<B, 2>(v, op, 2, s_wn);
        mn = v[0];
        sm = v[1];
    }
    if (t == 0) {
        c->scale_now = 0;
        if (check) {
            if (isfinite(mn)) {
                const bool scale = c->dt_floor > 0.0 && mn < c->dt_floor;
                if (scale) {
                    c->added_total_max = fmax(c->added_total_max, sm);
                    c->scale_now = 1;
                }
                c->dt_cfl = c->dt_scale * fmin(scale ? c->dt_floor : mn, c->contact_dt);
            }
        }
        c->phase = PH_HALF | (closed ? PH_GATHER : 0);
    }
    ctl_store<B>(D.ctl, c);
}

template <int B>
__global__ __launch_bounds__(B) void k_ctl(const InstDev* __restrict__ table, const int32_t* __restrict__ list,
                                           const int32_t* __restrict__ n_list,
                                           uint8_t* __restrict__ done_flag, int* __restrict__ n_done) {
    ctl_block<B>(table, list, n_list, done_flag, n_done, (int)blockIdx.x);
}

// Contact (src/contact.rs:92-160), in four or five launches over all open
// steps.  (1) geometry of the faces of one contact of one instance (done by
// k_facegeo_tiles for large contacts), and its face grid in the contact's
// device-memory region (8 cells and 64 entries per face, at least
// BATCH_GRID_CELLS and BATCH_GRID_ENTRIES).  A grid is kept over steps
// while every face stays inside its box at the build grown by its margin,
// so builds are rare; workgroups of 512 or more threads build in LDS.  The grid layout
// never changes which face a node hits: the search keeps the lowest
// accepting face index.
template <int B>
__device__ __forceinline__ void grid_block(const InstDev* __restrict__ table, const int2* __restrict__ tiles,
                                           const int32_t* __restrict__ n_tiles, int bid, bool any_phase) {
    __shared__ double s_w[B / 64];
    __shared__ int s_tmp[B / 64];
    if (bid >= *n_tiles) return;
    const int2 tile = tiles[bid];
    const InstDev& D = table[tile.x];
    if (any_phase ? D.ctl->done != 0 : !(D.ctl->phase & PH_HALF)) return;
    const int ci = tile.y;
    const ContactDesc cd = D.cdesc[ci];
    // faces of fixed nodes never move
    const FaceGrid prev = D.cgp[ci];
    const bool reusable = prev.built && !prev.overflow;
    if (reusable && cd.all_static) return;
    if (large_moving_contact(cd)) {
        // face pass and check done by k_facegeo_tiles, which marks the grid
        // stale (FaceGrid::stale)
        if (reusable && !prev.stale) return;
    } else {
        for (int f = threadIdx.x; f < cd.n_faces; f += B) contact_face(D, cd.face_begin + f);
        __syncthreads();
        if (reusable && grid_still_valid<B>(D, cd, prev.margin)) return;
    }
    __syncthreads();  // the face geometry and the margin levels
    int32_t* gs = D.cgrid + cd.grid_off;
    int32_t* ge = gs + cd.grid_cells + 1;
    FaceGrid g;
    if constexpr (B >= 512) {
        // large contacts: cell counts, scan and fill cursors in LDS (the
        // fill's returning atomics are the rebuild's latency chain), the
        // starts copied out once
        __shared__ int l_start[LDS_GRID_CELLS + 1];
        __shared__ int l_cursor[LDS_GRID_CELLS];
        g = build_face_grid<B>(D, cd, D.cell_scale, min(cd.grid_cells, LDS_GRID_CELLS), cd.grid_entries, l_start,
                               l_cursor, ge, s_w, s_tmp, GRID_MARGIN);
        const int nc = g.overflow ? 1 : g.dim[0] * g.dim[1] * g.dim[2];
        for (int i = threadIdx.x; i <= nc; i += B) gs[i] = l_start[i];
    } else {
        g = build_face_grid<B>(D, cd, D.cell_scale, cd.grid_cells, cd.grid_entries, gs, ge + cd.grid_entries, ge, s_w,
                               s_tmp, GRID_MARGIN);
    }
    if (threadIdx.x == 0) D.cgp[ci] = g;
}

template <int B>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_grid(const InstDev* __restrict__ table, const int2* __restrict__ tiles,
                                            const int32_t* __restrict__ n_tiles) {
    grid_block<B>(table, tiles, n_tiles, (int)blockIdx.x, false);
}

// (1a) face geometry of large moving contacts, tiles of T faces spread over
// the GPU (one workgroup per contact would run their f64 face pass on one
// CU each), four lanes per face (one triangle of the candidate region
// each: contact_face_quad); tile.y = contact << 20 | first face in the
// contact.  Each face is checked against its contact's kept grid, and a
// face outside it marks the grid stale for k_grid.
template <int T>
__device__ __forceinline__ void facegeo_block(const InstDev* __restrict__ table, const int2* __restrict__ tiles,
                              const int32_t* __restrict__ n_tiles, int chunk, int bid, int grid, bool any_phase) {
    const int b = xcd_tile(bid, grid, chunk);
    if (b >= *n_tiles) return;
    const int2 tile = tiles[b];
    const InstDev& D = table[tile.x];
    if (any_phase ? D.ctl->done != 0 : !(D.ctl->phase & PH_HALF)) return;
    const int ci = tile.y >> 20;
    const ContactDesc cd = D.cdesc[ci];
    // the four lanes of a face leave together
    const int fl = (tile.y & 0xFFFFF) + (int)(threadIdx.x >> 2);
    if (fl >= cd.n_faces) return;
    const int f = cd.face_begin + fl;
    contact_face_quad(D, f, (int)(threadIdx.x & 3));
    if (threadIdx.x & 3) return;
    const FaceGrid prev = D.cgp[ci];
    if (!prev.built || prev.overflow || D.cf_static[f]) return;
    if (!kept_box_holds(D, cd, f, prev.margin)) atomicOr(&D.cgp[ci].stale, 1);
}

template <int T>
__global__ __launch_bounds__(4 * T) void k_facegeo_tiles(const InstDev* __restrict__ table, const int2* __restrict__ tiles,
                                                         const int32_t* __restrict__ n_tiles, int chunk) {
    facegeo_block<T>(table, tiles, n_tiles, chunk, (int)blockIdx.x, (int)gridDim.x, false);
}

// The control pass and the large contacts' face geometry in one launch,
// when the control workgroup has the face tiles' 4T threads: the first
// n_ctl blocks (a multiple of 8, so that the face tiles keep their XCD
// chunks) run k_ctl's work, the rest k_facegeo_tiles'.  The two touch
// disjoint data (the face pass reads the displacement the last node pass
// left, which the control pass does not write), so they overlap.  The face
// pass cannot see the phase the control pass is setting, so it runs for
// every instance not yet done: on a tick that holds a step for a gather
// alone it recomputes the same face records from the same displacement,
// and at most raises a face's grid margin level once more, which changes
// the grid's layout but never a hit (the search keeps the lowest
// accepting face).
// The control pass and the grid pass in one launch, for batches with no
// contact of more than 256 moving faces (whose face tiles must mark their
// grids stale before the grid pass): the first n_ctl blocks run k_ctl's
// work, the rest k_grid's at the control workgroup's size (the face pass
// and the build loop over faces, so any size serves).  As in k_ctl_facegeo
// the grid pass runs for every instance not done, and on a tick that holds
// a step it recomputes the same face records and at most rebuilds a grid,
// which never changes a hit.
template <int B>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_ctl_grid(const InstDev* __restrict__ table, const int32_t* __restrict__ list,
                                                const int32_t* __restrict__ n_list, uint8_t* __restrict__ done_flag,
                                                int* __restrict__ n_done, int n_ctl, const int2* __restrict__ tiles,
                                                const int32_t* __restrict__ n_tiles) {
    if ((int)blockIdx.x < n_ctl)
        ctl_block<B>(table, list, n_list, done_flag, n_done, (int)blockIdx.x);
    else
        grid_block<B>(table, tiles, n_tiles, (int)blockIdx.x - n_ctl, true);
}

template <int T>
__global__ __launch_bounds__(4 * T) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_ctl_facegeo(const InstDev* __restrict__ table,
                                                       const int32_t* __restrict__ list,
                                                       const int32_t* __restrict__ n_list,
                                                       uint8_t* __restrict__ done_flag, int* __restrict__ n_done,
                                                       int n_ctl, const int2* __restrict__ tiles,
                                                       const int32_t* __restrict__ n_tiles, int chunk) {
    if ((int)blockIdx.x < n_ctl)
        ctl_block<4 * T>(table, list, n_list, done_flag, n_done, (int)blockIdx.x);
    else
        facegeo_block<T>(table, tiles, n_tiles, chunk, (int)blockIdx.x - n_ctl, (int)gridDim.x - n_ctl, true);
}

// (2) node search, tiles of CONTACT_TILE contacting nodes of one contact of
// one instance (tile.y = contact << 20 | first node in the contact), L lanes
// per node, T nodes per workgroup (CONTACT_TILE / T consecutive workgroups
// share a tile); penetrating nodes of pair contacts are counted per
// instance.
// At 5 waves per SIMD the search fits 94 VGPRs without spills (left free
// it takes 98, which allows 4; at 6 it spills 76).
#ifndef CRUSH_SEARCH_WAVES
#define CRUSH_SEARCH_WAVES 5
#endif
template <int T, int L>
__global__ __launch_bounds__(T * L) __attribute__((amdgpu_waves_per_eu(CRUSH_SEARCH_WAVES, 8))) void k_search_tiles(const InstDev* __restrict__ table,
                                                        const int2* __restrict__ tiles,
                                                        const int32_t* __restrict__ n_tiles, int chunk) {
    constexpr int P = CONTACT_TILE / T;
    static_assert(P * T == CONTACT_TILE, "a tile splits into whole workgroups");
    const int b = xcd_tile((int)blockIdx.x / P, (int)gridDim.x / P, chunk);
    if (b >= *n_tiles) return;
    const int2 tile = tiles[b];
    const InstDev& D = table[tile.x];
    StepCtl& c = *D.ctl;
    if (!(c.phase & PH_HALF)) return;
    const int ci = tile.y >> 20;
    const ContactDesc cd = D.cdesc[ci];
    const int q = (tile.y & 0xFFFFF) + ((int)blockIdx.x % P) * T + (int)threadIdx.x / L;
    int hit = 0;
    if (q < cd.n_nodes) {
        const FaceGrid g = D.cgp[ci];
        const int32_t* gs = D.cgrid + cd.grid_off;
        hit = contact_node<L>(D, cd, q, g, gs, gs + cd.grid_cells + 1, step_dt(c), (int)threadIdx.x % L);
    }
    const unsigned long long m = __ballot(hit != 0 && cd.pair);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(&c.pair_hits, (int)__popcll(m));
}

// (3) reactions of every contact face; (4) contact force of every touched
// node.
template <int T>
__global__ __launch_bounds__(T) void k_faceforce_tiles(const InstDev* __restrict__ table,
                                                       const int2* __restrict__ tiles,
                                                       const int32_t* __restrict__ n_tiles, int chunk) {
    const int b = xcd_tile(blockIdx.x, gridDim.x, chunk);
    if (b >= *n_tiles) return;
    const int2 tile = tiles[b];
    const InstDev& D = table[tile.x];
    if (!(D.ctl->phase & PH_HALF)) return;
    const int f = tile.y + (int)threadIdx.x;
    contact_face_force_wave(D, f, f < D.F);
}

// Element forces of the open steps: one workgroup per tile of T element
// slots of one instance and one material kind.  When the next step is a
// check step, each wave's stable-step partial is formed at the same
// configuration (the positions and erosion the next step starts from).
template <int KIND, int T>
__device__ __forceinline__ void force_tile(const InstDev* __restrict__ table, const int2* __restrict__ tiles,
                                           const int32_t* __restrict__ n_tiles, int chunk, float* s_q) {
    const int b = xcd_tile(blockIdx.x, gridDim.x, chunk);
    if (b >= *n_tiles) return;
    const int2 tile = tiles[b];
    const InstDev& D = table[tile.x];
    StepCtl& c = *D.ctl;
    if (!(c.phase & PH_HALF)) return;
    const int end = KIND == CRUSH_MAT_HONEYCOMB ? D.n_hc : D.E;
    const int s = tile.y + (int)threadIdx.x;
    int er = 0;
    // the elastic law parks the hourglass coordinates in LDS (see
    // force_element_class); the honeycomb kernel keeps them in registers
    constexpr int LS = KIND == CRUSH_MAT_HONEYCOMB ? 0 : T;
    if (s < end) er = force_element<KIND, true, LS>(D, s, LS > 0 ? s_q + threadIdx.x : nullptr);
    const unsigned long long m = __ballot(er != 0);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(&c.eroded_step, (int)__popcll(m));
    if (c.adaptive && (c.step + 1) % c.chk == 0)
        dt_partial<KIND>(D, c.dt_floor, s, tile.y + (int)(threadIdx.x & ~63u), end);
}

// Lower bound of the waves per SIMD the honeycomb kernel is compiled for:
// at 4 it fits 128 VGPRs with no spills (the class constants sit in SGPRs).
#ifndef CRUSH_HC_WAVES
#define CRUSH_HC_WAVES 4
#endif

template <int T>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(CRUSH_HC_WAVES, 8))) void k_force_hc(
    const InstDev* __restrict__ table, const int2* __restrict__ tiles, const int32_t* __restrict__ n_tiles, int chunk) {
    force_tile<CRUSH_MAT_HONEYCOMB, T>(table, tiles, n_tiles, chunk, nullptr);
}

// The elastic kernel at 4 waves per SIMD: with the hourglass coordinates in
// LDS through the Hencky law it fits 128 VGPRs without spills (8 % faster
// than the build that spilled 21 VGPRs, which ran 9 % faster than 3 waves).
template <int T>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_force_el(
    const InstDev* __restrict__ table, const int2* __restrict__ tiles, const int32_t* __restrict__ n_tiles, int chunk) {
    __shared__ float s_q[12 * T];
    force_tile<CRUSH_MAT_ELASTIC, T>(table, tiles, n_tiles, chunk, s_q);
}

// Stable-step partials of instances that have not taken a step yet.
template <int KIND, int T>
__global__ __launch_bounds__(T) void k_dt_tiles(const InstDev* __restrict__ table, const int2* __restrict__ tiles,
                                                const int32_t* __restrict__ n_tiles) {
    if ((int)blockIdx.x >= *n_tiles) return;
    const int2 tile = tiles[blockIdx.x];
    const InstDev& D = table[tile.x];
    const StepCtl& c = *D.ctl;
    if (c.done || c.phase != 0 || c.step != 0 || !c.adaptive) return;
    const int end = KIND == CRUSH_MAT_HONEYCOMB ? D.n_hc : D.E;
    dt_partial<KIND>(D, c.dt_floor, tile.y + (int)threadIdx.x, tile.y + (int)(threadIdx.x & ~63u), end);
}

// Node pass of a tick, tiles of T nodes of one instance: the gather and
// second half step of the step in flight (src/solver.rs:327-342) and the
// first half step of the next (src/solver.rs:313-323), after the node's
// scaling mass at steps that scale (src/kernel.rs:243-258).  Fused, the
// gathered force and full-step velocity stay in registers and only the
// half-step velocity and the displacement are written; the probe nodes
// record their acceleration and position for the pulses.  A gather alone
// writes f_int and v, which a half step alone (an instance's first step, or
// the step after a gather alone) reads.
template <int T>
__global__ __launch_bounds__(T) void k_node_tiles(const InstDev* __restrict__ table, const int2* __restrict__ tiles,
                                                  const int32_t* __restrict__ n_tiles, int chunk) {
    const int b = xcd_tile(blockIdx.x, gridDim.x, chunk);
    if (b >= *n_tiles) return;
    const int2 tile = tiles[b];
    const InstDev& D = table[tile.x];
    StepCtl& c = *D.ctl;
    const int ph = c.phase;
    if (ph == 0) return;
    const int n = tile.y + (int)threadIdx.x;
    if (n >= D.N) return;
    const uint8_t fl = gld(D.fixed + n);
    const bool fx = (fl & NODE_FIXED) != 0;
    const double dth = step_dt(c);
    double im = gld(D.inv_mass + n);
    double f[3], vd[3];
    {
        const Double3 v = ld_d3(D.v, n);
        vd[0] = v.x;
        vd[1] = v.y;
        vd[2] = v.z;
    }
    if (ph & PH_GATHER) {
        // the step in flight: closed by the control pass when a half step
        // follows, else still open
        const bool fused = (ph & PH_HALF) != 0;
        const long long qg = fused ? c.step : c.step + 1;
        const double dtg = fused ? c.dt : dth;
        gather_sum(D, n, f);
        double fe[3] = {0.0, 0.0, 0.0};
        if (fl & NODE_CONTACT) {
            const Double3 e = ld_d3(fext_step(D, qg), n);
            fe[0] = e.x;
            fe[1] = e.y;
            fe[2] = e.z;
        }
        double vn[3], a[3];
        if (!gather_update(f, fe, im, dtg, c.damping, fx, vd, vn, a)) atomicOr(&c.bad_step, BAD_GATHER);
        if (!fused) {
            st_d3(D.f_int, n, Double3{f[0], f[1], f[2]});
            st_d3(D.v, n, Double3{vn[0], vn[1], vn[2]});
            return;
        }
        if (fl & NODE_PROBE) {
            double* r = D.probe + 6 * (size_t)probe_slot(D, n);
            double x[3];
            node_position(D, n, x);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                r[d] = a[d];
                r[3 + d] = x[d];
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) vd[d] = vn[d];
    } else {
        const Double3 fi = ld_d3(D.f_int, n);
        f[0] = fi.x;
        f[1] = fi.y;
        f[2] = fi.z;
    }
    if (c.scale_now) im = mass_node(D, n);
    // the contact force of the next step, from this node's apply list (in
    // contact order, as the one-solve path's apply pass forms it), kept for
    // that step's gather
    double fe[3] = {0.0, 0.0, 0.0};
    if (fl & NODE_CONTACT) {
        const Double3 e = contact_apply_sum(D, gld(D.tn_of + n));
        st_d3(fext_step(D, c.step + 1), n, e);
        fe[0] = e.x;
        fe[1] = e.y;
        fe[2] = e.z;
    }
    if (fx) {
        st_d3(D.u, n, Double3{0.0, 0.0, 0.0});
        st_d3(D.v, n, Double3{0.0, 0.0, 0.0});
        return;
    }
    const Double3 u = ld_d3(D.u, n);
    const double ud[3] = {u.x, u.y, u.z};
    double vh[3], un[3];
    if (!half_update(f, fe, im, dth, vd, ud, vh, un)) atomicOr(&c.bad_step, BAD_HALF);
    st_d3(D.v, n, Double3{vh[0], vh[1], vh[2]});
    st_d3(D.u, n, Double3{un[0], un[1], un[2]});
}

// Compact per-instance summary for the host.
__global__ void k_collect(const InstDev* __restrict__ table, int n, BatchSummary* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const InstDev& D = table[i];
    const StepCtl& c = *D.ctl;
    BatchSummary& r = out[i];
    r.t = c.t;
    r.steps = c.step;
    r.eroded = (long long)c.eroded_total;
    r.added_max = c.added_total_max;
    r.done = c.done;
    r.error = c.error;
    r.separated = c.separated;
    r.pulse_samples = c.pulse_k;
    r.hist_samples = c.hist_k;
    r.t_retire = c.t_retire;
    for (int g = 0; g < 8; ++g) {
        r.crush_max[g] = g < D.n_groups ? c.crush_max[g] : 0.0;
        r.crush_end[g] = g < D.n_groups ? c.crush_end[g] : 0.0;
        for (int d = 0; d < 3; ++d) r.dv[g][d] = g < D.n_groups ? D.g_dv[3 * g + d] : 0.0;
    }
}

__global__ void k_stamp(unsigned long long* p) { *p = wall_clock64(); }

// ------------------------------------------------------------- host


static int inst_block(const crush_batch_options& o) {
    const int b = o.threads_per_instance;
    return (b == 64 || b == 128 || b == 512) ? b : 256;
}

static int force_tile(const crush_batch_options& o) {
    const int t = o.force_tile;
    return (t == 64 || t == 128) ? t : 256;
}

int steps_per_graph(const crush_batch_options& o) { return o.steps_per_graph > 0 ? o.steps_per_graph : 32; }

double cell_scale_of(const crush_batch_options& o) { return o.contact_cell_scale > 0.0 ? o.contact_cell_scale : 1.0; }


// Tiles of instance i for list L with force tile T.
static void add_tiles(const Prepared& P, int i, int T, std::vector<int2>* out) {
    for (int s = 0; s < P.n_hc; s += T) out[L_HC].push_back(make_int2(i, s));
    for (int s = P.n_hc; s < P.n_elems; s += T) out[L_EL].push_back(make_int2(i, s));
    for (int k = 0; k < P.n_nodes; k += NODE_TILE) out[L_NODE].push_back(make_int2(i, k));
    for (int f = 0; f < P.n_cfaces; f += CONTACT_TILE) out[L_FACE].push_back(make_int2(i, f));
    for (int ci = 0; ci < (int)P.cdesc.size(); ++ci) {
        const ContactDesc& cd = P.cdesc[ci];
        if (cd.n_faces == 0) continue;
        out[L_GRID].push_back(make_int2(i, ci));
        if (large_moving_contact(cd)) {
            if (ci >= (1 << 11) || cd.n_faces >= (1 << 20))
                fail(CRUSH_E_INPUT, "contact %d of %d faces: beyond the face tiles' 2048 contacts / 2^20 faces", ci,
                     cd.n_faces);
            for (int f = 0; f < cd.n_faces; f += CONTACT_TILE) out[L_FGEO].push_back(make_int2(i, (ci << 20) | f));
        }
        if (cd.n_nodes > 0 && (ci >= (1 << 11) || cd.n_nodes >= (1 << 20)))
            fail(CRUSH_E_INPUT, "contact %d of %d nodes: beyond the search tiles' 2048 contacts / 2^20 nodes", ci,
                 cd.n_nodes);
        for (int q = 0; q < cd.n_nodes; q += CONTACT_TILE) out[L_SEARCH].push_back(make_int2(i, (ci << 20) | q));
    }
}

// Rebuild the active lists from the instances not yet retired.
void rebuild_lists(BatchImpl& bi) {
    const int n = (int)bi.prep.size(), T = force_tile(bi.opt);
    std::vector<int32_t> list;
    std::vector<int2> tiles[N_LISTS];
    int max_faces = 0;
    for (int i = 0; i < n; ++i) {
        if (bi.host_done[i]) continue;
        list.push_back(i);
        add_tiles(bi.prep[i], i, T, tiles);
        for (const ContactDesc& cd : bi.prep[i].cdesc)
            if (!cd.all_static) max_faces = std::max(max_faces, cd.n_faces);
    }
    // one k_grid thread per face of the largest contact whose faces move,
    // from one wave up to sixteen (larger contacts loop); a server (headroom
    // above 1) shrinks it only to a quarter or less, so arrivals and
    // retirements do not force captures back and forth
    int gb = 64;
    while (gb < 1024 && gb < max_faces) gb *= 2;
    if (gb > bi.grid_block || (gb < bi.grid_block && (bi.headroom <= 1.0 || 4 * gb <= bi.grid_block))) {
        bi.grid_block = gb;
        bi.need_capture = true;
    }
    // few search tiles: 4 lanes per contacting node, fewer still: 16; a
    // server moves to more lanes at half a threshold going down and to fewer
    // at the threshold going up
    const int ns = (int)tiles[L_SEARCH].size();
    auto lanes_at = [&](int scale2) {  // thresholds x scale2 / 2
        return 2 * ns < SEARCH_WIDE_TILES * scale2 ? 16 : 2 * ns < SEARCH_SPLIT_TILES * scale2 ? 4 : 1;
    };
    int sl = lanes_at(2);
    if (bi.headroom > 1.0) sl = std::min(std::max(bi.search_lanes, lanes_at(1)), lanes_at(2));
    if (sl != bi.search_lanes) {
        bi.search_lanes = sl;
        bi.need_capture = true;
    }
    // grow the list buffers (a graph captured with the old ones is stale)
    auto need = [&](auto& buf, size_t want) {
        if (want > buf.n) {
            HIP_OK(hipStreamSynchronize(bi.stream));
            buf.alloc(want + want / 2);
            bi.need_capture = true;
        }
    };
    need(bi.d_list, list.size());
    for (int L = 1; L < N_LISTS; ++L) need(bi.d_tiles[L], tiles[L].size());
    int32_t counts[N_LISTS];
    counts[L_INST] = (int32_t)list.size();
    if (!list.empty())
        HIP_OK(hipMemcpyAsync(bi.d_list.p, list.data(), list.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                              bi.stream));
    for (int L = 1; L < N_LISTS; ++L) {
        counts[L] = (int32_t)tiles[L].size();
        if (!tiles[L].empty())
            HIP_OK(hipMemcpyAsync(bi.d_tiles[L].p, tiles[L].data(), tiles[L].size() * sizeof(int2),
                                  hipMemcpyHostToDevice, bi.stream));
    }
    HIP_OK(hipMemcpyAsync(bi.d_counts.p, counts, sizeof counts, hipMemcpyHostToDevice, bi.stream));
    HIP_OK(hipStreamSynchronize(bi.stream));
    for (int L = 0; L < N_LISTS; ++L) bi.cnt[L] = counts[L];
}

// Grid of a launch that deals its tiles to XCDs in chunks (xcd_tile): the
// listed tiles rounded up to a multiple of 8 chunks.
static unsigned grid8(int n, int chunk) {
    const int m = 8 * chunk;
    return (unsigned)((n + m - 1) / m * m);
}

template <int B, int T>
static void launch_step(BatchImpl& bi, hipEvent_t* ev) {
    const int32_t* cnt = bi.d_counts.p;
    const InstDev* tab = bi.table.p;
    hipStream_t st = bi.stream;
    auto mark = [&](int k) {
        if (ev) HIP_OK(hipEventRecord(ev[k], st));
    };
    mark(K_END);
    // with a control workgroup of the face tiles' size, the large contacts'
    // face geometry shares the control pass's launch
    const bool merged = B == 4 * CONTACT_TILE && bi.cap[L_FGEO] > 0;
    // with no large contact and a control workgroup of at most 128 threads
    // (3 waves per SIMD; at 256 the grid pass's share runs at 2 and half its
    // threads idle, slower than the two launches), the grid pass shares the
    // control pass's launch
    const bool merged_grid = B <= 128 && bi.cap[L_FGEO] == 0 && bi.cap[L_GRID] > 0;
    if (merged_grid) {
        if constexpr (B <= 128) {
            const int n_ctl = std::max(bi.cap[L_INST], 1);
            hipLaunchKernelGGL(k_ctl_grid<B>, dim3(n_ctl + bi.cap[L_GRID]), dim3(B), 0, st, tab, bi.d_list.p,
                               cnt + L_INST, bi.done_flag.p, bi.n_done.p, n_ctl, bi.d_tiles[L_GRID].p, cnt + L_GRID);
        }
    } else if (merged) {
        const int n_ctl = (std::max(bi.cap[L_INST], 1) + 7) / 8 * 8;
        hipLaunchKernelGGL(k_ctl_facegeo<CONTACT_TILE>, dim3(n_ctl + grid8(bi.cap[L_FGEO], bi.chunk[L_FGEO])),
                           dim3(4 * CONTACT_TILE), 0, st, tab, bi.d_list.p, cnt + L_INST, bi.done_flag.p, bi.n_done.p,
                           n_ctl, bi.d_tiles[L_FGEO].p, cnt + L_FGEO, bi.chunk[L_FGEO]);
    } else {
        hipLaunchKernelGGL(k_ctl<B>, dim3(std::max(bi.cap[L_INST], 1)), dim3(B), 0, st, tab, bi.d_list.p,
                           cnt + L_INST, bi.done_flag.p, bi.n_done.p);
    }
    mark(K_GRID);
    if (!merged && !merged_grid && bi.cap[L_FGEO] > 0)
        hipLaunchKernelGGL(k_facegeo_tiles<CONTACT_TILE>, dim3(grid8(bi.cap[L_FGEO], bi.chunk[L_FGEO])), dim3(4 * CONTACT_TILE), 0, st, tab,
                           bi.d_tiles[L_FGEO].p, cnt + L_FGEO, bi.chunk[L_FGEO]);
    if (!merged_grid && bi.cap[L_GRID] > 0) {
        const dim3 g(bi.cap[L_GRID]);
        if (bi.grid_block == 64)
            hipLaunchKernelGGL(k_grid<64>, g, dim3(64), 0, st, tab, bi.d_tiles[L_GRID].p, cnt + L_GRID);
        else if (bi.grid_block == 128)
            hipLaunchKernelGGL(k_grid<128>, g, dim3(128), 0, st, tab, bi.d_tiles[L_GRID].p, cnt + L_GRID);
        else if (bi.grid_block == 256)
            hipLaunchKernelGGL(k_grid<256>, g, dim3(256), 0, st, tab, bi.d_tiles[L_GRID].p, cnt + L_GRID);
        else if (bi.grid_block == 512)
            hipLaunchKernelGGL(k_grid<512>, g, dim3(512), 0, st, tab, bi.d_tiles[L_GRID].p, cnt + L_GRID);
        else
            hipLaunchKernelGGL(k_grid<1024>, g, dim3(1024), 0, st, tab, bi.d_tiles[L_GRID].p, cnt + L_GRID);
    }
    mark(K_SEARCH);
    if (bi.cap[L_SEARCH] > 0) {
        if (bi.search_lanes == 16)
            hipLaunchKernelGGL((k_search_tiles<CONTACT_TILE / 4, 16>), dim3(4 * grid8(bi.cap[L_SEARCH], bi.chunk[L_SEARCH])),
                               dim3(4 * CONTACT_TILE), 0, st, tab, bi.d_tiles[L_SEARCH].p, cnt + L_SEARCH,
                               bi.chunk[L_SEARCH]);
        else if (bi.search_lanes == 4)
            hipLaunchKernelGGL((k_search_tiles<CONTACT_TILE, 4>), dim3(grid8(bi.cap[L_SEARCH], bi.chunk[L_SEARCH])), dim3(4 * CONTACT_TILE), 0, st,
                               tab, bi.d_tiles[L_SEARCH].p, cnt + L_SEARCH, bi.chunk[L_SEARCH]);
        else
            hipLaunchKernelGGL((k_search_tiles<CONTACT_TILE, 1>), dim3(grid8(bi.cap[L_SEARCH], bi.chunk[L_SEARCH])), dim3(CONTACT_TILE), 0, st,
                               tab, bi.d_tiles[L_SEARCH].p, cnt + L_SEARCH, bi.chunk[L_SEARCH]);
    }
    mark(K_FACEFORCE);
    if (bi.cap[L_FACE] > 0)
        hipLaunchKernelGGL(k_faceforce_tiles<CONTACT_TILE>, dim3(grid8(bi.cap[L_FACE], bi.chunk[L_FACE])), dim3(CONTACT_TILE), 0, st, tab,
                           bi.d_tiles[L_FACE].p, cnt + L_FACE, bi.chunk[L_FACE]);
    mark(K_NODE);
    if (bi.cap[L_NODE] > 0)
        hipLaunchKernelGGL(k_node_tiles<NODE_TILE>, dim3(grid8(bi.cap[L_NODE], bi.chunk[L_NODE])), dim3(NODE_TILE), 0, st, tab,
                           bi.d_tiles[L_NODE].p, cnt + L_NODE, bi.chunk[L_NODE]);
    mark(K_FORCE_HC);
    if (bi.cap[L_HC] > 0)
        hipLaunchKernelGGL(k_force_hc<T>, dim3(grid8(bi.cap[L_HC], bi.chunk[L_HC])), dim3(T), 0, st, tab,
                           bi.d_tiles[L_HC].p, cnt + L_HC, bi.chunk[L_HC]);
    mark(K_FORCE_EL);
    if (bi.cap[L_EL] > 0)
        hipLaunchKernelGGL(k_force_el<T>, dim3(grid8(bi.cap[L_EL], bi.chunk[L_EL])), dim3(T), 0, st, tab,
                           bi.d_tiles[L_EL].p, cnt + L_EL, bi.chunk[L_EL]);
    mark(N_KERNELS);
}

template <int T>
static void launch_dt_init(BatchImpl& bi) {
    if (bi.cnt[L_HC] > 0)
        hipLaunchKernelGGL((k_dt_tiles<CRUSH_MAT_HONEYCOMB, T>), dim3(bi.cnt[L_HC]), dim3(T), 0, bi.stream, bi.table.p,
                           bi.d_tiles[L_HC].p, bi.d_counts.p + L_HC);
    if (bi.cnt[L_EL] > 0)
        hipLaunchKernelGGL((k_dt_tiles<CRUSH_MAT_ELASTIC, T>), dim3(bi.cnt[L_EL]), dim3(T), 0, bi.stream, bi.table.p,
                           bi.d_tiles[L_EL].p, bi.d_counts.p + L_EL);
}

template <int B>
static void launch_step_t(BatchImpl& bi, hipEvent_t* ev) {
    const int T = force_tile(bi.opt);
    if (T == 64)
        launch_step<B, 64>(bi, ev);
    else if (T == 128)
        launch_step<B, 128>(bi, ev);
    else
        launch_step<B, 256>(bi, ev);
}

constexpr int EV_PER_STEP = N_KERNELS + 1;

void dt_init(BatchImpl& bi) {
    const int T = force_tile(bi.opt);
    if (T == 64)
        launch_dt_init<64>(bi);
    else if (T == 128)
        launch_dt_init<128>(bi);
    else
        launch_dt_init<256>(bi);
}

void collect(BatchImpl& bi, int n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_collect, dim3((n + 255) / 256), dim3(256), 0, bi.stream, bi.table.p, n, bi.summary.p);
    bi.host_summary.resize(n);
    HIP_OK(hipMemcpyAsync(bi.host_summary.data(), bi.summary.p, n * sizeof(BatchSummary), hipMemcpyDeviceToHost,
                          bi.stream));
    HIP_OK(hipStreamSynchronize(bi.stream));
}

// One step's launches; ev (EV_PER_STEP events, or null) brackets the kernels.
static void launch_one(BatchImpl& bi, hipEvent_t* ev) {
    const int B = inst_block(bi.opt);
    if (B == 64)
        launch_step_t<64>(bi, ev);
    else if (B == 128)
        launch_step_t<128>(bi, ev);
    else if (B == 512)
        launch_step_t<512>(bi, ev);
    else
        launch_step_t<256>(bi, ev);
}

// Add the kernel durations of the last replay.
static void read_kernel_times(BatchImpl& bi) {
    if (!bi.time_kernels) return;
    for (int r = 0; r < steps_per_graph(bi.opt); ++r)
        for (int k = 0; k < N_KERNELS; ++k) {
            float ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&ms, bi.ev[EV_PER_STEP * r + k], bi.ev[EV_PER_STEP * r + k + 1]));
            bi.t_kernel[k] += ms;
        }
}

// Capture steps_per_graph steps with grids sized to the current lists.
void capture(BatchImpl& bi) {
    if (bi.graph) {
        HIP_OK(hipGraphExecDestroy(bi.graph));
        bi.graph = nullptr;
    }
    for (int L = 0; L < N_LISTS; ++L) bi.cap[L] = (int)((double)bi.cnt[L] * bi.headroom) + (bi.headroom > 1.0 ? 1 : 0);
    // XCD chunks of the force and node tiles: an eighth of an instance's
    // tiles (one instance spreads over the 8 L2s in 8 runs), at least 4
    // (many instances: a measured 2 % faster honeycomb pass than 1)
    for (int L : {L_HC, L_EL, L_NODE, L_FGEO, L_SEARCH, L_FACE}) {
        const int per_inst = (bi.cap[L] + std::max(bi.cap[L_INST], 1) - 1) / std::max(bi.cap[L_INST], 1);
        bi.chunk[L] = std::max(4, (per_inst + 7) / 8);
    }
    bi.need_capture = false;
    const int R = steps_per_graph(bi.opt);
    hipGraph_t g;
    HIP_OK(hipStreamBeginCapture(bi.stream, hipStreamCaptureModeThreadLocal));
    for (int r = 0; r < R; ++r) launch_one(bi, nullptr);
    HIP_OK(hipStreamEndCapture(bi.stream, &g));
    HIP_OK(hipGraphInstantiate(&bi.graph, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
}

// The launches of one graph replay: the graph, or with kernel timing on,
// the same steps launched directly with events around each kernel.
void replay(BatchImpl& bi) {
    const int R = steps_per_graph(bi.opt);
    if (bi.fresh) {
        dt_init(bi);
        bi.fresh = false;
    }
    if (!bi.time_kernels) {
        HIP_OK(hipGraphLaunch(bi.graph, bi.stream));
        return;
    }
    while ((int)bi.ev.size() < EV_PER_STEP * R) {
        hipEvent_t e;
        HIP_OK(hipEventCreate(&e));
        bi.ev.push_back(e);
    }
    for (int r = 0; r < R; ++r) launch_one(bi, &bi.ev[EV_PER_STEP * (size_t)r]);
}

void batch_defaults(crush_batch_options* o) {
    o->threads_per_instance = 256;
    o->force_tile = 256;
    o->steps_per_graph = 32;
    o->host_threads = 0;
    o->contact_cell_scale = 1.0;
    o->autotune = 0;
    o->reserved0 = 0;
}

static void upload_images(BatchImpl& bi) {
    const int n = (int)bi.prep.size();
    for (int i = 0; i < n; ++i) {
        std::vector<uint8_t> img;
        build_image(bi.prep[i], img);
        HIP_OK(hipMemcpyAsync(bi.pool.p + bi.base_off[i], img.data(), img.size(), hipMemcpyHostToDevice, bi.stream));
        HIP_OK(hipStreamSynchronize(bi.stream));
    }
}

// Restore the initial state and the full active lists (with `only`, the
// lists of the instances it marks).
static void reset_state(BatchImpl& bi, bool recapture, const std::vector<uint8_t>* only = nullptr) {
    upload_images(bi);
    bi.n_done.zero(bi.stream);
    bi.done_flag.zero(bi.stream);
    for (size_t i = 0; i < bi.host_done.size(); ++i) bi.host_done[i] = only ? !(*only)[i] : 0;
    rebuild_lists(bi);
    bool same = !bi.need_capture;
    for (int L = 0; L < N_LISTS; ++L) same = same && bi.cap[L] == bi.cnt[L];
    if (recapture || !same) capture(bi);
    bi.fresh = true;
    bi.summary_valid = false;
}

static void autotune(BatchImpl& bi);

BatchImpl* batch_create(const crush_model* const* models, int n, const crush_batch_options* opt) {
    require_device();
    if (n <= 0 || !models) fail(CRUSH_E_INPUT, "batch needs at least one model");
    auto bi = std::make_unique<BatchImpl>();
    batch_defaults(&bi->opt);
    if (opt) bi->opt = *opt;
    auto t0 = std::chrono::steady_clock::now();
    bi->prep.resize(n);
    bi->offs.resize(n);
    std::vector<std::vector<uint8_t>> images(n);
    // host preparation in parallel (each model independent)
    int threads = bi->opt.host_threads > 0 ? bi->opt.host_threads : (int)std::thread::hardware_concurrency();
    threads = std::max(1, std::min({threads, 16, n}));
    std::atomic<int> next(0);
    std::vector<std::string> errors(threads);
    std::vector<int> codes(threads, CRUSH_OK);
    std::vector<std::thread> pool;
    for (int w = 0; w < threads; ++w)
        pool.emplace_back([&, w] {
            for (int i = next++; i < n; i = next++) {
                if (codes[w] != CRUSH_OK) break;
                try {
                    if (!models[i]) fail(CRUSH_E_INPUT, "model %d is null", i);
                    prepare(*models[i], bi->prep[i]);
                    bi->offs[i] = build_image(bi->prep[i], images[i]);
                } catch (const Error& e) {
                    codes[w] = e.code;
                    errors[w] = "model " + std::to_string(i) + ": " + e.what();
                } catch (const std::exception& e) {
                    codes[w] = CRUSH_E_INPUT;
                    errors[w] = "model " + std::to_string(i) + ": " + e.what();
                }
            }
        });
    for (auto& th : pool) th.join();
    for (int w = 0; w < threads; ++w)
        if (codes[w] != CRUSH_OK) fail(codes[w], "%s", errors[w].c_str());
    auto t1 = std::chrono::steady_clock::now();
    bi->prepare_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    // pool layout
    bi->base_off.resize(n);
    size_t total = 0;
    for (int i = 0; i < n; ++i) {
        bi->base_off[i] = total;
        total += (bi->offs[i].total + 255) & ~size_t(255);
    }
    HIP_OK(hipStreamCreateWithFlags(&bi->stream, hipStreamNonBlocking));
    bi->pool.alloc(total);
    std::vector<InstDev>& tab = bi->host_tab;
    tab.resize(n);
    const uintptr_t base = reinterpret_cast<uintptr_t>(bi->pool.p);
    for (int i = 0; i < n; ++i) {
        tab[i] = make_inst(bi->prep[i], bi->offs[i], base + bi->base_off[i]);
        tab[i].cell_scale = cell_scale_of(bi->opt);
        HIP_OK(hipMemcpyAsync(bi->pool.p + bi->base_off[i], images[i].data(), images[i].size(), hipMemcpyHostToDevice,
                              bi->stream));
    }
    bi->table.upload(tab, bi->stream);
    bi->d_counts.alloc(N_LISTS);
    bi->n_done.alloc(1);
    bi->n_done.zero(bi->stream);
    bi->done_flag.alloc(n);
    bi->done_flag.zero(bi->stream);
    bi->host_done.assign(n, 0);
    bi->summary.alloc(n);
    bi->t_start.alloc(1);
    {
        int dev = 0, khz = 0;
        HIP_OK(hipGetDevice(&dev));
        HIP_OK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        bi->clock_khz = khz > 0 ? (double)khz : 1.0;
    }
    HIP_OK(hipStreamSynchronize(bi->stream));
    bi->upload_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
    rebuild_lists(*bi);
    capture(*bi);
    bi->fresh = true;
    if (bi->opt.autotune) autotune(*bi);
    return bi.release();
}

void batch_reset(BatchImpl& bi) { reset_state(bi, false); }

// Milliseconds from the start of the last crush_batch_run to the retirement
// of instance i (device wall clock).
static double retire_ms(const BatchImpl& bi, int i) {
    const unsigned long long tr = bi.host_summary[i].t_retire;
    return tr > bi.host_t_start ? (double)(tr - bi.host_t_start) / bi.clock_khz : 0.0;
}

// Replays the step graph until every instance has retired (or, when
// max_replays > 0, at most that many replays).  Between replays the active
// lists drop retired instances once a quarter of the listed ones have
// retired, and the graph is captured again with smaller grids once the
// lists have shrunk to half of what it was captured for.
static void drive(BatchImpl& bi, long long max_replays) {
    const int n = (int)bi.prep.size();
    int done = 0;
    long long guard = 0;
    while (done < n) {
        replay(bi);
        bi.launches += steps_per_graph(bi.opt);
        bi.replays += 1;
        HIP_OK(hipMemcpyAsync(&done, bi.n_done.p, sizeof(int), hipMemcpyDeviceToHost, bi.stream));
        HIP_OK(hipStreamSynchronize(bi.stream));
        read_kernel_times(bi);
        if (max_replays > 0 && bi.replays >= max_replays) break;
        if (++guard > 100000000LL) fail(CRUSH_E_INPUT, "batch did not finish");
        const int active = n - done;
        const int listed = bi.cnt[L_INST];
        if (active > 0 && 4 * (listed - active) >= listed) {
            HIP_OK(hipMemcpyAsync(bi.host_done.data(), bi.done_flag.p, n, hipMemcpyDeviceToHost, bi.stream));
            HIP_OK(hipStreamSynchronize(bi.stream));
            rebuild_lists(bi);
            if (bi.need_capture || 2 * bi.cnt[L_INST] <= bi.cap[L_INST]) capture(bi);
        }
    }
}

void batch_run(BatchImpl& bi, crush_batch_stats* st) {
    const int n = (int)bi.prep.size();
    auto t0 = std::chrono::steady_clock::now();
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, bi.stream));
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(1), 0, bi.stream, bi.t_start.p);
    bi.launches = 0;
    bi.replays = 0;
    for (double& x : bi.t_kernel) x = 0.0;
    drive(bi, 0);
    HIP_OK(hipEventRecord(e1, bi.stream));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    collect(bi, n);
    HIP_OK(hipMemcpyAsync(&bi.host_t_start, bi.t_start.p, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          bi.stream));
    HIP_OK(hipStreamSynchronize(bi.stream));
    bi.summary_valid = true;
    if (st) {
        st->prepare_ms = bi.prepare_ms;
        st->upload_ms = bi.upload_ms;
        st->device_ms = ms;
        st->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        st->instances = n;
        long long es = 0, hc = 0;
        for (int i = 0; i < n; ++i) {
            es += (long long)bi.prep[i].n_elems * bi.host_summary[i].steps;
            hc += (long long)bi.prep[i].n_hc * bi.host_summary[i].steps;
        }
        st->element_steps = es;
        st->hc_element_steps = hc;
        st->launches = bi.launches;
        st->graph_replays = bi.replays;
        st->pool_bytes = (double)bi.pool.n;
        std::vector<double> lat(n);
        for (int i = 0; i < n; ++i) lat[i] = retire_ms(bi, i);
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double q) { return lat[std::min<size_t>(n - 1, (size_t)std::ceil(q * n) - 1)]; };
        st->latency_p50_ms = pct(0.50);
        st->latency_p99_ms = pct(0.99);
        const double* k = bi.t_kernel;
        st->k_end_ms = k[K_END];
        st->k_contact_ms = k[K_GRID] + k[K_SEARCH] + k[K_FACEFORCE];
        st->k_node_ms = k[K_NODE];
        st->k_force_hc_ms = k[K_FORCE_HC];
        st->k_force_el_ms = k[K_FORCE_EL];
        for (int j = 0; j < N_KERNELS && j < 12; ++j) st->k_ms[j] = k[j];
    }
    for (int i = 0; i < n; ++i)
        if (bi.host_summary[i].error)
            fail(CRUSH_E_NONFINITE, "instance %d: displacement became non-finite at step %lld", i,
                 bi.host_summary[i].steps);
}

// Copy the host table into the device table (same buffer: captured
// graphs stay valid).
static void put_table(BatchImpl& bi) {
    HIP_OK(hipMemcpyAsync(bi.table.p, bi.host_tab.data(), bi.host_tab.size() * sizeof(InstDev), hipMemcpyHostToDevice,
                          bi.stream));
    HIP_OK(hipStreamSynchronize(bi.stream));
}

// Milliseconds of `replays` graph replays from the initial state of the
// instances `only` marks (all when null), after one replay to warm up.
static double trial_ms(BatchImpl& bi, const std::vector<uint8_t>* only, long long replays) {
    reset_state(bi, true, only);
    bi.replays = 0;
    drive(bi, 1);
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, bi.stream));
    bi.replays = 0;
    drive(bi, replays);
    HIP_OK(hipEventRecord(e1, bi.stream));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

// Timed trials over the knobs that do not change results (the reference
// fixes its own, src/kernel_simd.rs:562).  (1) The launch shapes, over the
// whole batch: workgroup size of the per-instance kernel and force tile.
// (2) Per size class (instances in up to four classes of equal count by
// element count), the contact grids' cell scale, each class timed alone.
// Each candidate runs a few graph replays from the initial state; the
// fastest is kept and the state restored.
static void autotune(BatchImpl& bi) {
    const int n = (int)bi.prep.size();
    const crush_batch_options base = bi.opt;
    const long long trial_replays = 3;
    const int shapes[][2] = {{256, 256}, {128, 256}, {512, 256}, {256, 128}, {256, 64}};
    double best = INFINITY;
    crush_batch_options best_opt = base;
    for (const auto& sh : shapes) {
        bi.opt = base;
        bi.opt.threads_per_instance = sh[0];
        bi.opt.force_tile = sh[1];
        const double ms = trial_ms(bi, nullptr, trial_replays);
        if (ms < best) {
            best = ms;
            best_opt = bi.opt;
        }
    }
    bi.opt = best_opt;
    bi.opt.autotune = 0;
    // size classes
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return bi.prep[a].n_elems < bi.prep[b].n_elems; });
    const int classes = n >= 8 ? 4 : 1;
    bi.class_scale.assign(classes, cell_scale_of(bi.opt));
    bi.class_elems.assign(classes, 0);
    const double scales[] = {0.5, 1.0, 2.0};
    for (int k = 0; k < classes; ++k) {
        std::vector<uint8_t> only(n, 0);
        const int lo = (int)((long long)n * k / classes), hi = (int)((long long)n * (k + 1) / classes);
        for (int j = lo; j < hi; ++j) only[order[j]] = 1;
        bi.class_elems[k] = bi.prep[order[hi - 1]].n_elems;
        double best_k = INFINITY;
        for (double cs : scales) {
            for (int j = lo; j < hi; ++j) bi.host_tab[order[j]].cell_scale = cs;
            put_table(bi);
            const double ms = trial_ms(bi, &only, trial_replays);
            if (ms < best_k) {
                best_k = ms;
                bi.class_scale[k] = cs;
            }
        }
        for (int j = lo; j < hi; ++j) bi.host_tab[order[j]].cell_scale = bi.class_scale[k];
    }
    put_table(bi);
    reset_state(bi, true);
}

void batch_options_of(const BatchImpl& bi, crush_batch_options* o) { *o = bi.opt; }

void batch_time_kernels(BatchImpl& bi, bool on) { bi.time_kernels = on; }

void batch_result(BatchImpl& bi, int i, crush_results* out) {
    if (i < 0 || i >= (int)bi.prep.size()) fail(CRUSH_E_INPUT, "instance %d out of range", i);
    if (!bi.summary_valid) fail(CRUSH_E_INPUT, "batch has not been run");
    const Prepared& P = bi.prep[i];
    const BatchSummary& r = bi.host_summary[i];
    const ImageOffsets& o = bi.offs[i];
    uint8_t* base = bi.pool.p + bi.base_off[i];
    out->time = r.t;
    out->steps = r.steps;
    out->eroded_elements = r.eroded;
    out->added_mass_max = r.added_max;
    out->separated = r.separated;
    out->wall_ms = 0.0;
    out->device_ms = retire_ms(bi, i);
    if (out->displacement)
        HIP_OK(hipMemcpy(out->displacement, base + o.u, sizeof(double) * 3 * P.n_nodes, hipMemcpyDeviceToHost));
    if (out->velocity)
        HIP_OK(hipMemcpy(out->velocity, base + o.v, sizeof(double) * 3 * P.n_nodes, hipMemcpyDeviceToHost));
    if (out->masses) std::memcpy(out->masses, P.mass.data(), sizeof(double) * P.n_nodes);
    out->n_groups = P.n_groups;
    if (out->groups)
        for (int g = 0; g < P.n_groups; ++g) {
            crush_group_result& gr = out->groups[g];
            gr.mass = P.g_mass[g];
            for (int d = 0; d < 3; ++d) gr.delta_v[d] = g < 8 ? r.dv[g][d] : 0.0;
            gr.crush_end = g < 8 ? r.crush_end[g] : 0.0;
            gr.crush_max = g < 8 ? r.crush_max[g] : 0.0;
        }
    pulse_results(P, reinterpret_cast<const double*>(base + o.pulse), r.pulse_samples, r.t, out);
    const std::vector<double> h =
        history_download(P, reinterpret_cast<const double*>(base + o.hist), r.hist_samples, bi.stream);
    history_results(P, h.data(), r.hist_samples, out);
    out->frame_samples = 0;
}

void batch_free(BatchImpl* bi) { delete bi; }

}  // namespace crush

// ------------------------------------------------------------------- ABI

struct crush_batch {
    crush::BatchImpl* impl;
};

using namespace crush;

extern "C" {

void crush_batch_defaults(crush_batch_options* opt) {
    if (opt) batch_defaults(opt);
}

int32_t crush_batch_create(const crush_model* const* models, int32_t n, const crush_batch_options* opt,
                           crush_batch** out) {
    return guarded([&] {
        if (!out) fail(CRUSH_E_INPUT, "null argument");
        BatchImpl* bi = batch_create(models, n, opt);
        *out = new crush_batch{bi};
    });
}

int32_t crush_batch_run(crush_batch* b, crush_batch_stats* stats) {
    return guarded([&] {
        if (!b) fail(CRUSH_E_INPUT, "null batch");
        batch_run(*b->impl, stats);
    });
}

int32_t crush_batch_options_get(const crush_batch* b, crush_batch_options* opt) {
    return guarded([&] {
        if (!b || !opt) fail(CRUSH_E_INPUT, "null argument");
        batch_options_of(*b->impl, opt);
    });
}

int32_t crush_batch_class_scales(const crush_batch* b, int32_t cap, double* scale, int32_t* max_elems,
                                 int32_t* n_classes) {
    return guarded([&] {
        if (!b || !n_classes || (cap > 0 && (!scale || !max_elems))) fail(CRUSH_E_INPUT, "null argument");
        const BatchImpl& bi = *b->impl;
        const int k = (int)bi.class_scale.size();
        for (int i = 0; i < k && i < cap; ++i) {
            scale[i] = bi.class_scale[i];
            max_elems[i] = bi.class_elems[i];
        }
        *n_classes = k;
    });
}

int32_t crush_batch_time_kernels(crush_batch* b, int32_t on) {
    return guarded([&] {
        if (!b) fail(CRUSH_E_INPUT, "null batch");
        batch_time_kernels(*b->impl, on != 0);
    });
}

int32_t crush_batch_reset(crush_batch* b) {
    return guarded([&] {
        if (!b) fail(CRUSH_E_INPUT, "null batch");
        batch_reset(*b->impl);
    });
}

int32_t crush_batch_result(crush_batch* b, int32_t i, crush_results* out) {
    return guarded([&] {
        if (!b || !out) fail(CRUSH_E_INPUT, "null argument");
        batch_result(*b->impl, i, out);
    });
}

void crush_batch_free(crush_batch* b) {
    if (!b) return;
    batch_free(b->impl);
    delete b;
}

}  // extern "C"
