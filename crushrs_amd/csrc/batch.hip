// Batched solves: thousands of independent models on one GPU.
//
// Each instance is one contiguous image in a device pool (its arrays plus a
// control block) described by an InstDev pointer table.  A persistent grid
// of workgroups pulls instance indices from a device work queue; the
// workgroup that takes an instance runs up to S explicit steps of it, every
// phase of a step (stable step, contact, half step, element forces, gather,
// end of step) as block-strided loops separated by barriers.  No phase
// depends on other instances, so an instance's results are bitwise the same
// alone or in any mix.  The launch sequence (queue reset + persistent
// kernel, R times) is captured in one hipGraph and replayed until every
// instance has retired.
#include "batch.hpp"

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

constexpr int BATCH_GRID_CELLS = 1024;
constexpr int BATCH_GRID_ENTRIES = 4096;

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

// Offsets of every array of an instance image relative to its start; the
// first `init_bytes` must be uploaded, the rest is scratch.
struct ImageOffsets {
    size_t conn, ecls, cls_mat, cgrad, cvol, cgam, chs, cdtf, emass, mats, state, eroded, extra;
    size_t X, u, v, f_int, f_ext, inv_mass, mass, added, fixed, n2e_ptr, n2e_idx;
    size_t cdesc, cn_node, cn_scale, cf_nodes, cf_contact, tn_node, tn_ptr, tn_ent, rec_face, fh_count;
    size_t g_nodes_ptr, g_nodes, g_front_ptr, g_front, g_accel, g_axis, g_mass, g_v0, g_dv, ctl;
    size_t init_bytes;
    size_t fe, fgeo, rec_f, fh_list, face_force;
    size_t total;
};

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

// Host image of one instance (initialized part only).
static ImageOffsets build_image(const Prepared& P, std::vector<uint8_t>& out) {
    ImageBuilder b;
    ImageOffsets o{};
    const int N = P.n_nodes, E = P.n_elems;
    o.conn = b.put(P.conn);
    o.ecls = b.put(P.ecls);
    o.cls_mat = b.put(P.cls_mat);
    o.cgrad = b.put(P.cls_grad);
    o.cvol = b.put(P.cls_vol);
    o.cgam = b.put(P.cls_gamma);
    o.chs = b.put(P.cls_H);
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
    {
        std::vector<double> im(N);
        for (int i = 0; i < N; ++i) im[i] = P.mass[i] > 0.0 ? 1.0 / P.mass[i] : 0.0;
        o.inv_mass = b.put(im);
    }
    o.mass = b.put(P.mass);
    o.added = b.zeros(sizeof(double) * N);
    o.fixed = b.put(P.fixed);
    o.n2e_ptr = b.put(P.n2e_ptr);
    o.n2e_idx = b.put(P.n2e_idx);
    o.cdesc = b.put(P.cdesc);
    o.cn_node = b.put(P.cn_node);
    o.cn_scale = b.put(P.cn_scale);
    o.cf_nodes = b.put(P.cf_nodes);
    o.cf_contact = b.put(P.cf_contact);
    o.tn_node = b.put(P.tn_node);
    o.tn_ptr = b.put(P.tn_ptr);
    o.tn_ent = b.put(P.tn_ent);
    o.rec_face = b.zeros(sizeof(int32_t) * std::max(P.n_cnodes, 1));
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
    o.fe = scratch(sizeof(float) * 24 * (size_t)E);
    o.fgeo = scratch(sizeof(double) * FG * (size_t)P.n_cfaces);
    o.rec_f = scratch(sizeof(double) * 3 * (size_t)P.n_cnodes);
    o.fh_list = scratch(sizeof(int32_t) * FACE_HITS_CAP * (size_t)P.n_cfaces);
    o.face_force = scratch(sizeof(double) * 3 * (size_t)P.n_cfaces);
    o.total = off;
    out.swap(b.bytes);
    return o;
}

static InstDev make_inst(const Prepared& P, const ImageOffsets& o, uintptr_t base) {
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
    I.cgrad = at<float>(base, o.cgrad);
    I.cvol = at<float>(base, o.cvol);
    I.cgam = at<float>(base, o.cgam);
    I.chs = at<float>(base, o.chs);
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
    I.tn_node = at<int32_t>(base, o.tn_node);
    I.tn_ptr = at<int32_t>(base, o.tn_ptr);
    I.tn_ent = at<int32_t>(base, o.tn_ent);
    I.sblk_contact = nullptr;
    I.sblk_start = nullptr;
    I.fgeo = at<double>(base, o.fgeo);
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
    I.ctl = at<StepCtl>(base, o.ctl);
    return I;
}

// ------------------------------------------------------------- kernel

// Phases of one step for one instance, block-strided, each ending with a
// barrier.
template <int B>
__device__ __forceinline__ void ph_dt(const InstDev* Dp, StepCtl* c, double* s_w) {
    const InstDev& D = *Dp;
    const int t = threadIdx.x;
    double mn = INFINITY, sm = 0.0;
    const double floor_ = c->dt_floor;
    for (int s = t; s < D.E; s += B) {
        double dte, ex;
        dt_element(D, s, floor_, &dte, &ex);
        mn = fmin(mn, dte);
        sm += ex;
    }
    mn = block_reduce<B>(mn, 1, s_w);
    sm = block_reduce<B>(sm, 0, s_w);
    if (isfinite(mn)) {
        double dt_use = mn;
        const bool scale = floor_ > 0.0 && mn < floor_;
        if (scale) {
            dt_use = floor_;
            for (int n = t; n < D.N; n += B) mass_node(D, n);
        }
        const double cfl = c->dt_scale * fmin(dt_use, c->contact_dt);
        __syncthreads();
        if (t == 0) {
            if (scale) c->added_total_max = fmax(c->added_total_max, sm);
            c->dt_cfl = cfl;
        }
        __syncthreads();
    }
}

template <int B>
__device__ __forceinline__ double ph_contact(const InstDev* Dp, double dt, double cell_scale, int* s_start,
                                                        int* s_cursor, int* s_entries, double* s_w, int* s_tmp) {
    const InstDev& D = *Dp;
    const int t = threadIdx.x;
    double pair_hits = 0.0;
    for (int f = t; f < D.F; f += B) contact_face(D, f);
    __syncthreads();
    for (int ci = 0; ci < D.n_contacts; ++ci) {
        const ContactDesc cd = D.cdesc[ci];
        const FaceGrid g = build_face_grid<B, BATCH_GRID_CELLS, BATCH_GRID_ENTRIES>(D, cd, cell_scale, s_start, s_cursor,
                                                                                     s_entries, s_w, s_tmp);
        for (int q = t; q < cd.n_nodes; q += B)
            if (contact_node(D, cd, q, g, s_start, s_entries, dt) && cd.pair) pair_hits += 1.0;
        __syncthreads();
    }
    for (int f = t; f < D.F; f += B) contact_face_force(D, f);
    __syncthreads();
    for (int i = t; i < D.T; i += B) contact_apply_node(D, i);
    __syncthreads();
    return pair_hits;
}

template <int B>
__device__ __forceinline__ void ph_half(const InstDev* Dp, double dt) {
    const InstDev& D = *Dp;
    for (int n = threadIdx.x; n < D.N; n += B) half_node(D, n, dt);
    __syncthreads();
}

template <int B>
__device__ __forceinline__ double ph_force(const InstDev* Dp) {
    const InstDev& D = *Dp;
    double eroded = 0.0;
    for (int s = threadIdx.x; s < D.n_hc; s += B) eroded += force_element<CRUSH_MAT_HONEYCOMB>(D, s);
    for (int s = D.n_hc + threadIdx.x; s < D.E; s += B) eroded += force_element<CRUSH_MAT_ELASTIC>(D, s);
    __syncthreads();
    return eroded;
}

template <int B>
__device__ __forceinline__ double ph_gather(const InstDev* Dp, double dt, double damp) {
    const InstDev& D = *Dp;
    double bad = 0.0;
    for (int n = threadIdx.x; n < D.N; n += B)
        if (!gather_node(D, n, dt, damp)) bad = 1.0;
    __syncthreads();
    return bad;
}

// Up to S steps of one instance by the whole block (the loop body of
// src/solver.rs:287-352).  Returns true when the instance has retired.
template <int B>
__device__ bool run_instance(const InstDev* Dp, int S, double cell_scale, int* s_start, int* s_cursor, int* s_entries,
                             double* s_w, int* s_tmp, StepCtl* c) {
    const InstDev& D = *Dp;
    const int t = threadIdx.x;
    if (t == 0) *c = *D.ctl;
    __syncthreads();
    if (c->done) return false;
    for (int k = 0; k < S && !c->done; ++k) {
        if (c->adaptive && c->step % c->chk == 0) ph_dt<B>(Dp, c, s_w);
        const double dt = step_dt(*c);
        double pair_hits = 0.0;
        if (D.F > 0) pair_hits = ph_contact<B>(Dp, dt, cell_scale, s_start, s_cursor, s_entries, s_w, s_tmp);
        ph_half<B>(Dp, dt);
        double eroded = ph_force<B>(Dp);
        double bad = ph_gather<B>(Dp, dt, c->damping);
        pair_hits = block_reduce<B>(pair_hits, 0, s_w);
        eroded = block_reduce<B>(eroded, 0, s_w);
        bad = block_reduce<B>(bad, 2, s_w);
        if (t == 0) c->eroded_total += (unsigned long long)eroded;
        step_end<B>(D, c, dt, (int)pair_hits, bad > 0.0, s_w);
    }
    const bool done = c->done != 0;
    if (done) {
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
    __syncthreads();
    if (t == 0) *D.ctl = *c;
    __syncthreads();
    return done;
}

template <int B>
__global__ __launch_bounds__(B) void k_batch(const InstDev* __restrict__ table, const int32_t* __restrict__ order, int n,
                                             int S, double cell_scale, int* __restrict__ counter,
                                             int* __restrict__ n_done) {
    __shared__ int s_start[BATCH_GRID_CELLS + 1];
    __shared__ int s_cursor[BATCH_GRID_CELLS];
    __shared__ int s_entries[BATCH_GRID_ENTRIES];
    __shared__ double s_w[B / 64];
    __shared__ int s_tmp[B / 64];
    __shared__ int s_next;
    __shared__ StepCtl s_ctl;
    __shared__ InstDev s_inst;
    for (;;) {
        if (threadIdx.x == 0) s_next = atomicAdd(counter, 1);
        __syncthreads();
        const int i = s_next;
        if (i >= n) return;
        // the instance's pointer table lives in LDS and is read where used
        constexpr int W = (int)(sizeof(InstDev) / sizeof(int));
        const int idx = order[i];
        for (int w = threadIdx.x; w < W; w += B)
            reinterpret_cast<int*>(&s_inst)[w] = reinterpret_cast<const int*>(&table[idx])[w];
        __syncthreads();
        if (run_instance<B>(&s_inst, S, cell_scale, s_start, s_cursor, s_entries, s_w, s_tmp, &s_ctl) && threadIdx.x == 0)
            atomicAdd(n_done, 1);
        __syncthreads();
    }
}

__global__ void k_queue_reset(int* counter) { *counter = 0; }

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
    for (int g = 0; g < 8; ++g) {
        r.crush_max[g] = g < D.n_groups ? c.crush_max[g] : 0.0;
        r.crush_end[g] = g < D.n_groups ? c.crush_end[g] : 0.0;
        for (int d = 0; d < 3; ++d) r.dv[g][d] = g < D.n_groups ? D.g_dv[3 * g + d] : 0.0;
    }
}

// ------------------------------------------------------------- host

struct BatchImpl {
    crush_batch_options opt{};
    std::vector<Prepared> prep;
    std::vector<ImageOffsets> offs;
    std::vector<size_t> base_off;
    std::vector<int32_t> order;
    DevBuf<uint8_t> pool;
    DevBuf<InstDev> table;
    DevBuf<int32_t> d_order, counter, n_done;
    DevBuf<BatchSummary> summary;
    std::vector<BatchSummary> host_summary;
    bool summary_valid = false;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int grid = 0;
    double prepare_ms = 0, upload_ms = 0;
    long long launches = 0, replays = 0;

    ~BatchImpl() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static int block_of(const crush_batch_options& o) {
    int b = o.threads_per_instance;
    return (b == 512 || b == 1024) ? b : 256;
}

template <int B>
static void launch_batch(BatchImpl& bi, int n) {
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(1), 0, bi.stream, bi.counter.p);
    const int S = bi.opt.steps_per_launch > 0 ? bi.opt.steps_per_launch : 0x7fffffff;
    const double cs = bi.opt.contact_cell_scale > 0.0 ? bi.opt.contact_cell_scale : 1.0;
    hipLaunchKernelGGL(k_batch<B>, dim3(bi.grid), dim3(B), 0, bi.stream, bi.table.p, bi.d_order.p, n, S, cs,
                       bi.counter.p, bi.n_done.p);
}

template <int B>
static int occupancy_grid() {
    int dev = 0, cus = 0, per_cu = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_batch<B>), B, 0));
    return std::max(1, cus * std::max(per_cu, 1));
}

void batch_defaults(crush_batch_options* o) {
    o->threads_per_instance = 256;
    o->steps_per_launch = 0;
    o->launches_per_graph = 1;
    o->host_threads = 0;
    o->contact_cell_scale = 1.0;
}

BatchImpl* batch_create(const crush_model* const* models, int n, const crush_batch_options* opt) {
    require_device();
    if (n <= 0 || !models) fail(CRUSH_E_INPUT, "batch needs at least one model");
    auto bi = std::make_unique<BatchImpl>();
    batch_defaults(&bi->opt);
    if (opt) bi->opt = *opt;
    if (bi->opt.launches_per_graph < 1) bi->opt.launches_per_graph = 1;
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
    std::vector<InstDev> tab(n);
    const uintptr_t base = reinterpret_cast<uintptr_t>(bi->pool.p);
    for (int i = 0; i < n; ++i) {
        tab[i] = make_inst(bi->prep[i], bi->offs[i], base + bi->base_off[i]);
        HIP_OK(hipMemcpyAsync(bi->pool.p + bi->base_off[i], images[i].data(), images[i].size(), hipMemcpyHostToDevice,
                              bi->stream));
    }
    bi->table.upload(tab, bi->stream);
    // largest instances first (longest-processing-time order)
    bi->order.resize(n);
    std::iota(bi->order.begin(), bi->order.end(), 0);
    std::vector<double> cost(n);
    for (int i = 0; i < n; ++i)
        cost[i] = (double)bi->prep[i].n_elems * bi->prep[i].settings.end_time / std::max(bi->prep[i].dt0, 1e-12);
    std::stable_sort(bi->order.begin(), bi->order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    bi->d_order.upload(bi->order, bi->stream);
    bi->counter.alloc(1);
    bi->n_done.alloc(1);
    bi->n_done.zero(bi->stream);
    bi->summary.alloc(n);
    HIP_OK(hipStreamSynchronize(bi->stream));
    bi->upload_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
    const int B = block_of(bi->opt);
    bi->grid = B == 1024 ? occupancy_grid<1024>() : (B == 512 ? occupancy_grid<512>() : occupancy_grid<256>());
    // capture the launch sequence
    hipGraph_t g;
    HIP_OK(hipStreamBeginCapture(bi->stream, hipStreamCaptureModeThreadLocal));
    for (int r = 0; r < bi->opt.launches_per_graph; ++r) {
        if (B == 1024)
            launch_batch<1024>(*bi, n);
        else if (B == 512)
            launch_batch<512>(*bi, n);
        else
            launch_batch<256>(*bi, n);
    }
    HIP_OK(hipStreamEndCapture(bi->stream, &g));
    HIP_OK(hipGraphInstantiate(&bi->graph, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    return bi.release();
}

void batch_reset(BatchImpl& bi) {
    const int n = (int)bi.prep.size();
    for (int i = 0; i < n; ++i) {
        std::vector<uint8_t> img;
        build_image(bi.prep[i], img);
        HIP_OK(hipMemcpyAsync(bi.pool.p + bi.base_off[i], img.data(), img.size(), hipMemcpyHostToDevice, bi.stream));
        HIP_OK(hipStreamSynchronize(bi.stream));
    }
    bi.n_done.zero(bi.stream);
    HIP_OK(hipStreamSynchronize(bi.stream));
    bi.summary_valid = false;
}

void batch_run(BatchImpl& bi, crush_batch_stats* st) {
    const int n = (int)bi.prep.size();
    auto t0 = std::chrono::steady_clock::now();
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, bi.stream));
    int done = 0;
    long long guard = 0;
    while (done < n) {
        HIP_OK(hipGraphLaunch(bi.graph, bi.stream));
        bi.launches += bi.opt.launches_per_graph;
        bi.replays += 1;
        HIP_OK(hipMemcpyAsync(&done, bi.n_done.p, sizeof(int), hipMemcpyDeviceToHost, bi.stream));
        HIP_OK(hipStreamSynchronize(bi.stream));
        if (++guard > 100000000LL) fail(CRUSH_E_INPUT, "batch did not finish");
    }
    HIP_OK(hipEventRecord(e1, bi.stream));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    hipLaunchKernelGGL(k_collect, dim3((n + 255) / 256), dim3(256), 0, bi.stream, bi.table.p, n, bi.summary.p);
    bi.host_summary.resize(n);
    HIP_OK(hipMemcpyAsync(bi.host_summary.data(), bi.summary.p, n * sizeof(BatchSummary), hipMemcpyDeviceToHost,
                          bi.stream));
    HIP_OK(hipStreamSynchronize(bi.stream));
    bi.summary_valid = true;
    if (st) {
        st->prepare_ms = bi.prepare_ms;
        st->upload_ms = bi.upload_ms;
        st->device_ms = ms;
        st->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        st->instances = n;
        long long es = 0;
        for (int i = 0; i < n; ++i) es += (long long)bi.prep[i].n_elems * bi.host_summary[i].steps;
        st->element_steps = es;
        st->launches = bi.launches;
        st->graph_replays = bi.replays;
        st->pool_bytes = (double)bi.pool.n;
    }
    for (int i = 0; i < n; ++i)
        if (bi.host_summary[i].error)
            fail(CRUSH_E_NONFINITE, "instance %d: displacement became non-finite at step %lld", i,
                 bi.host_summary[i].steps);
}

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
    out->device_ms = 0.0;
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
