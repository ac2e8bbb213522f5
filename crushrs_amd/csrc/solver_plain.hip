// One solve at a time on the GPU: the reference's central-difference loop
// (src/solver.rs:287-352) as a sequence of kernels per step, one model per
// HIP stream.  Nodal state is f64, element arithmetic f32, contact and the
// stable-step pass f64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "device_common.hpp"
#include "hip_util.hpp"
#include "outputs.hpp"
#include "phases.hpp"
#include "prep.hpp"
#include "signal.hpp"
#include "solver_plain.hpp"

namespace crush {

struct DeviceModel {
    const Prepared& P;
    hipStream_t stream;
    int dt_blocks = 0;
    DevBuf<int32_t> conn, ecls, cls_mat, n2e_ptr, n2e_idx;
    DevBuf<float> cls_pack, state, fe;
    DevBuf<double> cls_dtf, elem_mass, extra, X, u, v, f_int, f_ext, inv_mass, mass, added, part_min, part_sum;
    DevBuf<uint8_t> eroded, fixed, cf_static;
    DevBuf<MatParams> mats;
    DevBuf<ContactDesc> cdesc;
    DevBuf<int32_t> cn_node, cn_contact, cf_nodes, cf_contact, tn_node, tn_ptr, tn_ent, sblk_contact, sblk_start;
    DevBuf<double> cn_scale, fgeo, rec_f, face_force;
    DevBuf<float4> fbox;
    DevBuf<int32_t> rec_face, fh_count, fh_list;
    DevBuf<int32_t> g_nodes_ptr, g_nodes, g_front_ptr, g_front, g_accel, g_axis;
    DevBuf<double> g_mass, g_v0, g_dv, pulse, hist;
    DevBuf<int32_t> g_frame, p_nodes_ptr, p_nodes;
    DevBuf<StepCtl> ctl;
    InstDev inst{};

    DeviceModel(const Prepared& p, hipStream_t s) : P(p), stream(s) {
        const int E = P.n_elems, N = P.n_nodes;
        conn.upload(P.conn, s);
        ecls.upload(P.ecls, s);
        cls_mat.upload(P.cls_mat, s);
        n2e_ptr.upload(P.n2e_ptr, s);
        n2e_idx.upload(P.n2e_idx, s);
        cls_pack.upload(P.cls_pack, s);
        cls_dtf.upload(P.cls_dtf, s);
        elem_mass.upload(P.elem_mass, s);
        mats.upload(P.mats, s);
        // honeycomb state: H_prev = 0, R = I, sigma = 0, c = 0
        std::vector<float> st((size_t)25 * std::max(P.n_hc, 1), 0.0f);
        for (int k : {9, 13, 17})
            for (int i = 0; i < P.n_hc; ++i) st[(size_t)k * P.n_hc + i] = 1.0f;
        state.upload(st, s);
        fe.alloc(fe_floats(E));
        extra.alloc(E);
        eroded.alloc(E);
        eroded.zero(s);
        X.upload(P.X, s);
        u.alloc(3 * (size_t)N);
        u.zero(s);
        v.upload(P.v0, s);
        f_int.alloc(3 * (size_t)N);
        f_int.zero(s);
        f_ext.alloc(3 * (size_t)N);
        f_ext.zero(s);
        mass.upload(P.mass, s);
        std::vector<double> im(N);
        for (int i = 0; i < N; ++i) im[i] = P.mass[i] > 0.0 ? 1.0 / P.mass[i] : 0.0;
        inv_mass.upload(im, s);
        added.alloc(N);
        added.zero(s);
        fixed.upload(P.fixed, s);
        dt_blocks = (E + 255) / 256;
        part_min.alloc(dt_blocks);
        part_sum.alloc(dt_blocks);
        cdesc.upload(P.cdesc, s);
        cn_node.upload(P.cn_node, s);
        cn_contact.upload(P.cn_contact, s);
        cn_scale.upload(P.cn_scale, s);
        cf_nodes.upload(P.cf_nodes, s);
        cf_contact.upload(P.cf_contact, s);
        cf_static.upload(P.cf_static, s);
        tn_node.upload(P.tn_node, s);
        tn_ptr.upload(P.tn_ptr, s);
        tn_ent.upload(P.tn_ent, s);
        sblk_contact.upload(P.sblk_contact, s);
        sblk_start.upload(P.sblk_start, s);
        fgeo.alloc((size_t)FG * P.n_cfaces);
        fbox.alloc(2 * (size_t)P.n_cfaces);
        rec_f.alloc(3 * (size_t)P.n_cnodes);
        rec_face.alloc(P.n_cnodes);
        face_force.alloc(3 * (size_t)P.n_cfaces);
        fh_count.alloc(P.n_cfaces);
        fh_count.zero(s);
        fh_list.alloc((size_t)FACE_HITS_CAP * P.n_cfaces);
        g_nodes_ptr.upload(P.g_nodes_ptr, s);
        g_nodes.upload(P.g_nodes, s);
        g_front_ptr.upload(P.g_front_ptr, s);
        g_front.upload(P.g_front, s);
        g_accel.upload(P.g_accel, s);
        g_axis.upload(P.g_axis, s);
        g_mass.upload(P.g_mass, s);
        g_v0.upload(P.g_v0, s);
        g_dv.alloc(std::max<size_t>(3 * (size_t)P.n_groups, 1));
        g_frame.upload(P.g_frame, s);
        // written in order from sample 0 as the steps end
        pulse.alloc((size_t)std::min(P.n_groups, 8) * (size_t)P.n_pulse);
        p_nodes_ptr.upload(P.p_nodes_ptr, s);
        p_nodes.upload(P.p_nodes, s);
        hist.alloc(hist_record_len(P.n_parts, P.n_groups) * (size_t)P.n_hist);
        StepCtl c;
        std::memset(&c, 0, sizeof c);
        c.t = 0.0;
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
        ctl.alloc(1);
        HIP_OK(hipMemcpyAsync(ctl.p, &c, sizeof c, hipMemcpyHostToDevice, s));
        InstDev& I = inst;
        I.N = N;
        I.E = E;
        I.n_hc = P.n_hc;
        I.C = P.n_classes;
        I.F = P.n_cfaces;
        I.K = P.n_cnodes;
        I.T = (int)P.tn_node.size();
        I.n_contacts = (int)P.cdesc.size();
        I.n_groups = P.n_groups;
        I.n_sblk = (int)P.sblk_contact.size();
        I.conn = conn.p;
        I.ecls = ecls.p;
        I.cls_mat = cls_mat.p;
        I.cpk = cls_pack.p;
        I.cdtf = cls_dtf.p;
        I.emass = elem_mass.p;
        I.mats = mats.p;
        I.state = state.p;
        I.eroded = eroded.p;
        I.fe = fe.p;
        I.extra = extra.p;
        I.X = X.p;
        I.u = u.p;
        I.v = v.p;
        I.f_int = f_int.p;
        I.f_ext = f_ext.p;
        I.f_ext2 = f_ext.p;
        I.inv_mass = inv_mass.p;
        I.mass = mass.p;
        I.added = added.p;
        I.fixed = fixed.p;
        I.n2e_ptr = n2e_ptr.p;
        I.n2e_idx = n2e_idx.p;
        I.cdesc = cdesc.p;
        I.cn_node = cn_node.p;
        I.cn_scale = cn_scale.p;
        I.cf_nodes = cf_nodes.p;
        I.cf_contact = cf_contact.p;
        I.cf_static = cf_static.p;
        I.tn_node = tn_node.p;
        I.tn_ptr = tn_ptr.p;
        I.tn_ent = tn_ent.p;
        I.sblk_contact = sblk_contact.p;
        I.sblk_start = sblk_start.p;
        I.fgeo = fgeo.p;
        I.fbox = fbox.p;
        I.rec_face = rec_face.p;
        I.rec_f = rec_f.p;
        I.fh_count = fh_count.p;
        I.fh_list = fh_list.p;
        I.face_force = face_force.p;
        I.g_nodes_ptr = g_nodes_ptr.p;
        I.g_nodes = g_nodes.p;
        I.g_front_ptr = g_front_ptr.p;
        I.g_front = g_front.p;
        I.g_accel = g_accel.p;
        I.g_axis = g_axis.p;
        I.g_mass = g_mass.p;
        I.g_v0 = g_v0.p;
        I.g_dv = g_dv.p;
        I.g_frame = g_frame.p;
        I.pulse = pulse.p;
        I.n_pulse = P.n_pulse;
        I.pulse_dt = P.settings.pulse_dt;
        I.p_nodes_ptr = p_nodes_ptr.p;
        I.p_nodes = p_nodes.p;
        I.hist = hist.p;
        I.n_parts = P.n_parts;
        I.n_hist = P.n_hist;
        I.hist_steps = std::max(P.settings.history_steps, 1);
        I.ctl = ctl.p;
    }
};

// ------------------------------------------------------------- kernels

// Per-element stable step and scaling mass; block min / sum partials.
__global__ __launch_bounds__(256) void k_dt_elem(InstDev D, double* __restrict__ part_min,
                                                 double* __restrict__ part_sum) {
    __shared__ double s_w[4];
    const StepCtl* ctl = D.ctl;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double dte = INFINITY, ex = 0.0;
    if (s < D.E && !ctl->done) dt_element(D, s, ctl->dt_floor, &dte, &ex);
    const double mn = block_reduce<256>(dte, 1, s_w);
    const double sm = block_reduce<256>(ex, 0, s_w);
    if (threadIdx.x == 0) {
        part_min[blockIdx.x] = mn;
        part_sum[blockIdx.x] = sm;
    }
}

// The adaptive step decision of src/solver.rs:288-299.
__global__ __launch_bounds__(256) void k_dt_reduce(int nb, const double* __restrict__ part_min,
                                                   const double* __restrict__ part_sum, StepCtl* __restrict__ ctl) {
    __shared__ double s_w[4];
    double mn = INFINITY, sm = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) {
        mn = fmin(mn, part_min[i]);
        sm += part_sum[i];
    }
    mn = block_reduce<256>(mn, 1, s_w);
    sm = block_reduce<256>(sm, 0, s_w);
    if (threadIdx.x == 0) {
        ctl->scale_now = 0;
        if (ctl->done) return;
        if (isfinite(mn)) {
            double dt_use = mn;
            if (ctl->dt_floor > 0.0 && mn < ctl->dt_floor) {
                ctl->scale_now = 1;
                ctl->added_total_max = fmax(ctl->added_total_max, sm);
                dt_use = ctl->dt_floor;
            }
            ctl->dt_cfl = ctl->dt_scale * fmin(dt_use, ctl->contact_dt);
        }
    }
}

__global__ __launch_bounds__(256) void k_mass(InstDev D) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= D.N || !D.ctl->scale_now) return;
    mass_node(D, n);
}

__global__ void k_step_begin(StepCtl* __restrict__ ctl) {
    if (ctl->done) return;
    ctl->dt = step_dt(*ctl);
    ctl->pair_hits = 0;
}

__global__ __launch_bounds__(256) void k_contact_faces(InstDev D) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= D.F || D.ctl->done) return;
    contact_face(D, f);
}

constexpr int PLAIN_GRID_CELLS = 2048;
constexpr int PLAIN_GRID_ENTRIES = 8192;

// Bucketed node-to-face search: each block builds its contact's face grid
// in LDS and searches SEARCH_BLOCK contacting nodes.
__global__ __launch_bounds__(256) void k_contact_search(InstDev D, double cell_scale) {
    __shared__ int s_start[PLAIN_GRID_CELLS + 1];
    __shared__ int s_cursor[PLAIN_GRID_CELLS];
    __shared__ int s_entries[PLAIN_GRID_ENTRIES];
    __shared__ double s_w[4];
    __shared__ int s_tmp[4];
    StepCtl* ctl = D.ctl;
    if (ctl->done) return;
    const ContactDesc cd = D.cdesc[D.sblk_contact[blockIdx.x]];
    const FaceGrid g = build_face_grid<256>(D, cd, cell_scale, PLAIN_GRID_CELLS, PLAIN_GRID_ENTRIES, s_start, s_cursor,
                                                                                   s_entries, s_w, s_tmp);
    const int k = D.sblk_start[blockIdx.x] + threadIdx.x;
    if (k >= cd.n_nodes) return;
    if (contact_node(D, cd, k, g, s_start, s_entries, ctl->dt) && cd.pair) atomicAdd(&ctl->pair_hits, 1);
}

__global__ __launch_bounds__(256) void k_contact_faceforce(InstDev D) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (D.ctl->done) return;
    contact_face_force_wave(D, f, f < D.F);
}

__global__ __launch_bounds__(256) void k_contact_apply(InstDev D) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.T || D.ctl->done) return;
    contact_apply_node(D, i, D.f_ext);
}

__global__ __launch_bounds__(256) void k_half(InstDev D) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= D.N || D.ctl->done) return;
    half_node(D, n, D.ctl->dt, D.f_ext);
}

// Element forces of slots [begin, end) (one material kind), with per-lane
// loads of the class constants (the same arithmetic as the batched tiles).
template <int KIND>
__global__ __launch_bounds__(256) void k_force(InstDev D, int begin, int end) {
    constexpr int LS = KIND == CRUSH_MAT_HONEYCOMB ? 0 : 256;
    __shared__ float s_q[12 * (LS > 0 ? LS : 1)];
    const int s = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= end || D.ctl->done) return;
    if (force_element<KIND, false, LS>(D, s, s_q + threadIdx.x)) atomicAdd(&D.ctl->eroded_total, 1ULL);
}

__global__ __launch_bounds__(256) void k_gather_update(InstDev D) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= D.N || D.ctl->done) return;
    if (!gather_node(D, n, D.ctl->dt, D.ctl->damping, D.f_ext)) atomicOr(&D.ctl->error, 1);
}

// End of step (one block).
__global__ __launch_bounds__(256) void k_step_end(InstDev D) {
    __shared__ double s_w[4];
    __shared__ StepCtl c;
    if (D.ctl->done) return;
    if (threadIdx.x == 0) c = *D.ctl;
    __syncthreads();
    step_end<256>(D, &c, c.dt, c.pair_hits, c.error != 0, s_w, D.f_ext);
    if (threadIdx.x == 0) *D.ctl = c;
}

// ------------------------------------------------------------ host loop

static unsigned nblk(long long n, int b = 256) { return (unsigned)std::max<long long>((n + b - 1) / b, 1); }

struct PlainSolver {
    const Prepared& P;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::unique_ptr<DeviceModel> D;
    long long host_step = 0;
    double cell_scale = 1.0;

    // With an arena, the device buffers come from it and the solve runs on
    // the null stream (nothing is created); otherwise on a stream of its own.
    explicit PlainSolver(const Prepared& p, DevArena* arena = nullptr) : P(p) {
        require_device();
        if (!arena) {
            HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
            own_stream = true;
        }
        ArenaScope scope(arena);
        D.reset(new DeviceModel(P, stream));
    }
    ~PlainSolver() {
        D.reset();
        if (own_stream) (void)hipStreamDestroy(stream);
    }

    // Launch one step (the kernels no-op once the solve is done).
    void launch_step() {
        DeviceModel& d = *D;
        const InstDev& I = d.inst;
        const int E = P.n_elems, N = P.n_nodes;
        if (P.adaptive && host_step % P.settings.adaptive_check_steps == 0) {
            hipLaunchKernelGGL(k_dt_elem, dim3(d.dt_blocks), dim3(256), 0, stream, I, d.part_min.p, d.part_sum.p);
            hipLaunchKernelGGL(k_dt_reduce, dim3(1), dim3(256), 0, stream, d.dt_blocks, d.part_min.p, d.part_sum.p, d.ctl.p);
            hipLaunchKernelGGL(k_mass, dim3(nblk(N)), dim3(256), 0, stream, I);
        }
        hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, stream, d.ctl.p);
        if (P.n_cfaces > 0) {
            hipLaunchKernelGGL(k_contact_faces, dim3(nblk(P.n_cfaces)), dim3(256), 0, stream, I);
            if (I.n_sblk > 0)
                hipLaunchKernelGGL(k_contact_search, dim3((unsigned)I.n_sblk), dim3(256), 0, stream, I, cell_scale);
            hipLaunchKernelGGL(k_contact_faceforce, dim3(nblk(P.n_cfaces)), dim3(256), 0, stream, I);
            if (I.T > 0) hipLaunchKernelGGL(k_contact_apply, dim3(nblk(I.T)), dim3(256), 0, stream, I);
        }
        hipLaunchKernelGGL(k_half, dim3(nblk(N)), dim3(256), 0, stream, I);
        if (P.n_hc > 0)
            hipLaunchKernelGGL(k_force<CRUSH_MAT_HONEYCOMB>, dim3(nblk(P.n_hc)), dim3(256), 0, stream, I, 0, P.n_hc);
        if (E > P.n_hc)
            hipLaunchKernelGGL(k_force<CRUSH_MAT_ELASTIC>, dim3(nblk(E - P.n_hc)), dim3(256), 0, stream, I, P.n_hc, E);
        hipLaunchKernelGGL(k_gather_update, dim3(nblk(N)), dim3(256), 0, stream, I);
        hipLaunchKernelGGL(k_step_end, dim3(1), dim3(256), 0, stream, I);
        HIP_OK(hipGetLastError());
        ++host_step;
    }

    StepCtl read_ctl() {
        StepCtl c;
        HIP_OK(hipMemcpyAsync(&c, D->ctl.p, sizeof c, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        return c;
    }
};

static void fill_results(const Prepared& P, PlainSolver& S, const StepCtl& c, crush_results* out) {
    const int N = P.n_nodes;
    std::vector<double> u(3 * (size_t)N), v(3 * (size_t)N);
    HIP_OK(hipMemcpyAsync(u.data(), S.D->u.p, u.size() * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    HIP_OK(hipMemcpyAsync(v.data(), S.D->v.p, v.size() * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    HIP_OK(hipStreamSynchronize(S.stream));
    out->time = c.t;
    out->steps = c.step;
    out->eroded_elements = c.eroded_total;
    out->added_mass_max = c.added_total_max;
    out->separated = c.separated;
    if (out->displacement) std::memcpy(out->displacement, u.data(), u.size() * sizeof(double));
    if (out->velocity) std::memcpy(out->velocity, v.data(), v.size() * sizeof(double));
    if (out->masses) std::memcpy(out->masses, P.mass.data(), P.mass.size() * sizeof(double));
    out->n_groups = P.n_groups;
    if (out->groups) {
        for (int g = 0; g < P.n_groups; ++g) {
            crush_group_result& r = out->groups[g];
            double mt = 0.0, pv[3] = {0, 0, 0};
            for (int i = P.g_nodes_ptr[g]; i < P.g_nodes_ptr[g + 1]; ++i) {
                int n = P.g_nodes[i];
                mt += P.mass[n];
                for (int d = 0; d < 3; ++d) pv[d] += P.mass[n] * v[3 * (size_t)n + d];
            }
            r.mass = mt;
            for (int d = 0; d < 3; ++d) r.delta_v[d] = mt > 0 ? pv[d] / mt - P.g_v0[3 * g + d] : 0.0;
            const int b = P.g_front_ptr[g], e = P.g_front_ptr[g + 1], ax = P.g_axis[g];
            double cr = 0.0;
            if (P.g_accel[g] >= 0 && e > b) {
                double s = 0.0;
                for (int i = b; i < e; ++i) s += u[3 * (size_t)P.g_front[i] + ax];
                cr = std::fabs(u[3 * (size_t)P.g_accel[g] + ax] - s / (double)(e - b));
            }
            r.crush_end = cr;
            r.crush_max = g < 8 ? std::max(c.crush_max[g], cr) : cr;
        }
    }
    pulse_results(P, S.D->pulse.p, c.pulse_k, c.t, out);
    const std::vector<double> h = history_download(P, S.D->hist.p, c.hist_k, S.stream);
    history_results(P, h.data(), c.hist_k, out);
}

// Guard against a solve that never reaches end_time (dt is floored at 1e-15).
static const long long kMaxSteps = 100000000LL;

// arena: device buffers from it, no events (device_ms is then the host's
// time around the step loop, which ends with a synchronisation)
static void run_plain_in(const crush_model& m, crush_results* out, DevArena* arena) {
    auto t0 = std::chrono::steady_clock::now();
    Prepared P;
    prepare(m, P);
    if (arena && P.settings.frame_steps > 0 && out && out->frame_capacity > 0)
        fail(CRUSH_E_UNSUPPORTED, "crush_run_noalloc records no frames (frame_steps > 0): frames need device "
                                  "memory beyond its scratch; use crush_run");
    PlainSolver S(P, arena);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!arena) {
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        HIP_OK(hipEventRecord(e0, S.stream));
    }
    const auto t_loop = std::chrono::steady_clock::now();
    const int chunk = 32;
    // frames: the initial state, every frame_steps steps, the last step
    const int fs = P.settings.frame_steps;
    FrameStore frames;
    if (fs > 0 && out) frames.init(P, out->frame_capacity, S.stream);
    StepCtl c;
    while (true) {
        for (int i = 0; i < chunk; ++i) {
            S.launch_step();
            if (frames.cap > 0 && S.host_step % fs == 0)
                frames.record(S.D->inst, P, (int)(S.host_step / fs), S.host_step, S.stream);
        }
        c = S.read_ctl();
        if (c.done) break;
        if (S.host_step > kMaxSteps) fail(CRUSH_E_INPUT, "no end after %lld steps (t = %.3e)", S.host_step, c.t);
    }
    long long n_frames = 0;
    if (fs > 0) {
        n_frames = 1 + c.step / fs;
        if (c.step % fs != 0) {
            if (frames.cap > 0) frames.record(S.D->inst, P, (int)std::min<long long>(n_frames, INT32_MAX), c.step, S.stream);
            ++n_frames;
        }
    }
    float ms = 0.0f;
    if (!arena) {
        HIP_OK(hipEventRecord(e1, S.stream));
        HIP_OK(hipEventSynchronize(e1));
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    } else {
        HIP_OK(hipStreamSynchronize(S.stream));
        ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_loop).count();
    }
    if (c.error) fail(CRUSH_E_NONFINITE, "displacement became non-finite at step %lld (t = %.3e)", c.step, c.t);
    if (out) {
        fill_results(P, S, c, out);
        if (fs > 0) frames.fill(P, (int)std::min<long long>(n_frames, INT32_MAX), out, S.stream);
        else out->frame_samples = 0;
        out->device_ms = ms;
        out->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
}

void run_plain(const crush_model& m, crush_results* out) { run_plain_in(m, out, nullptr); }

// The static device scratch of crush_run_noalloc: every device buffer of a
// solve is carved from it.  Each GPU the library runs on holds its own copy
// for the life of the process.  In host code the array's name is not a
// device address: its address on the current device comes from
// hipGetSymbolAddress, once per device.
constexpr size_t kNoAllocBytes = (size_t)CRUSH_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {
struct NoAllocDevice {
    uint8_t* ptr = nullptr;  // device address of `scratch` on this device
    std::mutex run;          // calls on one device take turns
};
std::mutex g_noalloc_mu;
std::map<int, std::unique_ptr<NoAllocDevice>> g_noalloc;

// The current device's scratch (its address resolved on first use).
NoAllocDevice& noalloc_device() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    std::unique_ptr<NoAllocDevice>& d = g_noalloc[dev];
    if (!d) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(CRUSH_E_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        d.reset(new NoAllocDevice);
        d->ptr = static_cast<uint8_t*>(p);
    }
    return *d;
}
}  // namespace

void run_plain_noalloc(const crush_model& m, crush_results* out) {
    require_device();
    NoAllocDevice& d = noalloc_device();
    std::lock_guard<std::mutex> lk(d.run);
    DevArena arena;
    arena.base = d.ptr;
    arena.cap = kNoAllocBytes;
    arena.owner = "crush_run_noalloc";
    run_plain_in(m, out, &arena);
}

void run_plain_taps(const crush_model& m, int n_steps, double* f_int, double* f_ext, double* u_out, int32_t* hit,
                    uint8_t* eroded,
                    double* dt, int32_t* steps_done) {
    Prepared P;
    prepare(m, P);
    PlainSolver S(P);
    const int N = P.n_nodes, E = P.n_elems, K = P.n_cnodes;
    std::vector<int32_t> rf(std::max(K, 1));
    std::vector<uint8_t> er(E);
    int done_steps = 0;
    for (int s = 0; s < n_steps; ++s) {
        StepCtl before = S.read_ctl();
        if (before.done) break;
        S.launch_step();
        StepCtl c = S.read_ctl();
        if (f_int)
            HIP_OK(hipMemcpy(f_int + (size_t)s * 3 * N, S.D->f_int.p, 3 * (size_t)N * sizeof(double), hipMemcpyDeviceToHost));
        if (f_ext)
            HIP_OK(hipMemcpy(f_ext + (size_t)s * 3 * N, S.D->f_ext.p, 3 * (size_t)N * sizeof(double), hipMemcpyDeviceToHost));
        if (u_out)
            HIP_OK(hipMemcpy(u_out + (size_t)s * 3 * N, S.D->u.p, 3 * (size_t)N * sizeof(double), hipMemcpyDeviceToHost));
        if (hit && K > 0) {
            HIP_OK(hipMemcpy(rf.data(), S.D->rec_face.p, K * sizeof(int32_t), hipMemcpyDeviceToHost));
            for (int k = 0; k < K; ++k) {
                const int gf = rf[k];
                hit[(size_t)s * K + k] = gf < 0 ? -1 : gf - P.cdesc[P.cn_contact[k]].face_begin;
            }
        }
        if (eroded) {
            HIP_OK(hipMemcpy(er.data(), S.D->eroded.p, E, hipMemcpyDeviceToHost));
            for (int sl = 0; sl < E; ++sl) eroded[(size_t)s * E + P.slot_elem[sl]] = er[sl];
        }
        if (dt) dt[s] = c.dt;
        done_steps = s + 1;
        if (c.error) {
            // the taps of the failing step are kept for inspection
            if (steps_done) *steps_done = done_steps;
            fail(CRUSH_E_NONFINITE, "displacement became non-finite at step %lld", c.step);
        }
    }
    if (steps_done) *steps_done = done_steps;
}

}  // namespace crush
