// One solve at a time on the GPU: the reference's central-difference loop
// (src/solver.rs:287-352) as a sequence of kernels per step, one model per
// HIP stream.  Nodal state is f64, element arithmetic f32, contact and the
// stable-step pass f64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "common.hpp"
#include "device_common.hpp"
#include "prep.hpp"
#include "solver_plain.hpp"

namespace crush {

#define HIP_OK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) fail(CRUSH_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));       \
    } while (0)

// ------------------------------------------------------------- buffers

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t count) {
        n = std::max<size_t>(count, 1);
        hipError_t e = hipMalloc(&p, n * sizeof(T));
        if (e != hipSuccess) fail(CRUSH_E_OOM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
    }
    void upload(const std::vector<T>& v, hipStream_t s) {
        alloc(v.size());
        if (!v.empty()) HIP_OK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void zero(hipStream_t s) { HIP_OK(hipMemsetAsync(p, 0, n * sizeof(T), s)); }
};

struct DeviceModel {
    const Prepared& P;
    hipStream_t stream;
    int dt_blocks = 0;
    DevBuf<int32_t> conn, ecls, cls_mat, n2e_ptr, n2e_idx;
    DevBuf<float> cls_grad, cls_vol, cls_gamma, cls_H, state, fe;
    DevBuf<double> cls_dtf, elem_mass, extra, X, u, v, f_int, f_ext, inv_mass, mass, added, part_min, part_sum;
    DevBuf<uint8_t> eroded, fixed;
    DevBuf<MatParams> mats;
    DevBuf<ContactDesc> cdesc;
    DevBuf<int32_t> cn_node, cn_contact, cf_nodes, cf_contact, tn_node, tn_ptr, tn_ent, sblk_contact, sblk_start;
    DevBuf<double> cn_scale, fgeo, rec_f, face_force;
    DevBuf<int32_t> rec_face, fh_count, fh_list;
    DevBuf<int32_t> g_nodes_ptr, g_nodes, g_front_ptr, g_front, g_accel, g_axis;
    DevBuf<double> g_mass;
    DevBuf<StepCtl> ctl;

    DeviceModel(const Prepared& p, hipStream_t s) : P(p), stream(s) {
        const int E = P.n_elems, N = P.n_nodes;
        conn.upload(P.conn, s);
        ecls.upload(P.ecls, s);
        cls_mat.upload(P.cls_mat, s);
        n2e_ptr.upload(P.n2e_ptr, s);
        n2e_idx.upload(P.n2e_idx, s);
        cls_grad.upload(P.cls_grad, s);
        cls_vol.upload(P.cls_vol, s);
        cls_gamma.upload(P.cls_gamma, s);
        cls_H.upload(P.cls_H, s);
        cls_dtf.upload(P.cls_dtf, s);
        elem_mass.upload(P.elem_mass, s);
        mats.upload(P.mats, s);
        // honeycomb state: H_prev = 0, R = I, sigma = 0, c = 0
        std::vector<float> st((size_t)25 * std::max(P.n_hc, 1), 0.0f);
        for (int k : {9, 13, 17})
            for (int i = 0; i < P.n_hc; ++i) st[(size_t)k * P.n_hc + i] = 1.0f;
        state.upload(st, s);
        fe.alloc((size_t)24 * E);
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
        tn_node.upload(P.tn_node, s);
        tn_ptr.upload(P.tn_ptr, s);
        tn_ent.upload(P.tn_ent, s);
        sblk_contact.upload(P.sblk_contact, s);
        sblk_start.upload(P.sblk_start, s);
        fgeo.alloc((size_t)FG * P.n_cfaces);
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
    }
};

// ------------------------------------------------------------- kernels

// Per-element stable step (src/kernel.rs:214-237) and the mass each element
// would need under selective mass scaling (src/kernel.rs:243-258); block
// minimum / sum partials in a fixed order.
__global__ __launch_bounds__(256) void k_dt_elem(int E, const int32_t* __restrict__ conn, const double* __restrict__ X,
                                                 const double* __restrict__ u, const uint8_t* __restrict__ eroded,
                                                 const int32_t* __restrict__ ecls, const double* __restrict__ cdtf,
                                                 const double* __restrict__ emass, double* __restrict__ extra,
                                                 double* __restrict__ part_min, double* __restrict__ part_sum,
                                                 const StepCtl* __restrict__ ctl) {
#pragma clang fp contract(off)
    __shared__ double smin[256], ssum[256];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double dte = INFINITY, ex = 0.0;
    if (s < E && !ctl->done) {
        if (!eroded[s]) {
            double p[8][3];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int n = conn[(size_t)i * E + s];
#pragma unroll
                for (int d = 0; d < 3; ++d) p[i][d] = X[3 * (size_t)n + d] + u[3 * (size_t)n + d];
            }
            double l, vol;
            char_length_d(p, &l, &vol);
            if (isfinite(l) && vol > 0.0) dte = l * cdtf[ecls[s]];
        }
        const double floor_ = ctl->dt_floor;
        if (dte < floor_) {
            double r = floor_ / dte;
            ex = (fmin(r * r, 100.0) - 1.0) * emass[s];
        }
        extra[s] = ex;
    }
    smin[threadIdx.x] = dte;
    ssum[threadIdx.x] = ex;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            ssum[threadIdx.x] = ssum[threadIdx.x] + ssum[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part_min[blockIdx.x] = smin[0];
        part_sum[blockIdx.x] = ssum[0];
    }
}

// The adaptive step decision of src/solver.rs:288-299.
__global__ __launch_bounds__(256) void k_dt_reduce(int nb, const double* __restrict__ part_min,
                                                   const double* __restrict__ part_sum, StepCtl* __restrict__ ctl) {
    __shared__ double smin[256], ssum[256];
    double mn = INFINITY, sm = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) {
        mn = fmin(mn, part_min[i]);
        sm += part_sum[i];
    }
    smin[threadIdx.x] = mn;
    ssum[threadIdx.x] = sm;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            ssum[threadIdx.x] += ssum[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ctl->scale_now = 0;
        if (ctl->done) return;
        const double dts = smin[0];
        if (isfinite(dts)) {
            double dt_use = dts;
            if (ctl->dt_floor > 0.0 && dts < ctl->dt_floor) {
                ctl->scale_now = 1;
                ctl->added_total_max = fmax(ctl->added_total_max, ssum[0]);
                dt_use = ctl->dt_floor;
            }
            ctl->dt_cfl = ctl->dt_scale * fmin(dt_use, ctl->contact_dt);
        }
    }
}

// Added mass per node, summed over its elements in element order, and the
// inverse masses (src/kernel.rs:243-258, src/solver.rs:202-210).
__global__ __launch_bounds__(256) void k_mass(int N, const int32_t* __restrict__ n2e_ptr, const int32_t* __restrict__ n2e_idx,
                                              const double* __restrict__ extra, const double* __restrict__ mass,
                                              double* __restrict__ added, double* __restrict__ inv_mass,
                                              const StepCtl* __restrict__ ctl) {
#pragma clang fp contract(off)
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N || !ctl->scale_now) return;
    double a = 0.0;
    for (int k = n2e_ptr[n]; k < n2e_ptr[n + 1]; ++k) a += extra[n2e_idx[k] >> 3] / 8.0;
    added[n] = a;
    const double m = mass[n] + a;
    inv_mass[n] = m > 0.0 ? 1.0 / m : 0.0;
}

__global__ void k_step_begin(StepCtl* __restrict__ ctl) {
    if (ctl->done) return;
    ctl->dt = fmax(fmin(ctl->dt_cfl, ctl->end_time - ctl->t), 1e-15);
    ctl->pair_hits = 0;
}

// Current face geometry of every contact face.
__global__ __launch_bounds__(256) void k_contact_faces(int F, const int32_t* __restrict__ cf_nodes,
                                                       const int32_t* __restrict__ cf_contact,
                                                       const ContactDesc* __restrict__ cdesc, const double* __restrict__ X,
                                                       const double* __restrict__ u, double* __restrict__ fgeo,
                                                       int32_t* __restrict__ fh_count, const StepCtl* __restrict__ ctl) {
#pragma clang fp contract(off)
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F || ctl->done) return;
    double x[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int n = cf_nodes[(size_t)k * F + f];
#pragma unroll
        for (int d = 0; d < 3; ++d) x[k][d] = X[3 * (size_t)n + d] + u[3 * (size_t)n + d];
    }
    double g[FG];
    face_geometry(x, cdesc[cf_contact[f]].max_distance, g);
#pragma unroll
    for (int k = 0; k < FG; ++k) fgeo[(size_t)k * F + f] = g[k];
    fh_count[f] = 0;
}

constexpr int GRID_MAX_CELLS = 2048;
constexpr int GRID_MAX_ENTRIES = 8192;

// Block-wide exclusive scan of s[0..n) in place, 256 threads (4 wave64s):
// each thread scans a run of consecutive entries, the run totals are
// scanned across each wave with shuffles and across waves through LDS.
__device__ void block_exclusive_scan(int* s, int n, int* s_tmp /* [>= 4] */) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int per = (n + 255) / 256;
    const int b = min(t * per, n), e = min(b + per, n);
    int sum = 0;
    for (int i = b; i < e; ++i) sum += s[i];
    int incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) s_tmp[wave] = incl;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wave; ++w) base += s_tmp[w];
    int acc = base + incl - sum;
    for (int i = b; i < e; ++i) {
        int v = s[i];
        s[i] = acc;
        acc += v;
    }
    __syncthreads();
}

// Bucketed node-to-face search: the block builds a uniform grid over its
// contact's face candidate boxes in LDS, then each thread finds, for one
// contacting node, the lowest-index face that accepts it (the reference's
// first hit in face order, src/contact.rs:113-157) and its penalty force
// (src/contact.rs:136-144).
__global__ __launch_bounds__(256) void k_contact_search(
    const int32_t* __restrict__ sblk_contact, const int32_t* __restrict__ sblk_start,
    const ContactDesc* __restrict__ cdesc, int F, const double* __restrict__ fgeo, const int32_t* __restrict__ cn_node,
    const double* __restrict__ cn_scale, const double* __restrict__ X, const double* __restrict__ u,
    const double* __restrict__ mass, int32_t* __restrict__ rec_face, double* __restrict__ rec_f,
    int32_t* __restrict__ fh_count, int32_t* __restrict__ fh_list, StepCtl* __restrict__ ctl) {
#pragma clang fp contract(off)
    __shared__ int s_count[GRID_MAX_CELLS + 1];
    __shared__ int s_cursor[GRID_MAX_CELLS];
    __shared__ int s_entries[GRID_MAX_ENTRIES];
    __shared__ int s_tmp[256];
    __shared__ double s_red[6][256];
    __shared__ double s_glo[3], s_h;
    __shared__ int s_dim[3], s_overflow;
    if (ctl->done) return;
    const int c = sblk_contact[blockIdx.x];
    const ContactDesc cd = cdesc[c];
    const int t = threadIdx.x;
    const int fb = cd.face_begin, nf = cd.n_faces;
    // grid bounds and cell size
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY}, sz = 0.0;
    for (int f = t; f < nf; f += 256) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], fgeo[(size_t)(FG_CLO + d) * F + fb + f]);
            hi[d] = fmax(hi[d], fgeo[(size_t)(FG_CHI + d) * F + fb + f]);
        }
        sz = fmax(sz, fgeo[(size_t)FG_SZ * F + fb + f]);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        s_red[d][t] = lo[d];
        s_red[3 + d][t] = -hi[d];
    }
    s_tmp[t] = 0;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 6; ++k) s_red[k][t] = fmin(s_red[k][t], s_red[k][t + w]);
        __syncthreads();
    }
    double* s_sz = &s_red[0][0];  // reuse after the bounds are read
    double glo[3] = {s_red[0][0], s_red[1][0], s_red[2][0]};
    double ghi[3] = {-s_red[3][0], -s_red[4][0], -s_red[5][0]};
    __syncthreads();
    s_sz[t] = sz;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) s_sz[t] = fmax(s_sz[t], s_sz[t + w]);
        __syncthreads();
    }
    if (t == 0) {
        double h = fmax(s_sz[0], 1e-9);
        int dim[3];
        for (int iter = 0; iter < 64; ++iter) {
            long long cells = 1;
            for (int d = 0; d < 3; ++d) {
                double ext = ghi[d] - glo[d];
                dim[d] = ext > 0.0 ? (int)fmin(ceil(ext / h), 1e6) : 1;
                if (dim[d] < 1) dim[d] = 1;
                cells *= dim[d];
            }
            if (cells <= GRID_MAX_CELLS) break;
            h *= 1.25;
        }
        s_h = h;
        for (int d = 0; d < 3; ++d) {
            s_dim[d] = dim[d];
            s_glo[d] = glo[d];
        }
        s_overflow = (nf == 0 || !(isfinite(glo[0]) && isfinite(ghi[0]))) ? 1 : 0;
    }
    __syncthreads();
    const double h = s_h;
    const int dx = s_dim[0], dy = s_dim[1], dz = s_dim[2];
    const int ncell = dx * dy * dz;
    auto cell_of = [&](double x, int d, int dimd) {
        int ci = (int)floor((x - s_glo[d]) / h);
        return min(max(ci, 0), dimd - 1);
    };
    for (int i = t; i < ncell + 1; i += 256) s_count[i] = 0;
    __syncthreads();
    if (!s_overflow) {
        for (int f = t; f < nf; f += 256) {
            const double* g = fgeo + fb + f;
            double flo[3], fhi[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                flo[d] = g[(size_t)(FG_CLO + d) * F];
                fhi[d] = g[(size_t)(FG_CHI + d) * F];
            }
            if (!(flo[0] <= fhi[0] && flo[1] <= fhi[1] && flo[2] <= fhi[2])) continue;
            int a0 = cell_of(flo[0], 0, dx), a1 = cell_of(fhi[0], 0, dx);
            int b0 = cell_of(flo[1], 1, dy), b1 = cell_of(fhi[1], 1, dy);
            int c0 = cell_of(flo[2], 2, dz), c1 = cell_of(fhi[2], 2, dz);
            for (int kz = c0; kz <= c1; ++kz)
                for (int ky = b0; ky <= b1; ++ky)
                    for (int kx = a0; kx <= a1; ++kx) atomicAdd(&s_count[(kz * dy + ky) * dx + kx], 1);
        }
    }
    __syncthreads();
    block_exclusive_scan(s_count, ncell + 1, s_tmp);
    if (t == 0 && s_count[ncell] > GRID_MAX_ENTRIES) s_overflow = 1;
    for (int i = t; i < ncell; i += 256) s_cursor[i] = s_count[i];
    __syncthreads();
    if (!s_overflow) {
        for (int f = t; f < nf; f += 256) {
            const double* g = fgeo + fb + f;
            double flo[3], fhi[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                flo[d] = g[(size_t)(FG_CLO + d) * F];
                fhi[d] = g[(size_t)(FG_CHI + d) * F];
            }
            if (!(flo[0] <= fhi[0] && flo[1] <= fhi[1] && flo[2] <= fhi[2])) continue;
            int a0 = cell_of(flo[0], 0, dx), a1 = cell_of(fhi[0], 0, dx);
            int b0 = cell_of(flo[1], 1, dy), b1 = cell_of(fhi[1], 1, dy);
            int c0 = cell_of(flo[2], 2, dz), c1 = cell_of(fhi[2], 2, dz);
            for (int kz = c0; kz <= c1; ++kz)
                for (int ky = b0; ky <= b1; ++ky)
                    for (int kx = a0; kx <= a1; ++kx) {
                        int pos = atomicAdd(&s_cursor[(kz * dy + ky) * dx + kx], 1);
                        s_entries[pos] = f;
                    }
        }
    }
    __syncthreads();
    const int k = sblk_start[blockIdx.x] + t;
    if (k >= cd.n_nodes) return;
    const int gk = cd.node_begin + k;
    const int node = cn_node[gk];
    double p[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) p[d] = X[3 * (size_t)node + d] + u[3 * (size_t)node + d];
    int best = -1;
    double best_sd = 0.0;
    if (s_overflow) {
        for (int f = 0; f < nf; ++f) {
            double sd;
            if (face_accepts(fgeo, F, fb + f, p, cd.max_distance, &sd)) {
                best = f;
                best_sd = sd;
                break;
            }
        }
    } else {
        bool inside = true;
        int ci[3];
        const int dims[3] = {dx, dy, dz};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double r = (p[d] - s_glo[d]) / h;
            inside = inside && r >= 0.0 && r < (double)dims[d] + 1.0;
            int q = (int)floor(r);
            ci[d] = min(max(q, 0), dims[d] - 1);
        }
        if (inside) {
            const int cell = (ci[2] * dy + ci[1]) * dx + ci[0];
            for (int q = s_count[cell]; q < s_count[cell + 1]; ++q) {
                const int f = s_entries[q];
                if (best >= 0 && f >= best) continue;
                double sd;
                if (face_accepts(fgeo, F, fb + f, p, cd.max_distance, &sd)) {
                    best = f;
                    best_sd = sd;
                }
            }
        }
    }
    if (best < 0) {
        rec_face[gk] = -1;
        return;
    }
    const int gf = fb + best;
    const double m = mass[node];
    const double dt = ctl->dt;
    const double depth = fmin(-best_sd, cd.depth_cap);
    const double kk = fmin(cd.stiffness * cn_scale[gk], C_SOFT * m / (dt * dt));
    const double mag = fmin(kk * depth + 0.0, m * C_MAX_ACCEL);
#pragma unroll
    for (int d = 0; d < 3; ++d) rec_f[3 * (size_t)gk + d] = mag * fgeo[(size_t)(FG_N + d) * F + gf];
    rec_face[gk] = gf;
    const int slot = atomicAdd(&fh_count[gf], 1);
    if (slot < FACE_HITS_CAP) fh_list[(size_t)gf * FACE_HITS_CAP + slot] = gk;
    if (cd.pair) atomicAdd(&ctl->pair_hits, 1);
}

// Sum of the quarter reactions each face receives, in contacting-node
// order (deterministic whatever order the hits were recorded in).
__global__ __launch_bounds__(256) void k_contact_faceforce(int F, const int32_t* __restrict__ cf_contact,
                                                           const ContactDesc* __restrict__ cdesc,
                                                           const int32_t* __restrict__ fh_count,
                                                           const int32_t* __restrict__ fh_list,
                                                           const int32_t* __restrict__ rec_face,
                                                           const double* __restrict__ rec_f,
                                                           double* __restrict__ face_force, const StepCtl* __restrict__ ctl) {
#pragma clang fp contract(off)
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F || ctl->done) return;
    const int cnt = fh_count[f];
    double s[3] = {0.0, 0.0, 0.0};
    if (cnt > 0 && cnt <= FACE_HITS_CAP) {
        int ids[FACE_HITS_CAP];
        for (int i = 0; i < cnt; ++i) ids[i] = fh_list[(size_t)f * FACE_HITS_CAP + i];
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
            for (int d = 0; d < 3; ++d) s[d] += 0.25 * rec_f[3 * (size_t)ids[i] + d];
    } else if (cnt > FACE_HITS_CAP) {
        const ContactDesc cd = cdesc[cf_contact[f]];
        for (int k = cd.node_begin; k < cd.node_begin + cd.n_nodes; ++k)
            if (rec_face[k] == f)
#pragma unroll
                for (int d = 0; d < 3; ++d) s[d] += 0.25 * rec_f[3 * (size_t)k + d];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) face_force[3 * (size_t)f + d] = s[d];
}

// External force of every node touched by a contact: its own penalty
// forces and the reactions of the faces it belongs to, in contact order.
__global__ __launch_bounds__(256) void k_contact_apply(int T, const int32_t* __restrict__ tn_node,
                                                       const int32_t* __restrict__ tn_ptr, const int32_t* __restrict__ tn_ent,
                                                       const int32_t* __restrict__ rec_face, const double* __restrict__ rec_f,
                                                       const double* __restrict__ face_force, double* __restrict__ f_ext,
                                                       const StepCtl* __restrict__ ctl) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T || ctl->done) return;
    double s[3] = {0.0, 0.0, 0.0};
    for (int k = tn_ptr[i]; k < tn_ptr[i + 1]; ++k) {
        const int e = tn_ent[k];
        if (e >= 0) {
            if (rec_face[e] >= 0)
#pragma unroll
                for (int d = 0; d < 3; ++d) s[d] += rec_f[3 * (size_t)e + d];
        } else {
            const int f = -e - 1;
#pragma unroll
            for (int d = 0; d < 3; ++d) s[d] -= face_force[3 * (size_t)f + d];
        }
    }
    const int n = tn_node[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) f_ext[3 * (size_t)n + d] = s[d];
}

// First half step: v_1/2 = v + dt/2 a(f_ext, f_int); u += dt v_1/2
// (src/solver.rs:313-323).  v holds v_1/2 afterwards.
__global__ __launch_bounds__(256) void k_half(int N, const double* __restrict__ f_ext, const double* __restrict__ f_int,
                                              const double* __restrict__ inv_mass, double* __restrict__ v,
                                              double* __restrict__ u, const uint8_t* __restrict__ fixed,
                                              const StepCtl* __restrict__ ctl) {
#pragma clang fp contract(off)
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N || ctl->done) return;
    const double dt = ctl->dt;
    const double im = inv_mass[n];
    if (fixed[n]) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            u[3 * (size_t)n + d] = 0.0;
            v[3 * (size_t)n + d] = 0.0;
        }
        return;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const size_t k = 3 * (size_t)n + d;
        const double a = (f_ext[k] - f_int[k]) * im;
        const double vh = v[k] + 0.5 * dt * a;
        v[k] = vh;
        u[k] += dt * vh;
    }
}

// Element forces at the new configuration, f32 (src/kernel_simd.rs:300-491
// for honeycomb, src/kernel.rs:116-156 for the elastic scalar path).
// Slots [begin, end) share one material kind.
template <int KIND>
__global__ __launch_bounds__(256) void k_force(int begin, int end, int E, int C, const int32_t* __restrict__ conn,
                                               const int32_t* __restrict__ ecls, const int32_t* __restrict__ cls_mat,
                                               const float* __restrict__ cgrad, const float* __restrict__ cvol,
                                               const float* __restrict__ cgam, const float* __restrict__ chs,
                                               const MatParams* __restrict__ mats, const double* __restrict__ u,
                                               float* __restrict__ state, int n_hc, uint8_t* __restrict__ eroded,
                                               float* __restrict__ fe, StepCtl* __restrict__ ctl) {
    const int s = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= end || ctl->done) return;
    if (eroded[s]) {
#pragma unroll
        for (int k = 0; k < 24; ++k) fe[(size_t)k * E + s] = 0.0f;
        return;
    }
    int nd[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) nd[i] = conn[(size_t)i * E + s];
    const int c = ecls[s];
    const double u0x = u[3 * (size_t)nd[0]], u0y = u[3 * (size_t)nd[0] + 1], u0z = u[3 * (size_t)nd[0] + 2];
    float dl[8][3];
    dl[0][0] = dl[0][1] = dl[0][2] = 0.0f;
#pragma unroll
    for (int i = 1; i < 8; ++i) {
        dl[i][0] = (float)(u[3 * (size_t)nd[i]] - u0x);
        dl[i][1] = (float)(u[3 * (size_t)nd[i] + 1] - u0y);
        dl[i][2] = (float)(u[3 * (size_t)nd[i] + 2] - u0z);
    }
    auto grad = [&](int k) { return cgrad[(size_t)k * C + c]; };
    auto gam = [&](int k) { return cgam[(size_t)k * C + c]; };
    auto hsym = [&](int k) { return chs[(size_t)k * C + c]; };
    // F = I + sum_i (u_i - u_0) (x) grad_i
    float hd[9];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            float sum = 0.0f;
#pragma unroll
            for (int i = 1; i < 8; ++i) sum += dl[i][a] * grad(3 * i + b);
            hd[3 * a + b] = sum;
        }
    float F[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = hd[k] + ((k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f);
    const float J = det3f(F);
    if (!(J > 0.0f)) {
        eroded[s] = 1;
        atomicAdd(&ctl->eroded_total, 1ULL);
#pragma unroll
        for (int k = 0; k < 24; ++k) fe[(size_t)k * E + s] = 0.0f;
        return;
    }
    const MatParams& mp = mats[cls_mat[c]];
    float tau[9];
    if (KIND == CRUSH_MAT_HONEYCOMB) {
        float st[25];
#pragma unroll
        for (int k = 0; k < 25; ++k) st[k] = state[(size_t)k * n_hc + s];
        honeycomb_update(mp, hd, F, st, tau);
#pragma unroll
        for (int k = 0; k < 25; ++k) state[(size_t)k * n_hc + s] = st[k];
    } else {
        hencky_elastic(mp, hd, tau);
    }
    float fo[8][3];
    element_nodal_forces(tau, F, J, cvol[c], dl, grad, gam, hsym, fo);
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d) fe[(size_t)(3 * i + d) * E + s] = fo[i][d];
}

// Node-centric gather of the element forces (fixed element order, f64
// sums) and the second half step with the stale f_ext
// (src/solver.rs:327-342).
__global__ __launch_bounds__(256) void k_gather_update(int N, int E, const int32_t* __restrict__ n2e_ptr,
                                                       const int32_t* __restrict__ n2e_idx, const float* __restrict__ fe,
                                                       const double* __restrict__ f_ext, const double* __restrict__ inv_mass,
                                                       double* __restrict__ v, double* __restrict__ f_int,
                                                       const double* __restrict__ u, const uint8_t* __restrict__ fixed,
                                                       StepCtl* __restrict__ ctl) {
#pragma clang fp contract(off)
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N || ctl->done) return;
    double f[3] = {0.0, 0.0, 0.0};
    for (int k = n2e_ptr[n]; k < n2e_ptr[n + 1]; ++k) {
        const int code = n2e_idx[k];
        const size_t base = (size_t)(3 * (code & 7)) * E + (code >> 3);
#pragma unroll
        for (int d = 0; d < 3; ++d) f[d] += (double)fe[base + (size_t)d * E];
    }
    const double dt = ctl->dt, damp = ctl->damping, im = inv_mass[n];
    bool bad = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const size_t k = 3 * (size_t)n + d;
        f_int[k] = f[d];
        const double a = (f_ext[k] - f[d]) * im;
        double vn = (v[k] + 0.5 * dt * a) * damp;
        if (fixed[n]) vn = 0.0;
        v[k] = vn;
        bad = bad || !isfinite(vn) || !isfinite(u[k]);
    }
    if (bad) atomicOr(&ctl->error, 1);
}

// End of step: time, step count, crush samples, separation and the stop
// decision.  One block.
__global__ __launch_bounds__(256) void k_step_end(int n_groups, const int32_t* __restrict__ g_nodes_ptr,
                                                  const int32_t* __restrict__ g_nodes, const int32_t* __restrict__ g_front_ptr,
                                                  const int32_t* __restrict__ g_front, const int32_t* __restrict__ g_accel,
                                                  const int32_t* __restrict__ g_axis, const double* __restrict__ mass,
                                                  const double* __restrict__ X, const double* __restrict__ u,
                                                  const double* __restrict__ v, StepCtl* __restrict__ ctl) {
#pragma clang fp contract(off)
    __shared__ double red[7][256];
    __shared__ int s_done;
    const int t = threadIdx.x;
    if (ctl->done) return;
    const long long step = ctl->step + 1;
    // crush of each group: |u_axis(accel) - mean u_axis(front)|
    for (int g = 0; g < n_groups && g < 8; ++g) {
        const int b = g_front_ptr[g], e = g_front_ptr[g + 1], ax = g_axis[g];
        double s = 0.0;
        for (int i = b + t; i < e; i += 256) s += u[3 * (size_t)g_front[i] + ax];
        red[0][t] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (t < w) red[0][t] += red[0][t + w];
            __syncthreads();
        }
        if (t == 0 && g_accel[g] >= 0 && e > b) {
            double cr = fabs(u[3 * (size_t)g_accel[g] + ax] - red[0][0] / (double)(e - b));
            ctl->crush_max[g] = fmax(ctl->crush_max[g], cr);
        }
        __syncthreads();
    }
    if (t == 0) {
        const int hits = ctl->pair_hits;
        if (hits > 0) {
            ctl->had_pair = 1;
            ctl->quiet = 0;
        } else {
            ctl->quiet += 1;
        }
        s_done = 0;
    }
    __syncthreads();
    const bool check_sep = ctl->sep_steps > 0 && n_groups >= 2 && ctl->had_pair && ctl->quiet >= ctl->sep_steps &&
                           step % ctl->chk == 0;
    if (check_sep) {
        double acc[2][7];
        for (int g = 0; g < 2; ++g) {
            double s[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int i = g_nodes_ptr[g] + t; i < g_nodes_ptr[g + 1]; i += 256) {
                const int n = g_nodes[i];
                const double m = mass[n];
                s[0] += m;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    s[1 + d] += m * (X[3 * (size_t)n + d] + u[3 * (size_t)n + d]);
                    s[4 + d] += m * v[3 * (size_t)n + d];
                }
            }
            for (int k = 0; k < 7; ++k) red[k][t] = s[k];
            __syncthreads();
            for (int w = 128; w > 0; w >>= 1) {
                if (t < w)
                    for (int k = 0; k < 7; ++k) red[k][t] += red[k][t + w];
                __syncthreads();
            }
            for (int k = 0; k < 7; ++k) acc[g][k] = red[k][0];
            __syncthreads();
        }
        if (t == 0) {
            double sep = 0.0;
            for (int d = 0; d < 3; ++d) {
                double dx = acc[1][1 + d] / acc[1][0] - acc[0][1 + d] / acc[0][0];
                double dv = acc[1][4 + d] / acc[1][0] - acc[0][4 + d] / acc[0][0];
                sep += dv * dx;
            }
            if (sep > 0.0) {
                ctl->separated = 1;
                s_done = 1;
            }
        }
    }
    __syncthreads();
    if (t == 0) {
        ctl->t = ctl->t + ctl->dt;
        ctl->step = step;
        if (s_done || ctl->error || !(ctl->t < ctl->end_time)) ctl->done = 1;
    }
}

// ------------------------------------------------------------ host loop

static unsigned nblk(long long n, int b = 256) { return (unsigned)std::max<long long>((n + b - 1) / b, 1); }

struct PlainSolver {
    const Prepared& P;
    hipStream_t stream = nullptr;
    std::unique_ptr<DeviceModel> D;
    long long host_step = 0;

    explicit PlainSolver(const Prepared& p) : P(p) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) fail(CRUSH_E_DEVICE, "no HIP device available");
        HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        D.reset(new DeviceModel(P, stream));
    }
    ~PlainSolver() {
        D.reset();
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Launch one step (the kernels no-op once the solve is done).
    void launch_step() {
        DeviceModel& d = *D;
        const int E = P.n_elems, N = P.n_nodes;
        if (P.adaptive && host_step % P.settings.adaptive_check_steps == 0) {
            hipLaunchKernelGGL(k_dt_elem, dim3(d.dt_blocks), dim3(256), 0, stream, E, d.conn.p, d.X.p, d.u.p, d.eroded.p,
                               d.ecls.p, d.cls_dtf.p, d.elem_mass.p, d.extra.p, d.part_min.p, d.part_sum.p, d.ctl.p);
            hipLaunchKernelGGL(k_dt_reduce, dim3(1), dim3(256), 0, stream, d.dt_blocks, d.part_min.p, d.part_sum.p, d.ctl.p);
            hipLaunchKernelGGL(k_mass, dim3(nblk(N)), dim3(256), 0, stream, N, d.n2e_ptr.p, d.n2e_idx.p, d.extra.p,
                               d.mass.p, d.added.p, d.inv_mass.p, d.ctl.p);
        }
        hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, stream, d.ctl.p);
        if (P.n_cfaces > 0) {
            const int F = P.n_cfaces, T = (int)P.tn_node.size();
            hipLaunchKernelGGL(k_contact_faces, dim3(nblk(F)), dim3(256), 0, stream, F, d.cf_nodes.p, d.cf_contact.p,
                               d.cdesc.p, d.X.p, d.u.p, d.fgeo.p, d.fh_count.p, d.ctl.p);
            if (!P.sblk_contact.empty())
                hipLaunchKernelGGL(k_contact_search, dim3((unsigned)P.sblk_contact.size()), dim3(256), 0, stream,
                                   d.sblk_contact.p, d.sblk_start.p, d.cdesc.p, F, d.fgeo.p, d.cn_node.p, d.cn_scale.p,
                                   d.X.p, d.u.p, d.mass.p, d.rec_face.p, d.rec_f.p, d.fh_count.p, d.fh_list.p, d.ctl.p);
            hipLaunchKernelGGL(k_contact_faceforce, dim3(nblk(F)), dim3(256), 0, stream, F, d.cf_contact.p, d.cdesc.p,
                               d.fh_count.p, d.fh_list.p, d.rec_face.p, d.rec_f.p, d.face_force.p, d.ctl.p);
            if (T > 0)
                hipLaunchKernelGGL(k_contact_apply, dim3(nblk(T)), dim3(256), 0, stream, T, d.tn_node.p, d.tn_ptr.p,
                                   d.tn_ent.p, d.rec_face.p, d.rec_f.p, d.face_force.p, d.f_ext.p, d.ctl.p);
        }
        hipLaunchKernelGGL(k_half, dim3(nblk(N)), dim3(256), 0, stream, N, d.f_ext.p, d.f_int.p, d.inv_mass.p, d.v.p,
                           d.u.p, d.fixed.p, d.ctl.p);
        const int C = P.n_classes;
        if (P.n_hc > 0)
            hipLaunchKernelGGL(k_force<CRUSH_MAT_HONEYCOMB>, dim3(nblk(P.n_hc)), dim3(256), 0, stream, 0, P.n_hc, E, C,
                               d.conn.p, d.ecls.p, d.cls_mat.p, d.cls_grad.p, d.cls_vol.p, d.cls_gamma.p, d.cls_H.p,
                               d.mats.p, d.u.p, d.state.p, P.n_hc, d.eroded.p, d.fe.p, d.ctl.p);
        if (E > P.n_hc)
            hipLaunchKernelGGL(k_force<CRUSH_MAT_ELASTIC>, dim3(nblk(E - P.n_hc)), dim3(256), 0, stream, P.n_hc, E, E, C,
                               d.conn.p, d.ecls.p, d.cls_mat.p, d.cls_grad.p, d.cls_vol.p, d.cls_gamma.p, d.cls_H.p,
                               d.mats.p, d.u.p, d.state.p, P.n_hc, d.eroded.p, d.fe.p, d.ctl.p);
        hipLaunchKernelGGL(k_gather_update, dim3(nblk(N)), dim3(256), 0, stream, N, E, d.n2e_ptr.p, d.n2e_idx.p, d.fe.p,
                           d.f_ext.p, d.inv_mass.p, d.v.p, d.f_int.p, d.u.p, d.fixed.p, d.ctl.p);
        hipLaunchKernelGGL(k_step_end, dim3(1), dim3(256), 0, stream, P.n_groups, d.g_nodes_ptr.p, d.g_nodes.p,
                           d.g_front_ptr.p, d.g_front.p, d.g_accel.p, d.g_axis.p, d.mass.p, d.X.p, d.u.p, d.v.p, d.ctl.p);
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
}

// Guard against a solve that never reaches end_time (dt is floored at 1e-15).
static const long long kMaxSteps = 100000000LL;

void run_plain(const crush_model& m, crush_results* out) {
    auto t0 = std::chrono::steady_clock::now();
    Prepared P;
    prepare(m, P);
    PlainSolver S(P);
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, S.stream));
    const int chunk = 32;
    StepCtl c;
    while (true) {
        for (int i = 0; i < chunk; ++i) S.launch_step();
        c = S.read_ctl();
        if (c.done) break;
        if (S.host_step > kMaxSteps) fail(CRUSH_E_INPUT, "no end after %lld steps (t = %.3e)", S.host_step, c.t);
    }
    HIP_OK(hipEventRecord(e1, S.stream));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (c.error) fail(CRUSH_E_NONFINITE, "displacement became non-finite at step %lld (t = %.3e)", c.step, c.t);
    if (out) {
        fill_results(P, S, c, out);
        out->device_ms = ms;
        out->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
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
