// Histories and frames of a solve (src/solver.rs:8-62, 114-187, 265-357):
// the device records one history record per sample after the initial state
// (phases.hpp history_sample); the host adds the initial sample and lays
// the records out as crush_results asks.  Frames are the one-model path's.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/crushrs.h"
#include "phases.hpp"
#include "prep.hpp"

namespace crush {

// Copies the first `n` device history records (of hist_record_len doubles)
// to the host.
std::vector<double> history_download(const Prepared& P, const double* d_hist, int n, hipStream_t s);

// Fills out->history_* from the initial state and the device records
// (`n_taken` records taken, of which min(n_taken, P.n_hist) were kept).
void history_results(const Prepared& P, const double* recs, int n_taken, crush_results* out);

// Device frame buffers of one solve: `cap` frames of nodal displacement,
// plastic strain, stress and erosion (f32 / u8, element order of the
// model) and their times.  Frame 0 is the initial state.
struct FrameStore {
    int cap = 0;
    float *u = nullptr, *plastic = nullptr, *stress = nullptr;
    uint8_t* eroded = nullptr;
    int32_t* slot_elem = nullptr;
    double* t_dev = nullptr;
    ~FrameStore();
    void init(const Prepared& P, int capacity, hipStream_t s);
    // Frame k of the state after step `step`; the kernel does nothing
    // unless that step ran (ctl->step == step).
    void record(const InstDev& I, const Prepared& P, int k, long long step, hipStream_t s);
    // Copies min(n_frames, cap, out->frame_capacity) frames out.
    void fill(const Prepared& P, int n_frames, crush_results* out, hipStream_t s);
};

}  // namespace crush
