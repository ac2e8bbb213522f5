// The one-solve-at-a-time GPU path.
#pragma once

#include <cstdint>

#include "../../include/crushrs.h"

namespace crush {

// Whole solve on the GPU (crush_run).
void run_plain(const crush_model& m, crush_results* out);

// The same solve with every device buffer in the library's static device
// scratch (crush_run_noalloc): no device memory is allocated.
void run_plain_noalloc(const crush_model& m, crush_results* out);

// First n_steps with per-step taps (crush_debug_run_steps).
void run_plain_taps(const crush_model& m, int n_steps, double* f_int, double* f_ext, double* u, int32_t* hit,
                    uint8_t* eroded,
                    double* dt, int32_t* steps_done);

}  // namespace crush
