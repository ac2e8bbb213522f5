// Batched solves (crush_batch_* in include/crushrs.h).
#pragma once

#include <memory>

#include "../../include/crushrs.h"
#include "prep.hpp"

namespace crush {

// Per-instance summary gathered on the device after a run.
struct BatchSummary {
    double t;
    long long steps;
    long long eroded;
    double added_max;
    int done, error, separated, pulse_samples;
    int hist_samples, reserved0;
    double dv[8][3];
    double crush_end[8];
    double crush_max[8];
    unsigned long long t_retire;
};

struct BatchImpl;

void batch_defaults(crush_batch_options* o);
BatchImpl* batch_create(const crush_model* const* models, int n, const crush_batch_options* opt);
void batch_run(BatchImpl& bi, crush_batch_stats* st);
void batch_reset(BatchImpl& bi);
void batch_result(BatchImpl& bi, int i, crush_results* out);
void batch_free(BatchImpl* bi);
void batch_options_of(const BatchImpl& bi, crush_batch_options* o);
void batch_time_kernels(BatchImpl& bi, bool on);

}  // namespace crush
