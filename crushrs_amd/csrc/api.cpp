// C ABI of the core entry points (include/crushrs.h).
#include <hip/hip_runtime.h>

#include <string>

#include "common.hpp"
#include "prep.hpp"
#include "solver_plain.hpp"

namespace crush {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }

}  // namespace crush

using namespace crush;

extern "C" {

int32_t crush_abi_version(void) { return CRUSH_ABI_VERSION; }

const char* crush_last_error(void) { return g_last_error.c_str(); }

int32_t crush_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int32_t crush_run(const crush_model* model, crush_results* out) {
    return guarded([&] {
        if (!model) fail(CRUSH_E_INPUT, "null model");
        run_plain(*model, out);
    });
}

int32_t crush_run_noalloc(const crush_model* model, crush_results* out) {
    return guarded([&] {
        if (!model) fail(CRUSH_E_INPUT, "null model");
        run_plain_noalloc(*model, out);
    });
}

int32_t crush_debug_run_steps(const crush_model* model, int32_t n_steps, double* f_int, double* f_ext, double* u,
                              int32_t* hit, uint8_t* eroded, double* dt, int32_t* steps_done) {
    return guarded([&] {
        if (!model || n_steps < 0) fail(CRUSH_E_INPUT, "bad arguments");
        run_plain_taps(*model, n_steps, f_int, f_ext, u, hit, eroded, dt, steps_done);
    });
}

int32_t crush_contact_node_count(const crush_model* model, int32_t* count) {
    return guarded([&] {
        if (!model || !count) fail(CRUSH_E_INPUT, "null argument");
        Prepared p;
        prepare(*model, p);
        *count = p.n_cnodes;
    });
}

int32_t crush_output_capacity(const crush_model* model, int32_t* history_samples, int32_t* frame_samples) {
    return guarded([&] {
        if (!model) fail(CRUSH_E_INPUT, "null argument");
        Prepared p;
        prepare(*model, p);
        if (history_samples) *history_samples = p.settings.history_steps > 0 ? 1 + p.n_hist : 0;
        if (frame_samples)
            *frame_samples = p.settings.frame_steps > 0 ? 1 + history_capacity(p, p.settings.frame_steps) : 0;
    });
}

}  // extern "C"
