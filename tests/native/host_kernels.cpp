// Test-only host build of the product's f32 element math
// (crushrs_amd/csrc/device_common.hpp), so the per-element arithmetic the
// GPU kernels run can be checked against the f64 oracle on the CPU.
#include <hip/hip_runtime.h>

#include "../../crushrs_amd/csrc/device_common.hpp"

using namespace crush;

static MatParams make(float youngs, float nu, const float* knots, float trans) {
    MatParams mp{};
    mp.youngs = youngs;
    mp.bulk = youngs / (3.0f * (1.0f - 2.0f * nu));
    mp.shear = youngs / (2.0f * (1.0f + nu));
    mp.trans = trans;
    for (int i = 0; i < 8; ++i) {
        mp.knot_c[i] = knots ? knots[3 * i] : INFINITY;
        mp.knot_s[i] = knots ? knots[3 * i + 1] : 0.0f;
        mp.knot_h[i] = knots ? knots[3 * i + 2] : 0.0f;
    }
    return mp;
}

extern "C" void hk_hencky(const float* hd, float youngs, float nu, float* tau) {
    MatParams mp = make(youngs, nu, nullptr, 1.0f);
    hencky_elastic(mp, hd, tau);
}

extern "C" void hk_honeycomb(const float* hd, float youngs, const float* knots, float trans, float* st, float* tau) {
    MatParams mp = make(youngs, 0.0f, knots, trans);
    honeycomb_update(mp, hd, st, tau);
}

extern "C" void hk_eigen(const float* a, float* vals, float* vecs) { sym_eigen3f(a, vals, vecs); }
