"""The reference's envelope of answers for ill-conditioned models.

Some crash models sit at a bifurcation in the reference itself: a change
of an elastic modulus far below the precision of any material data flips
an element between surviving and eroding, and the step count, crush and
delta-v move with it.  For those models a single oracle solve is one
sample of the reference's answers; the GPU is compared with the spread of
several (the f64 oracle perturbed, and the reference's f32 lane path).
"""
import numpy as np

import oracle_ffi as of


def oracle_envelope(case):
    """The reference's own answers for a case: both of its element paths,
    the f64 scalar oracle and the default f32 lane path (kernel_simd.rs
    restated), each unperturbed and with every elastic part's Young's
    modulus scaled by 1 +- 1e-6 and 1 +- 1e-7 (far below the precision of
    any material data).  For a well-conditioned case these agree to a few
    mm/s; at a bifurcation they split between branches."""
    from concurrent.futures import ThreadPoolExecutor
    from crushrs_amd import _abi

    def one(arg):
        scale, lanes = arg
        m = case.build()
        v = m.view
        for p in range(v.n_parts):
            if v.materials[p].model == _abi.MAT_ELASTIC:
                v.materials[p].youngs_modulus *= scale
        return of.run_lanes(m, 1) if lanes else of.run(m)

    args = [(scale, lanes) for lanes in (False, True)
            for scale in (1.0, 1.0 + 1e-6, 1.0 - 1e-6, 1.0 + 1e-7, 1.0 - 1e-7)]
    with ThreadPoolExecutor(len(args)) as ex:
        return list(ex.map(one, args))


def within_envelope(r, ens, axis=0, dv_tol=0.05, dv_off_tol=0.1, crush_tol=2e-3):
    """The GPU result against the envelope of the reference's answers:
    delta-v along the impact axis (the reference's scalar delta-v, e.g.
    BarrierMetrics.delta_v, src/vehicle.rs:947) within dv_tol of the
    envelope, the other components within dv_off_tol; steps within 1 %,
    erosion count one the envelope holds, crush within crush_tol."""
    steps = [e.steps for e in ens]
    assert 0.99 * min(steps) - 1 <= r.steps <= 1.01 * max(steps) + 1, (r.steps, steps)
    assert r.eroded_elements in {e.eroded_elements for e in ens}, (r.eroded_elements, [e.eroded_elements for e in ens])
    for g in range(len(r.groups)):
        dv = np.array([e.groups[g].delta_v for e in ens])
        for d in range(3):
            tol = dv_tol if d == axis else dv_off_tol
            assert dv[:, d].min() - tol <= r.groups[g].delta_v[d] <= dv[:, d].max() + tol, (g, d, r.groups[g].delta_v,
                                                                                           dv[:, d])
        cm = [e.groups[g].crush_max for e in ens]
        assert min(cm) - crush_tol <= r.groups[g].crush_max <= max(cm) + crush_tol, (r.groups[g].crush_max, cm)
