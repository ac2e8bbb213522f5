"""The synthetic workloads' builders (CPU): the large single instance has
the sizes and knot table it is described with, and the calibration set
perturbs each test's crush curve and rail box as stated."""
import ctypes as C

import numpy as np

import crushrs_amd as ca
from crushrs_amd import _abi, workloads
import oracle_ffi as of


def test_large_pair_sizes_and_knots():
    m = workloads.large_pair()
    # 78x27x25 + 58x23x18 hexes of 0.075 m, (79x28x26) + (59x24x19) nodes
    assert m.n_hexes == 78 * 27 * 25 + 58 * 23 * 18 == 76662
    assert m.n_nodes == 79 * 28 * 26 + 59 * 24 * 19 == 84416
    assert m.settings.mass_scaling == 0.25
    assert m.view.n_contacts == 2  # no road
    v = m.view
    for p in range(v.n_parts):
        mat = v.materials[p]
        assert mat.model == _abi.MAT_HONEYCOMB
        knots, cap = np.zeros(24), np.zeros(3)
        assert of.load().oracle_honeycomb_params(C.byref(mat), of.ptr(knots), of.ptr(cap)) == 0
        k = knots.reshape(8, 3)
        # 7 tabulated knots plus the lock-up knot: all 8 in use
        assert np.all(np.isfinite(k[:, 0])) and np.all(np.diff(k[:, 0]) > 0)
        assert k[7, 0] == 1.4
        # lock-up slope = last slope + E (src/vehicle.rs:486, src/material.rs:444-447)
        assert abs(k[7, 2] - (k[6, 2] + mat.youngs_modulus)) < 1e-6 * mat.youngs_modulus
    # the tabulated curve is the preset's bilinear force F = F0 + k x at each knot:
    # sigma = F / A_front, compaction = -ln(1 - x / L_zone) (src/vehicle.rs:509-511)
    a = ca.vehicle("silverado")
    area = a.size[1] * a.size[2]
    mat = v.materials[0]
    for i, x in enumerate(workloads.LARGE_KNOTS_M):
        assert abs(mat.curve[i][0] - (-np.log(1.0 - x / a.crush_zone_length))) < 1e-12
        assert abs(mat.curve[i][1] - (a.curve[0] + a.curve[1] * x) / area) < 1e-9 * mat.curve[i][1]


def test_calibration_set_perturbs_curves_and_rails():
    cases = workloads.calibration_set(64, seed=3)
    assert all(0.8 <= c.yield_scale <= 1.25 for c in cases)
    assert all(0.9 <= k <= 1.1 for c in cases for k in c.knot_scales)
    assert all(0.5 <= c.rail_fraction <= 0.8 and 0.4 <= c.rail_width <= 0.8 for c in cases)
    assert len({c.vehicle for c in cases}) == 4
    # same seed, same set
    assert workloads.calibration_set(64, seed=3) == cases
    c = next(c for c in cases if c.vehicle == "neon_1996")
    base = workloads.BarrierCase(c.vehicle, 1.0, (1.0,) * 8, c.rail_fraction, c.rail_width).build()
    pert = c.build()
    assert pert.n_hexes == base.n_hexes
    names = pert.part_names()
    crush = names.index("neon_1996_crush")
    mb, mp = base.view.materials[crush], pert.view.materials[crush]
    for i in range(mb.n_curve):
        want = mb.curve[i][1] * c.yield_scale * c.knot_scales[i]
        assert abs(mp.curve[i][1] - want) < 1e-9 * want
    # a bilinear vehicle scales its yield force only
    s = next(c for c in cases if c.vehicle == "silverado_2007")
    mbase, mpert = workloads.BarrierCase(s.vehicle, 1.0, (1.0,) * 8, 0.6, 0.6).build(), s.build()
    sb, sp = mbase.view.materials[0], mpert.view.materials[0]  # views into the live models
    assert abs(sp.yield_stress - sb.yield_stress * s.yield_scale) < 1e-9 * sp.yield_stress
    assert sp.hardening == sb.hardening


def test_calibration_case_at_bifurcation_flips_with_bumper_modulus():
    """Instance 0 of calibration_set(12, seed=5) (Expedition barrier, yield
    x 1.147) is ill-conditioned in the oracle itself: scaling the bumper's
    Young's modulus by 1 + 1e-6 switches the solve from the branch where
    every element survives to the one where an element behind the bumper
    collapses during the rebound and erodes (about 10 % more steps, crush
    up by ~5 mm, delta-v by ~0.03 m/s).  The GPU parity test for this case
    therefore checks the GPU against the oracle's envelope."""
    from concurrent.futures import ThreadPoolExecutor
    case = workloads.calibration_set(12, seed=5)[0]

    def one(scale):
        m = case.build()
        bumper = m.part_names().index("expedition_1999_bumper")
        m.view.materials[bumper].youngs_modulus *= scale
        return of.run(m)

    with ThreadPoolExecutor(2) as ex:
        base, bumped = list(ex.map(one, [1.0, 1.0 + 1e-6]))
    assert base.eroded_elements == 0 and bumped.eroded_elements == 1
    assert bumped.steps > 1.05 * base.steps
    assert abs(bumped.groups[0].delta_v[0] - base.groups[0].delta_v[0]) < 0.05


def test_headon_toml_case_is_ill_conditioned_in_the_oracle():
    """The TOML head-on of the parity tests (Neon vs Expedition at 25/20
    mph, 0.3 m offset, road) sits at a bifurcation in the oracle itself:
    scaling every elastic part's Young's modulus by 1 + 1e-9 makes one
    crush-zone element erode that survives unperturbed, and the step count
    moves.  The GPU parity test for the full solve therefore checks the GPU
    against the oracle's envelope; delta-v stays well-conditioned."""
    from concurrent.futures import ThreadPoolExecutor
    from test_gpu_parity import headon_toml

    def one(scale):
        m = headon_toml()
        v = m.view
        for p in range(v.n_parts):
            if v.materials[p].model == _abi.MAT_ELASTIC:
                v.materials[p].youngs_modulus *= scale
        return of.run(m)

    with ThreadPoolExecutor(2) as ex:
        base, bumped = list(ex.map(one, [1.0, 1.0 + 1e-9]))
    assert base.eroded_elements == 0 and bumped.eroded_elements == 1
    assert base.steps != bumped.steps
    for a, b in zip(base.groups, bumped.groups):
        assert abs(a.delta_v[0] - b.delta_v[0]) < 0.01
