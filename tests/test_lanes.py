"""The AVX2 lane restatement of the reference's default element path
(src/kernel_simd.rs), which is the timed CPU baseline, pinned against the
f64 oracle: the reference's own lane-vs-scalar test
(src/kernel_simd.rs:605-685) and whole solves."""
import ctypes as C

import numpy as np
import pytest

import crushrs_amd as ca
import oracle_ffi as of
from test_oracle_units import XI, hc_update, honeycomb_material, params, virgin_state


def _lane_vs_scalar(m):
    """A crushing cube element stepped 200 times: the f32 lane kernel must
    track the f64 honeycomb update to single precision in compaction,
    stress and nodal force (src/kernel_simd.rs:605-653)."""
    L = of.load()
    rc, knots, cap = params(m)
    assert rc == 0
    e = m.youngs_modulus
    cube = 0.15 * (XI + 1.0)
    grad = (XI / (4.0 * 0.3)).ravel()
    zeros32, zeros144 = np.zeros(32), np.zeros(144)
    conn = np.arange(8, dtype=np.int32)
    blk = L.lane_blocks_create(1)
    try:
        L.lane_blocks_set(blk, 0, of.ptr(conn), of.ptr(grad), 0.027, of.ptr(zeros32), of.ptr(zeros144), e,
                          of.ptr(knots), of.ptr(cap), of.ptr(virgin_state()), 0)
        state = virgin_state()
        lane = np.zeros(25)
        pos = np.ascontiguousarray(cube.ravel())
        for step in range(1, 201):
            c = 0.5 * step / 200.0
            u = np.zeros(24)
            for i in (1, 2, 5, 6):
                u[3 * i] = -c * 0.3
                u[3 * i + 1] = 0.02 * c * 0.3
            glob = np.zeros(24)
            assert L.lane_blocks_forces(blk, of.ptr(pos), of.ptr(u), of.ptr(glob), 1) == 0
            x = cube + u.reshape(8, 3)
            F = x.T @ grad.reshape(8, 3)
            tau = hc_update(e, knots, cap, F, state)
            mm = tau @ np.linalg.inv(F).T * 0.027
            L.lane_blocks_state(blk, 0, of.ptr(lane))
            sy = np.zeros(1); h = np.zeros(1)
            L.oracle_knot_yield_at(of.ptr(knots), state[24], of.ptr(sy), of.ptr(h))
            floor = 1e-6 * e + 2e-3 * sy[0]
            assert abs(lane[24] - state[24]) < 1e-3 * max(state[24], 1e-6) + 1e-6, (step, lane[24], state[24])
            assert np.all(np.abs(lane[18:24] - state[18:24]) < floor), (step, lane[18:24], state[18:24])
            want = (mm @ grad.reshape(8, 3).T).T.ravel()
            assert np.all(np.abs(glob - want) < floor * 0.027 / 0.3), (step, np.abs(glob - want).max())
        assert state[24] > 0.3
    finally:
        L.lane_blocks_free(blk)


def test_lane_kernel_matches_scalar_reference():
    # src/kernel_simd.rs:655-658
    _lane_vs_scalar(honeycomb_material(e=125.8e6, sy=1234.0, h=800339.0))


@pytest.mark.parametrize("trans", [1.0, 4.0])
def test_lane_kernel_matches_scalar_with_tabulated_curve(trans):
    # src/kernel_simd.rs:660-684: plateau, dip, steep rise, flat, lock-up
    m = honeycomb_material(e=23.1e6, rho=100.0, curve=[[0.0, 7000.0], [0.05, 10000.0], [0.15, 6000.0],
                                                       [0.3, 30000.0], [0.45, 30000.0]], dens=(1.0, 2.0e6),
                           trans=trans)
    _lane_vs_scalar(m)


def test_lane_block_padding_and_inversion():
    """Padding lanes of a partly filled block carry no force, and a lane
    whose element inverts is eroded once and then stays at zero force
    (src/kernel_simd.rs:324-334,486-490)."""
    L = of.load()
    rc, knots, cap = params(honeycomb_material())
    assert rc == 0
    cube = 0.15 * (XI + 1.0)
    grad = (XI / (4.0 * 0.3)).ravel()
    n_el = 3
    pos = np.tile(cube.ravel(), n_el)
    blk = L.lane_blocks_create(n_el)
    try:
        for k in range(n_el):
            conn = (np.arange(8) + 8 * k).astype(np.int32)
            L.lane_blocks_set(blk, k, of.ptr(conn), of.ptr(grad), 0.027, of.ptr(np.zeros(32)), of.ptr(np.zeros(144)),
                              20e6, of.ptr(knots), of.ptr(cap), of.ptr(virgin_state()), 0)
        u = np.zeros(24 * n_el)
        u[24 + 3 * 1] = u[24 + 3 * 2] = u[24 + 3 * 5] = u[24 + 3 * 6] = -0.5  # element 1 turned inside out
        u[48 + 3 * 1] = -0.01                                              # element 2 mildly loaded
        glob = np.zeros(24 * n_el)
        assert L.lane_blocks_forces(blk, of.ptr(pos), of.ptr(u), of.ptr(glob), 1) == 1
        assert [L.lane_blocks_eroded(blk, k) for k in range(n_el)] == [0, 1, 0]
        assert np.all(glob[:24] == 0.0) and np.all(glob[24:48] == 0.0) and np.abs(glob[48:]).max() > 0.0
        glob[:] = 0.0
        assert L.lane_blocks_forces(blk, of.ptr(pos), of.ptr(u), of.ptr(glob), 2) == 0
        assert np.all(glob[24:48] == 0.0)
    finally:
        L.lane_blocks_free(blk)


@pytest.mark.parametrize("threads", [1, 4])
def test_lane_solve_matches_f64_oracle_crash(threads):
    """crash 30 30 on the lane path: delta-v within 0.05 m/s and crush
    within 2 mm of the f64 oracle; the thread count changes nothing."""
    m = ca.Model.crash(30.0, 30.0)
    o = of.run(m)
    r = of.run_lanes(m, threads)
    assert abs(r.steps - o.steps) <= max(2, o.steps // 100)
    for a, b in zip(r.groups, o.groups):
        assert abs(a.delta_v[0] - b.delta_v[0]) < 0.05, (a.delta_v, b.delta_v)
        assert abs(a.crush_max - b.crush_max) < 2e-3
    if threads > 1:
        r1 = of.run_lanes(m, 1)
        assert r1.steps == r.steps and np.array_equal(r1.velocity, r.velocity)


def test_lane_solve_with_elastic_parts_matches_f64_oracle():
    """A TOML head-on with bumper/body (elastic, scalar path) and road: the
    mixed lane + scalar force pass stays within 0.05 m/s of the oracle."""
    a = ca.vehicle("neon")
    b = ca.renamed(ca.vehicle("expedition"), "expedition_b")
    m = ca.Model.headon(a, b, 25.0 * ca.MPH, 20.0 * ca.MPH).set(end_time=0.06)
    o = of.run(m)
    r = of.run_lanes(m, 2)
    for x, y in zip(r.groups, o.groups):
        assert abs(x.delta_v[0] - y.delta_v[0]) < 0.05, (x.delta_v, y.delta_v)
