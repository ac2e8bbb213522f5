"""The GPU kernels' f32 element arithmetic (compiled for the host) against
the f64 oracle, element by element."""
import ctypes as C

import numpy as np
import pytest

import host_kernels as hk
import oracle_ffi as of
from test_oracle_units import honeycomb_material, params, hc_update, virgin_state


def f32(a):
    return np.ascontiguousarray(a, np.float32)


def hencky32(hd, e, nu):
    tau = np.zeros(9, np.float32)
    hk.load().hk_hencky(hk.fp(f32(hd).ravel()), e, nu, hk.fp(tau))
    return tau.reshape(3, 3).astype(float)


def hencky64(F, e, nu):
    tau = np.zeros(9)
    of.load().oracle_hencky_elastic(e, nu, of.ptr(np.ascontiguousarray(F, float).ravel()), of.ptr(tau))
    return tau.reshape(3, 3)


@pytest.mark.parametrize("scale", [1e-38, 1e-30, 1e-12, 1e-8, 1e-6, 1e-4, 1e-3, 1e-2, 0.02, 0.1, 0.3])
def test_hencky_f32_matches_f64(scale):
    rng = np.random.default_rng(int(-np.log10(scale) * 10))
    e, nu = 3e8, 0.3
    for _ in range(200):
        hd = rng.normal(size=(3, 3)) * scale
        hd32 = f32(hd).astype(float)
        t32 = hencky32(hd32, e, nu)
        assert np.isfinite(t32).all()
        t64 = hencky64(np.eye(3) + hd32, e, nu)
        n = np.linalg.norm(t64)
        # f32 resolution of the strain: eps ~ 6e-8 relative to |Hd|, plus the
        # oracle's own round-off of F F^T - I at tiny strains (~1e-16 E)
        assert np.linalg.norm(t32 - t64) <= 2e-6 * n + 1e-15 * e, (scale, np.linalg.norm(t32 - t64) / max(n, 1e-300))


def test_hencky_denormal_and_degenerate_inputs_stay_finite():
    e, nu = 4.9e8, 0.3
    cases = [np.zeros((3, 3)), np.diag([1e-40, 0, 0]), np.full((3, 3), 1e-39),
             np.array([[0, 0, 0], [-6.6e-41, -1.4e-39, -6.1e-41], [1.5e-40, -2.2e-40, 1.5e-41]]),
             np.eye(3) * 1e-3, np.eye(3) * 0.2]
    for hd in cases:
        assert np.isfinite(hencky32(hd, e, nu)).all()


def test_honeycomb_f32_tracks_f64_under_crush_with_shear():
    # the reference's lane-vs-scalar check (src/kernel_simd.rs:613-662): a
    # 200-step crush with shear; compaction within 1e-3 relative, stresses
    # within 1e-6 E + 2e-3 sigma_y(c)
    for curve, dens, trans in [(None, None, 1.0),
                               ([[0.0, 7000.0], [0.05, 10000.0], [0.15, 6000.0], [0.3, 30000.0], [0.45, 30000.0]],
                                (1.0, 2.0e6), 1.0),
                               ([[0.0, 7000.0], [0.05, 10000.0], [0.15, 6000.0], [0.3, 30000.0], [0.45, 30000.0]],
                                (1.0, 2.0e6), 4.0)]:
        if curve is None:
            m = honeycomb_material(e=125.8e6, sy=1234.0, h=800339.0)
        else:
            m = honeycomb_material(e=23.1e6, curve=curve, dens=dens, trans=trans)
        rc, knots, cap = params(m)
        assert rc == 0
        e = m.youngs_modulus
        st64 = virgin_state()
        st32 = np.zeros(25, np.float32)
        st32[[9, 13, 17]] = 1.0
        L = of.load()
        for step in range(1, 201):
            c = 0.5 * step / 200
            hd = np.zeros((3, 3))
            # the element of src/kernel_simd.rs:629-634: the +x face moves by
            # (-c L, 0.02 c L), i.e. F = I + [[-c, 0, 0], [0.02 c, 0, 0], 0]
            hd[0, 0] = -c
            hd[1, 0] = 0.02 * c
            tau64 = hc_update(e, knots, cap, np.eye(3) + hd, st64)
            tau32 = np.zeros(9, np.float32)
            hk.load().hk_honeycomb(hk.fp(f32(hd).ravel()), e, hk.fp(f32(knots)), trans, hk.fp(st32), hk.fp(tau32))
            sy = np.zeros(1); h = np.zeros(1)
            L.oracle_knot_yield_at(of.ptr(knots), st64[24], of.ptr(sy), of.ptr(h))
            floor = 1e-6 * e + 2e-3 * sy[0]
            assert abs(st32[24] - st64[24]) < 1e-3 * max(st64[24], 1e-6) + 1e-6, step
            assert np.all(np.abs(st32[18:24] - st64[18:24]) < floor), step
            assert np.all(np.abs(tau32 - tau64.ravel()) < floor * 1.0), step
        assert st64[24] > 0.3


@pytest.mark.parametrize("case", ["crash", "toml"])
def test_product_element_math_inside_oracle_solve(case):
    """The GPU kernels' f32 element force (honeycomb and Hencky, relative
    displacements, packed hourglass H; compiled for the host) substituted
    for the oracle's element force in the oracle's own solve loop: whole
    solves stay within the parity bars (delta-v 0.05 m/s, crush 2 mm)."""
    import crushrs_amd as ca
    if case == "crash":
        make = lambda: ca.Model.crash(30.0, 30.0)
    else:
        a = ca.renamed(ca.vehicle_file("neon_1996"), "neon_a")
        b = ca.renamed(ca.vehicle_file("expedition_1999"), "expedition_b")
        make = lambda: ca.Model.headon(a, b, 25 * ca.MPH, 20 * ca.MPH, offset=0.3, end_time=0.08)
    o = of.run(make())
    m = make()
    ctx, fn = hk.element_hook(m)
    L = of.load()
    L.oracle_set_element_hook(fn, ctx, 3)
    try:
        r = of.run(m)
    finally:
        L.oracle_set_element_hook(None, None, 0)
        hk.load().hk_hook_free(ctx)
    assert r.eroded_elements == o.eroded_elements
    assert abs(r.steps - o.steps) <= max(1, o.steps // 100)
    for x, y in zip(r.groups, o.groups):
        assert np.all(np.abs(x.delta_v - y.delta_v) < 0.05), (x.delta_v, y.delta_v)
        assert abs(x.crush_max - y.crush_max) < 2e-3


def test_candidate_box_f32_rounding_is_outward_and_tight():
    """The search's f32 candidate boxes (f32_down/f32_up, device_common.hpp)
    bracket every f64 bound and are the nearest floats on each side
    (checked with numpy's nextafter), so the f32 box is a superset of the
    f64 box and rejects nothing the f64 box accepts."""
    L = hk.load()
    rng = np.random.default_rng(3)
    base = np.concatenate([rng.normal(0.0, 1.0, 4000), rng.normal(0.0, 1e3, 2000), rng.uniform(-1e-30, 1e-30, 500)])
    exact = rng.normal(0.0, 5.0, 500).astype(np.float32).astype(np.float64)
    near = np.concatenate([np.nextafter(exact, np.inf), np.nextafter(exact, -np.inf)])
    special = np.array([0.0, -0.0, 3.4e38, -3.4e38, 1e39, -1e39, np.inf, -np.inf, np.nan, 1e-46, -1e-46])
    x = np.ascontiguousarray(np.concatenate([base, exact, near, special]))
    dn = np.zeros(x.size, np.float32)
    up = np.zeros(x.size, np.float32)
    L.hk_f32_round(x.ctypes.data_as(C.POINTER(C.c_double)), x.size, hk.fp(dn), hk.fp(up))
    fin = np.isfinite(x)
    assert np.isnan(dn[~fin & np.isnan(x)]).all() and np.isnan(up[~fin & np.isnan(x)]).all()
    xf = x[fin]
    d, u = dn[fin].astype(np.float64), up[fin].astype(np.float64)
    assert (d <= xf).all() and (u >= xf).all()
    # tight: the next float outward from each bound lies past x
    with np.errstate(over="ignore"):
        assert (np.nextafter(dn[fin], np.float32(np.inf)).astype(np.float64) > xf).all()
        assert (np.nextafter(up[fin], np.float32(-np.inf)).astype(np.float64) < xf).all()
    # floats map to themselves
    ex = np.isin(x, exact)
    assert (dn[ex] == x[ex]).all() and (up[ex] == x[ex]).all()
    assert dn[x == np.inf][0] == np.inf and up[x == -np.inf][0] == -np.inf


def test_char_length_bitwise_equals_oracle():
    """The device's characteristic length and volume (char_length_d, which
    takes one square root of the largest squared face norm) equal the
    oracle's restatement of src/element.rs:176-186 (a square root per face)
    bit for bit, on boxes, distorted, inverted and degenerate hexes."""
    L, O = hk.load(), of.load()
    rng = np.random.default_rng(11)
    cube = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    hexes = []
    for k in range(3000):
        scale = rng.uniform(0.01, 0.5, size=3)
        h = cube * scale + rng.normal(0.0, 0.05 * scale.min() * (k % 4), size=(8, 3)) + rng.normal(0.0, 5.0, size=3)
        hexes.append(h)
    inv = cube.copy()
    inv[[4, 5, 6, 7], 2] = -1.0
    hexes += [inv, np.zeros((8, 3)), np.repeat(cube[:1], 8, axis=0), cube * 1e-160]
    P = np.ascontiguousarray(np.array(hexes).reshape(len(hexes), 24))
    n = len(hexes)
    lg, vg = np.zeros(n), np.zeros(n)
    L.hk_char_length(P.ctypes.data_as(C.POINTER(C.c_double)), n, lg.ctypes.data_as(C.POINTER(C.c_double)),
                     vg.ctypes.data_as(C.POINTER(C.c_double)))
    lo, vo = np.zeros(n), np.zeros(n)
    for e in range(n):
        l, v = np.zeros(1), np.zeros(1)
        O.oracle_char_length(of.ptr(P[e].copy()), of.ptr(l), of.ptr(v))
        lo[e], vo[e] = l[0], v[0]
    assert np.array_equal(vg.view(np.int64), vo.view(np.int64))
    assert np.array_equal(lg.view(np.int64), lo.view(np.int64))


def test_candidate_boxes_hold_every_accepted_point():
    """The product's face geometry (face_geometry, the tight candidate
    region with its margin) and the search's f32 box record contain every
    point the acceptance test of src/contact.rs:116-128 (face_accepts)
    accepts, on flat, warped, sheared and nearly degenerate faces, with
    points sampled over the faces, near their edges and behind them."""
    L = hk.load()
    rng = np.random.default_rng(5)
    n, m, max_d = 600, 400, 0.05
    quad = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
    faces = []
    for k in range(n):
        size = rng.uniform(0.02, 0.4)
        q = quad * size
        q[:, 2] += rng.normal(0.0, 0.1 * size * (k % 3), size=4)      # warp
        q[:, 0] += rng.normal(0.0, 0.2 * size * (k % 2), size=4)      # shear
        if k % 50 == 0:
            q[3] = q[2] + 1e-9 * size                                  # nearly degenerate
        th = rng.uniform(0, 2 * np.pi, 3)
        Rz = np.array([[np.cos(th[0]), -np.sin(th[0]), 0], [np.sin(th[0]), np.cos(th[0]), 0], [0, 0, 1]])
        Rx = np.array([[1, 0, 0], [0, np.cos(th[1]), -np.sin(th[1])], [0, np.sin(th[1]), np.cos(th[1])]])
        faces.append(q @ (Rz @ Rx).T + rng.normal(0.0, 3.0, 3))
    X = np.ascontiguousarray(np.array(faces).reshape(n, 12))
    # points: barycentric-like samples over the face (including past its
    # edges by the 10 % tolerance) and offsets along and against the normal
    P = np.zeros((n, m, 3))
    for f in range(n):
        q = np.array(faces[f])
        w = rng.uniform(-0.3, 1.3, size=(m, 2))
        base = (q[0] * (1 - w[:, :1]) * (1 - w[:, 1:]) + q[1] * w[:, :1] * (1 - w[:, 1:]) + q[2] * w[:, :1] * w[:, 1:]
                + q[3] * (1 - w[:, :1]) * w[:, 1:])
        nrm = np.cross(q[1] - q[0], q[2] - q[0])
        nrm /= max(np.linalg.norm(nrm), 1e-300)
        P[f] = base + nrm * rng.uniform(-1.5 * max_d, 0.5 * max_d, size=(m, 1))
    P = np.ascontiguousarray(P.reshape(-1))
    G = np.zeros(n * 31)
    acc, in64, in32 = (np.zeros(n * m, np.int32) for _ in range(3))
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    L.hk_face_check(X.ctypes.data_as(dp), n, max_d, P.ctypes.data_as(dp), m, G.ctypes.data_as(dp),
                    acc.ctypes.data_as(ip), in64.ctypes.data_as(ip), in32.ctypes.data_as(ip))
    assert acc.sum() > 0.1 * n * m              # the sample exercises the test
    assert np.all(in64[acc == 1] == 1)
    assert np.all(in32[acc == 1] == 1)
    assert np.all(in32[in64 == 1] == 1)         # the f32 record holds the f64 box
