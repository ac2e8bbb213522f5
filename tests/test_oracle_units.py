"""Pin the f64 oracle's building blocks against closed forms, invariants and
library routines (the reference's own unit tests, src/element.rs:202-248,
src/kernel.rs:355-401, src/kernel_simd.rs:664-685, src/material.rs:783-1023)."""
import ctypes as C

import numpy as np
import pytest

from crushrs_amd import _abi
import oracle_ffi as of

XI = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
               [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)


def geometry(x, e=20e6, nu=0.0, rho=200.0, hg=0.1):
    L = of.load()
    grad = np.zeros(24); vol = np.zeros(1); lump = np.zeros(8); gam = np.zeros(32); H = np.zeros(144)
    rc = L.oracle_hex_geometry(of.ptr(np.ascontiguousarray(x, float).ravel()), e, nu, rho, hg, of.ptr(grad),
                               of.ptr(vol), of.ptr(lump), of.ptr(gam), of.ptr(H))
    assert rc == 0
    return grad, vol[0], lump, gam.reshape(8, 4), H.reshape(12, 12)


def dense_k(x, e=20e6, nu=0.0, hg=0.1):
    K = np.zeros(576)
    assert of.load().oracle_dense_hourglass_k(of.ptr(np.ascontiguousarray(x, float).ravel()), e, nu, hg, of.ptr(K)) == 0
    return K.reshape(24, 24)


def hourglass_force(gam, H, u):
    um = u.reshape(8, 3).T            # 3x8
    q = (um @ gam).T.ravel()          # q[3a+d]
    r = H @ q
    return (r.reshape(4, 3).T @ gam.T).T.ravel()


def honeycomb_material(e=20e6, rho=200.0, sy=0.2e6, h=0.5e6, curve=None, dens=None, trans=1.0):
    m = _abi.Material()
    m.youngs_modulus, m.poisson_ratio, m.density, m.model = e, 0.0, rho, _abi.MAT_HONEYCOMB
    m.yield_stress, m.hardening, m.transverse_factor = sy, h, trans
    if curve is not None:
        m.n_curve = len(curve)
        for i, (c, s) in enumerate(curve):
            m.curve[i][0], m.curve[i][1] = c, s
        m.yield_stress = curve[0][1]
        dc = curve[1][0] - curve[0][0]
        m.hardening = (curve[1][1] - curve[0][1]) / dc if dc else 0.0
    if dens is not None:
        m.has_densification = 1
        m.densification[0], m.densification[1] = dens
    return m


def params(m):
    knots = np.zeros(24); cap = np.zeros(3)
    rc = of.load().oracle_honeycomb_params(C.byref(m), of.ptr(knots), of.ptr(cap))
    return rc, knots, cap


def virgin_state():
    s = np.zeros(25)
    s[[0, 4, 8, 9, 13, 17]] = 1.0
    return s


def hc_update(e, knots, cap, F, state):
    tau = np.zeros(9)
    of.load().oracle_honeycomb_update(e, of.ptr(knots), of.ptr(cap), of.ptr(np.ascontiguousarray(F, float).ravel()),
                                      of.ptr(state), of.ptr(tau))
    return tau.reshape(3, 3)


def rotation(a, b, c):
    def rx(t): return np.array([[1, 0, 0], [0, np.cos(t), -np.sin(t)], [0, np.sin(t), np.cos(t)]])
    def ry(t): return np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])
    def rz(t): return np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1]])
    return rz(c) @ ry(b) @ rx(a)


# ---------------------------------------------------------------- geometry

def test_cube_geometry_closed_form():
    # src/element.rs:203-212: V = L^3, lumped rho V / 8, dN/dx = +-1/(4L)
    x = 0.5 * 0.3 * (XI + 1.0)
    grad, vol, lump, _, _ = geometry(x, e=1e6, rho=100.0)
    assert abs(vol - 0.027) < 1e-12
    assert np.allclose(lump, 100.0 * 0.027 / 8.0, atol=1e-12)
    assert abs(grad[0] + 1.0 / 1.2) < 1e-12
    assert np.allclose(np.abs(grad), 1.0 / 1.2, atol=1e-12)


def test_rank12_equals_dense_on_distorted_hex():
    # src/element.rs:215-248
    x = np.array([0.0, 0.0, 0.0, 1.2, 0.1, 0.0, 1.1, 0.9, 0.1, -0.1, 1.0, 0.0, 0.1, 0.0, 1.0, 1.0, 0.2, 1.1,
                  1.3, 1.1, 0.9, 0.0, 0.8, 1.2]).reshape(8, 3)
    _, _, _, gam, H = geometry(x, e=20e6, nu=0.0)
    K = dense_k(x, e=20e6, nu=0.0)
    u = np.array([((i * 7 % 11) - 5.0) * 1e-3 for i in range(24)])
    f = hourglass_force(gam, H, u)
    dense = K @ u
    assert np.linalg.norm(f - dense) < 1e-9 * np.linalg.norm(dense)
    # a uniform gradient field (rotation + stretch) gives no stabilisation force
    A = rotation(0.3, -0.7, 1.1) @ np.array([[1.1, 0.05, 0.0], [0.0, 0.9, 0.02], [0.0, 0.0, 1.05]])
    u2 = (x @ A.T - x).ravel()
    assert np.linalg.norm(hourglass_force(gam, H, u2)) < 1e-8 * np.linalg.norm(dense)


def test_rank12_with_poisson_ratio():
    # the same identity with nu = 0.3 (elastic body parts)
    x = np.array([0.0, 0.0, 0.0, 1.2, 0.1, 0.0, 1.1, 0.9, 0.1, -0.1, 1.0, 0.0, 0.1, 0.0, 1.0, 1.0, 0.2, 1.1,
                  1.3, 1.1, 0.9, 0.0, 0.8, 1.2]).reshape(8, 3)
    _, _, _, gam, H = geometry(x, e=3e8, nu=0.3)
    K = dense_k(x, e=3e8, nu=0.3)
    rng = np.random.default_rng(3)
    u = rng.normal(size=24) * 1e-3
    assert np.linalg.norm(hourglass_force(gam, H, u) - K @ u) < 1e-9 * np.linalg.norm(K @ u)
    assert np.allclose(H, H.T, rtol=0, atol=1e-9 * np.abs(H).max())


def test_char_length_box():
    L = of.load()
    x = 0.5 * (XI + 1.0) * np.array([0.29, 0.285, 0.27])
    l = np.zeros(1); v = np.zeros(1)
    L.oracle_char_length(of.ptr(x.ravel().copy()), of.ptr(l), of.ptr(v))
    assert abs(v[0] - 0.29 * 0.285 * 0.27) < 1e-14
    assert abs(l[0] - 0.27) < 1e-14  # V / A_max with A_max = 0.29 * 0.285
    # inverted -> infinite length
    xi = x.copy(); xi[:, 2] *= -1
    L.oracle_char_length(of.ptr(xi.ravel().copy()), of.ptr(l), of.ptr(v))
    assert np.isinf(l[0])


# ---------------------------------------------------------------- materials

def test_knot_table_values():
    # src/kernel_simd.rs:672-680
    m = honeycomb_material(e=23.1e6, rho=100.0, curve=[[0.0, 7000.0], [0.05, 10000.0], [0.15, 6000.0],
                                                       [0.3, 30000.0], [0.45, 30000.0]], dens=(1.0, 2.0e6))
    rc, knots, cap = params(m)
    assert rc == 0
    L = of.load()
    sy = np.zeros(1); h = np.zeros(1)
    for c, want in [(0.0, (7000.0, 60000.0)), (0.1, (8000.0, -40000.0)), (0.6, (30000.0, 0.0)),
                    (1.5, (30000.0 + 1.0e6, 2.0e6))]:
        L.oracle_knot_yield_at(of.ptr(knots), c, of.ptr(sy), of.ptr(h))
        assert abs(sy[0] - want[0]) < 1e-9 * max(abs(want[0]), 1.0)
        assert abs(h[0] - want[1]) < 1e-9 * max(abs(want[1]), 1.0)
        sy2 = np.zeros(1); h2 = np.zeros(1)
        L.oracle_plasticity_yield_at(C.byref(m), c, of.ptr(sy2), of.ptr(h2))
        if c in (0.1, 1.5):  # Plasticity::yield_at agrees (src/kernel_simd.rs:679-680)
            assert abs(sy2[0] - sy[0]) < 1e-9 * abs(sy[0]) and abs(h2[0] - h[0]) < 1e-9 * max(abs(h[0]), 1.0)


def test_invalid_curves_rejected():
    # Plasticity::check (src/material.rs:116-151)
    assert params(honeycomb_material(curve=[[0.1, 7000.0], [0.2, 8000.0]]))[0] != 0      # must start at 0
    assert params(honeycomb_material(curve=[[0.0, 7000.0], [0.0, 8000.0]]))[0] != 0      # increasing
    assert params(honeycomb_material(curve=[[0.0, 7000.0], [0.2, -1.0]]))[0] != 0        # positive
    eight = [[0.1 * i, 1000.0 + i] for i in range(8)]
    assert params(honeycomb_material(curve=eight))[0] == 0
    assert params(honeycomb_material(curve=eight, dens=(2.0, 1e6)))[0] != 0               # 8 + lock > 8
    assert params(honeycomb_material(curve=[[0.0, 7000.0], [0.5, 8000.0]], dens=(0.4, 1e6)))[0] != 0


def _single_element_force(x, u, kind, e, nu, knots, cap, state):
    grad, vol, _, gam, H = geometry(x, e=e, nu=nu)
    out = np.zeros(24)
    ok = of.load().oracle_element_force(of.ptr((x + u.reshape(8, 3)).ravel().copy()), of.ptr(u.copy()),
                                        of.ptr(grad), vol, of.ptr(gam.ravel().copy()), of.ptr(H.ravel().copy()),
                                        kind, e, nu, of.ptr(knots), of.ptr(cap), of.ptr(state), of.ptr(out))
    return ok, out


def test_rigid_motion_gives_zero_force():
    # src/kernel.rs:356-366 (honeycomb; the elastic Hencky branch too)
    x = 0.5 * (XI + 1.0)
    R = rotation(0.4, 1.2, -0.8)
    u = (x @ R.T + np.array([3.0, -1.0, 2.0]) - x).ravel()
    rc, knots, cap = params(honeycomb_material())
    ok, f = _single_element_force(x, u, _abi.MAT_HONEYCOMB, 20e6, 0.0, knots, cap, virgin_state())
    assert ok and np.linalg.norm(f) < 1e-3
    ok, f = _single_element_force(x, u, _abi.MAT_ELASTIC, 200e9, 0.3, knots, cap, virgin_state())
    assert ok and np.linalg.norm(f) < 1e-3


def test_uniform_crush_matches_material_law():
    # src/kernel.rs:369-386: 40 % crush in 100 steps, face force = -sigma_y(c) A,
    # compaction = -ln 0.6 - sigma_y / E
    x = 0.5 * (XI + 1.0)
    rc, knots, cap = params(honeycomb_material())
    st = virgin_state()
    for k in range(1, 101):
        A = np.diag([1 - 0.4 * k / 100, 1.0, 1.0])
        ok, f = _single_element_force(x, (x @ A.T - x).ravel(), _abi.MAT_HONEYCOMB, 20e6, 0.0, knots, cap, st)
    fx = sum(f[3 * i] for i in (1, 2, 5, 6))
    c = st[24]
    sy = 0.2e6 + 0.5e6 * c
    assert abs(fx + sy) < 1e-6 * sy
    assert abs(c - (-np.log(0.6) - sy / 20e6)) < 2e-3


def test_hourglass_mode_is_resisted():
    # src/kernel.rs:389-401
    x = 0.5 * (XI + 1.0)
    rc, knots, cap = params(honeycomb_material())
    u = np.zeros(24)
    u[0::3] = 1e-3 * np.array([1, -1, 1, -1, 1, -1, 1, -1])
    st = virgin_state()
    ok, f = _single_element_force(x, u, _abi.MAT_HONEYCOMB, 20e6, 0.0, knots, cap, st)
    assert np.linalg.norm(f) > 1.0
    assert st[24] == 0.0


def test_inverted_element_is_eroded():
    x = 0.5 * (XI + 1.0)
    rc, knots, cap = params(honeycomb_material())
    u = np.zeros(24)
    u[2::3] = np.where(XI[:, 2] > 0, -2.0, 0.0)  # top face pushed through the bottom
    ok, f = _single_element_force(x, u, _abi.MAT_HONEYCOMB, 20e6, 0.0, knots, cap, virgin_state())
    assert not ok and not f.any()


def test_honeycomb_is_objective_under_superposed_rotation():
    # src/material.rs:998-1023
    rc, knots, cap = params(honeycomb_material())
    st = virgin_state()
    f0 = np.diag([0.7, 1.0, 1.0])
    F = np.eye(3)
    for k in range(1, 101):
        F = np.eye(3) + (f0 - np.eye(3)) * (k / 100)
        hc_update(20e6, knots, cap, F, st)
    tau0 = hc_update(20e6, knots, cap, F, st.copy())
    c0 = st[24]
    axis = np.array([0.3, -0.5, 0.8]); axis /= np.linalg.norm(axis)

    def rot(t):
        K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
        return np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * K @ K

    for k in range(1, 201):
        tau = hc_update(20e6, knots, cap, rot(1.2 * k / 200) @ F, st)
    Rt = rot(1.2)
    assert np.linalg.norm(tau - Rt @ tau0 @ Rt.T) < 1e-4 * np.linalg.norm(tau0)
    assert abs(st[24] - c0) < 1e-6


def test_honeycomb_repeat_is_idempotent_and_unloading_elastic():
    # src/material.rs:985-995 (the honeycomb half of the foam comparison)
    rc, knots, cap = params(honeycomb_material())
    st = virgin_state()
    F = np.eye(3)
    for k in range(1, 201):
        F = np.diag([1 - 0.6 * k / 200, 1.0, 1.0])
        tau = hc_update(20e6, knots, cap, F, st)
    j = np.linalg.det(F)
    sy = 0.2e6 + 0.5e6 * st[24]
    assert abs(tau[0, 0] / j + sy) < 1e-6 * sy        # Cauchy stress on the hardened cap
    assert abs(tau[1, 1]) < 1e-6 * abs(tau[0, 0])
    # the Hencky foam's compaction for 60 % crush: -ln 0.4 - sigma_y / E
    assert abs(st[24] - (-np.log(0.4) - sy / 20e6)) < 2e-2 * st[24]
    c = st[24]
    tau2 = hc_update(20e6, knots, cap, F, st)
    assert np.linalg.norm(tau2 - tau) < 1e-9 * np.linalg.norm(tau) and st[24] == c
    F2 = F.copy(); F2[0, 0] += 0.01
    tau3 = hc_update(20e6, knots, cap, F2, st)
    assert tau3[0, 0] > tau[0, 0] and st[24] == c


def test_eigen3_matches_numpy():
    # src/material.rs:886-925 (closed form vs a library eigensolver)
    L = of.load()
    rng = np.random.default_rng(12345)
    vals = np.zeros(3); vecs = np.zeros(9)

    def check(a):
        L.oracle_eigen3(of.ptr(np.ascontiguousarray(a).ravel()), of.ptr(vals), of.ptr(vecs))
        v = vecs.reshape(3, 3)
        scale = max(np.linalg.norm(a), 1e-300)
        assert np.linalg.norm(v @ np.diag(vals) @ v.T - a) <= 1e-12 * scale
        assert np.linalg.norm(v.T @ v - np.eye(3)) < 1e-12
        assert vals[0] <= vals[1] <= vals[2]
        assert np.all(np.abs(vals - np.linalg.eigvalsh(a)) <= 1e-12 * scale)

    for k in range(3000):
        s = 10.0 ** ((k % 13) - 6)
        m = rng.uniform(-1, 1, (3, 3))
        check((m + m.T) * s)
    for k in range(600):
        e = 10.0 ** (-(k % 12))
        m = rng.uniform(-1, 1, (3, 3))
        check(np.eye(3) + (m + m.T) * e)
        R = rotation(*rng.uniform(-1, 1, 3))
        check(R @ np.diag([2.0, 2.0 + e * rng.uniform(-1, 1), 0.5]) @ R.T)
        check(R @ np.diag([0.3, 1.0, 1.0]) @ R.T)
    check(np.eye(3)); check(np.zeros((3, 3))); check(np.diag([3.0, -1.0, 2.0]))


def test_hencky_pressure_and_objectivity():
    # src/material.rs:878-884 (p = K ln J) and objectivity of the elastic branch
    L = of.load()
    e, nu = 20e6, 0.3
    k = e / (3 * (1 - 2 * nu))
    tau = np.zeros(9)
    F = np.eye(3) * 0.5 ** (1 / 3)
    L.oracle_hencky_elastic(e, nu, of.ptr(F.ravel().copy()), of.ptr(tau))
    assert abs(np.trace(tau.reshape(3, 3)) / 3 - k * np.log(0.5)) < 1e-6 * k
    F = np.array([[0.7, 0.2, 0.0], [0.1, 1.2, 0.05], [0.0, 0.0, 1.15]])
    R = rotation(0.4, -1.1, 2.0)
    t1 = np.zeros(9); t2 = np.zeros(9)
    L.oracle_hencky_elastic(e, nu, of.ptr(F.ravel().copy()), of.ptr(t1))
    L.oracle_hencky_elastic(e, nu, of.ptr((R @ F).ravel().copy()), of.ptr(t2))
    t1 = t1.reshape(3, 3); t2 = t2.reshape(3, 3)
    assert np.linalg.norm(R @ t1 @ R.T - t2) < 1e-9 * np.linalg.norm(t1)
    # small-strain limit: tau = lambda tr(eps) I + 2 mu eps
    h = np.array([[2e-5, 3e-6, 0.0], [3e-6, -1e-5, 1e-6], [0.0, 1e-6, -5e-6]])
    t1 = np.zeros(9)
    L.oracle_hencky_elastic(e, nu, of.ptr((np.eye(3) + h).ravel().copy()), of.ptr(t1))
    lam = e * nu / ((1 + nu) * (1 - 2 * nu)); mu = e / (2 * (1 + nu))
    lin = lam * np.trace(h) * np.eye(3) + 2 * mu * h
    assert np.linalg.norm(t1.reshape(3, 3) - lin) < 1e-4 * np.linalg.norm(lin)


def _penetrated_crash(push=0.05):
    """crash 30 30 with the Neon block pushed `push` m into the Silverado."""
    import crushrs_amd as ca
    m = ca.Model.crash(30.0, 30.0)
    x = m.nodes()
    u = np.zeros_like(x)
    neon = m.node_set("neon")
    u[neon, 0] -= push
    return m, u


def _lumped(m):
    """Lumped nodal masses (src/model.rs:268-278) from a one-step oracle solve."""
    import crushrs_amd as ca
    s = m.settings
    end = s.end_time
    s.end_time = 1e-12
    m.settings = s
    r = of.run(m)
    s.end_time = end
    m.settings = s
    return r.masses


def _contact(m, u, dt=1e-4):
    f = np.zeros(3 * m.n_nodes)
    hit = np.full(max(m.contact_node_count(), 1), -2, np.int32)
    masses = np.ascontiguousarray(_lumped(m))
    n = of.load().oracle_contact_apply(m.ptr, of.ptr(np.ascontiguousarray(u.reshape(-1))), of.ptr(masses), dt,
                                       of.ptr(f), of.ptr(hit))
    return n, f.reshape(-1, 3), hit


def test_contact_forces_balance_and_only_penetrating_nodes_push():
    """src/contact.rs:145-154: each node force +f n is balanced by -f n/4 on
    the four face nodes, so the contact forces sum to zero; untouched nodes
    get none."""
    m, u = _penetrated_crash(0.05)
    n, f, hit = _contact(m, u)
    assert n > 0
    total = f.sum(axis=0)
    assert np.all(np.abs(total) < 1e-9 * np.abs(f).sum())
    # force limit m * 1e5 (src/contact.rs:144) and penalty direction
    assert np.all(np.isfinite(f))
    assert (hit >= -1).all()


def test_contact_first_hit_is_lowest_face_index_brute_force():
    """src/contact.rs:113-157: for a penetrating node the face that takes
    it is the first face in order that accepts it.  The oracle's choice is
    re-derived here by brute force over every face with the same tests
    (plane distance within [-max_d, 0], projection inside one of the four
    triangles of either diagonal split, tolerance 0.1)."""
    m, u = _penetrated_crash(0.05)
    n, f, hit = _contact(m, u)
    v = m.view
    x = m.nodes() + u
    faces = np.ctypeslib.as_array(v.faces, shape=(v.n_faces, 4))
    off = 0
    checked = 0
    for ci in range(v.n_contacts):
        c = v.contacts[ci]
        cnodes = np.ctypeslib.as_array(c.nodes, shape=(c.n_nodes,))
        cfaces = np.ctypeslib.as_array(c.faces, shape=(c.n_faces,))
        # nodes kept by the static broad phase (src/contact.rs:48-70)
        x0 = m.nodes()
        fx = x0[faces[cfaces].reshape(-1)]
        lo, hi = fx.min(0) - c.max_distance, fx.max(0) + c.max_distance
        kept = [nd for nd in cnodes if np.all(x0[nd] >= lo) and np.all(x0[nd] <= hi)]
        for j, nd in enumerate(kept):
            want = -1
            for fi, face in enumerate(cfaces):
                q = x[faces[face]]
                nrm = np.cross(q[1] - q[0], q[2] - q[0])
                nrm /= np.linalg.norm(nrm)
                ctr = q.mean(0)
                bl, bh = q.min(0) - c.max_distance, q.max(0) + c.max_distance
                p = x[nd]
                if np.any(p < bl) or np.any(p > bh):
                    continue
                s = nrm @ (p - ctr)
                if not (-c.max_distance <= s <= 0.0):
                    continue
                pp = p - s * nrm
                inside = False
                for a, b, cc in ((0, 1, 2), (0, 2, 3), (1, 2, 3), (1, 3, 0)):
                    v0, v1, v2 = q[cc] - q[a], q[b] - q[a], pp - q[a]
                    d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
                    d20, d21 = v2 @ v0, v2 @ v1
                    den = d00 * d11 - d01 * d01
                    uu = (d11 * d20 - d01 * d21) / den
                    vv = (d00 * d21 - d01 * d20) / den
                    if uu >= -0.1 and vv >= -0.1 and uu + vv <= 1.1:
                        inside = True
                        break
                if inside:
                    want = fi
                    break
            got = hit[off + j]
            if want >= 0 or got >= 0:
                checked += 1
            assert got == want, (ci, j, nd, got, want)
        off += len(kept)
    assert checked > 0
