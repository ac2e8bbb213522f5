"""Histories, frames and their output formats on the CPU (the oracle's
restatement of src/solver.rs:114-187, 265-357 and the writers of
crushrs_amd.output, src/output/history.rs and src/output/vtk.rs).

Pins: momentum conservation of a model without external forces, the last
history sample against the final state, the reference's own history test
(tests/impact.rs:148-189: the accelerometer's acceleration integrates to
its velocity change within 2 %, the frame starts on the global axes and
stays close in a head-on, the Parquet node table has 25 columns and one
row per sample), its VTK test (tests/impact.rs:133-146: one DataSet per
frame), and closed forms of the frame stresses (a honeycomb cube crushed
uniaxially carries -(yield + hardening x compaction) along the crush; an
elastic cube's Cauchy stress is the Hencky stress of its deformation
gradient over J)."""
import ctypes as C
import math

import numpy as np
import pytest

import crushrs_amd as ca
from crushrs_amd import _abi, output
import oracle_ffi as of


def part_node_lists(m):
    hexes = m.hexes()
    hp = np.ctypeslib.as_array(m.view.hex_part, shape=(m.n_hexes,))
    return [np.unique(hexes[hp == p].ravel()) for p in range(m.view.n_parts)]


def test_history_samples_conserve_momentum_and_end_on_the_final_state():
    """crash 30 30 has no road or wall: the parts only push on each other,
    so the sum of part masses x mean velocities is the initial momentum at
    every sample.  Samples are taken at steps 0, 5, 10, ... and the last;
    the last equals the part means of the final state."""
    m = ca.Model.crash(30.0, 30.0).set(end_time=0.02, history_steps=5)
    r = of.run(m)
    steps = list(r.history_step)
    want = list(range(0, r.steps + 1, 5)) + ([r.steps] if r.steps % 5 else [])
    assert steps == want
    assert r.history_time[0] == 0.0 and r.history_time[-1] == r.time
    P = r.parts
    mom = (P[:, :, 0:1] * P[:, :, 4:7]).sum(axis=1)
    p0 = (r.masses[:, None] * m.initial_velocity()).sum(axis=0)
    assert np.abs(mom - p0).max() < 1e-9 * np.abs(p0).max()
    for p, nodes in enumerate(part_node_lists(m)):
        mass = r.masses[nodes].sum()
        assert abs(P[-1, p, 0] - mass) < 1e-12 * mass
        v = (r.masses[nodes, None] * r.velocity[nodes]).sum(0) / mass
        u = (r.masses[nodes, None] * r.displacement[nodes]).sum(0) / mass
        assert np.allclose(P[-1, p, 4:7], v, rtol=0, atol=1e-12 * max(1.0, np.abs(v).max()))
        assert np.allclose(P[-1, p, 1:4], u, rtol=0, atol=1e-12)
        ke = 0.5 * (r.masses[nodes] * (r.velocity[nodes] ** 2).sum(1)).sum()
        assert abs(P[-1, p, 10] - ke) < 1e-9 * ke


def test_accelerometer_history_integrates_to_delta_v_and_writes_parquet(tmp_path):
    """tests/impact.rs:148-189 with the crash 30 30 model: every step is a
    sample; the Neon accelerometer's trapezoid-integrated x acceleration
    equals its x velocity change within 2 % (and it slows by more than 3
    m/s); its frame starts on the global axes and local x stays within
    0.99 of global x; local = axes x global; the node table has 25 columns
    and a row per sample, the part table a row per sample and part."""
    pq = pytest.importorskip("pyarrow.parquet")
    m = ca.Model.crash(30.0, 30.0).set(end_time=0.06, history_steps=1)
    r = of.run(m)
    assert len(r.history_time) == r.steps + 1
    a = r.accel[:, 1]  # the Neon (group 1)
    t = r.history_time
    dv = np.sum(0.5 * (a[1:, 9] + a[:-1, 9]) * np.diff(t))
    dv_node = a[-1, 6] - a[0, 6]
    assert dv_node > 3.0
    assert abs(dv - dv_node) < 0.02 * abs(dv_node)
    axes = a[:, 15:24].reshape(-1, 3, 3)
    assert abs(axes[0, 0, 0] - 1.0) < 1e-9 and abs(axes[0, 1, 1] - 1.0) < 1e-9
    assert axes[-1, 0, 0] > 0.99
    assert np.allclose(a[:, 0:3], m.nodes()[m.view.groups[1].accel_node] + a[:, 3:6], atol=1e-12)
    paths = output.write_history(m, r, tmp_path / "run")
    assert [p.name for p in paths] == ["run.nodes.parquet", "run.frames.parquet", "run.parts.parquet"]
    nodes = pq.read_table(paths[0])
    assert nodes.num_columns == 25 and nodes.num_rows == len(t) * 2
    assert nodes.column_names[:4] == ["time", "step", "accelerometer", "node"]
    assert set(nodes.column("accelerometer").to_pylist()) == {"silverado_rear", "neon_rear"}
    la = np.stack([nodes.column(c).to_numpy() for c in ("lax", "lay", "laz")], 1)[1::2]
    assert np.allclose(la, np.einsum("kij,kj->ki", axes, a[:, 9:12]), rtol=0, atol=1e-9 * np.abs(la).max())
    parts = pq.read_table(paths[2])
    assert parts.num_rows == len(t) * m.view.n_parts and parts.num_columns == 13
    frames = pq.read_table(paths[1])
    assert frames.num_rows == len(t) * 2 and frames.num_columns == 15


def test_frames_and_vtk_series(tmp_path):
    """tests/impact.rs:133-146: frames at step 0, every frame_steps steps
    and the last; the .pvd lists one DataSet per frame; each .vtu holds the
    deformed points and the frame's fields; the last frame is the final
    state."""
    m = ca.Model.crash(30.0, 30.0).set(end_time=0.02, frame_steps=20)
    r = of.run(m)
    n = 1 + r.steps // 20 + (1 if r.steps % 20 else 0)
    assert len(r.frame_time) == n and r.frame_time[0] == 0.0 and r.frame_time[-1] == r.time
    assert np.array_equal(r.frame_displacement[-1], r.displacement.astype(np.float32))
    # frame 0 follows the initial force pass: round-off stresses only
    assert not r.frame_displacement[0].any() and np.abs(r.frame_stress[0]).max() < 1e-3
    pvd = output.write_series(m, r, tmp_path / "crash")
    text = pvd.read_text()
    assert text.count("<DataSet") == n
    assert (tmp_path / "crash_0000.vtu").exists()
    back = output.read_vtu(tmp_path / f"crash_{n - 1:04d}.vtu")
    X = m.nodes().astype(np.float32)
    assert np.array_equal(back["Points"].reshape(-1, 3), X + r.frame_displacement[-1])
    assert np.array_equal(back["stress"].reshape(-1, 6), r.frame_stress[-1])
    assert np.array_equal(back["connectivity"].reshape(-1, 8), m.hexes())
    assert back["types"].tolist() == [12] * m.n_hexes


def single_cube(material, v0, fixed_bottom: bool, end_time: float, frame_steps: int = 10):
    """A 0.3 m cube (one hexahedron) with initial velocities v0 [8, 3]."""
    xi = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                   [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    nodes = np.ascontiguousarray(0.15 * (xi + 1.0))
    hexes = np.arange(8, dtype=np.int32)
    part = np.zeros(1, np.int32)
    iv = np.ascontiguousarray(v0, dtype=float)
    fixed = np.arange(4, dtype=np.int32)
    mat = (_abi.Material * 1)()
    mat[0] = material
    mv = _abi.ModelView()
    mv.n_nodes, mv.n_hexes, mv.n_parts = 8, 1, 1
    mv.n_fixed = 4 if fixed_bottom else 0
    keep = [nodes, hexes, part, iv, fixed, mat]
    mv.nodes = nodes.ctypes.data_as(_abi.f64p)
    mv.hexes = hexes.ctypes.data_as(_abi.i32p)
    mv.hex_part = part.ctypes.data_as(_abi.i32p)
    mv.materials = C.cast(mat, C.POINTER(_abi.Material))
    mv.initial_velocity = iv.ctypes.data_as(_abi.f64p)
    mv.fixed_nodes = fixed.ctypes.data_as(_abi.i32p)
    mv.settings = ca.default_settings()
    mv.settings.end_time = end_time
    mv.settings.frame_steps = frame_steps
    return mv, keep, nodes


def run_view(mv, n_nodes, n_hexes, frames=64):
    r, u, v, m, g = ca.alloc_results(n_nodes, 0, ca.PULSE_CAPACITY, 1, 0, n_hexes, frames)
    assert of.load().oracle_run(C.byref(mv), C.byref(r)) == 0
    return ca.to_results(r, u, v, m, g)


def test_honeycomb_frame_stress_is_the_yield_stress_at_its_compaction():
    """A honeycomb cube, bottom fixed, top moving down at 30 m/s: while it
    flows, the crush-direction stress is -(yield + hardening x compaction)
    (the bilinear curve at the frame's plastic strain), the lateral stresses
    vanish (the rate form has no Poisson coupling), and the rotation stays
    the identity."""
    mat = _abi.Material()
    mat.youngs_modulus, mat.poisson_ratio, mat.density, mat.model = 20e6, 0.0, 200.0, _abi.MAT_HONEYCOMB
    mat.yield_stress, mat.hardening, mat.transverse_factor = 0.2e6, 0.5e6, 1.0
    v0 = np.zeros((8, 3))
    v0[4:, 2] = -30.0
    mv, keep, _ = single_cube(mat, v0, True, 0.002, frame_steps=1)
    r = run_view(mv, 8, 1)
    c = r.frame_plastic_strain[:, 0].astype(float)
    s = r.frame_stress[:, 0].astype(float)
    flowing = c > 1e-3
    assert flowing.sum() >= 3
    want = -(0.2e6 + 0.5e6 * c[flowing])
    assert np.allclose(s[flowing, 2], want, rtol=2e-3)
    assert np.abs(s[flowing][:, [0, 1, 3, 4, 5]]).max() < 1e-3 * 0.2e6


def test_elastic_frame_stress_is_the_hencky_stress_over_j():
    """An elastic cube stretched by an initial velocity field: each frame's
    Cauchy stress equals K ln(J) / J I + 2G dev(ln V) / J with V the left
    stretch of the frame's deformation gradient (numpy's eigen-solver on
    F F^T, independent of the oracle's)."""
    E, nu = 2.1e9, 0.3
    mat = _abi.Material()
    mat.youngs_modulus, mat.poisson_ratio, mat.density, mat.model = E, nu, 7800.0, _abi.MAT_ELASTIC
    xi = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                   [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    X = 0.15 * (xi + 1.0)
    L = np.array([[20.0, 3.0, 0.0], [0.0, -5.0, 2.0], [1.0, 0.0, 8.0]])  # velocity gradient, 1/s
    mv, keep, nodes = single_cube(mat, X @ L.T, False, 2e-3, frame_steps=1)
    r = run_view(mv, 8, 1)
    K, G = E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu))
    grad = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                     [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float) / (4 * 0.3)
    checked = 0
    for k in range(1, len(r.frame_time)):
        u = r.frame_displacement[k].astype(float)
        F = np.eye(3) + u.T @ grad  # the centre gradient of a cube: dN_i/dX = xi_i / (4 a)
        J = np.linalg.det(F)
        lam, vec = np.linalg.eigh(F @ F.T)
        eps = vec @ np.diag(0.5 * np.log(lam)) @ vec.T
        vol = np.trace(eps)
        tau = K * vol * np.eye(3) + 2 * G * (eps - vol / 3 * np.eye(3))
        sig = tau / J
        want = np.array([sig[0, 0], sig[1, 1], sig[2, 2], sig[0, 1], sig[1, 2], sig[0, 2]])
        got = r.frame_stress[k, 0].astype(float)
        scale = np.abs(want).max()
        assert scale > 1e5
        # f32 displacements (frame) against the f64 state: 1e-3 of the scale
        assert np.abs(got - want).max() < 1e-3 * scale, (k, got, want)
        checked += 1
    assert checked >= 3
