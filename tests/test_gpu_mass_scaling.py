"""GPU parity of the step-size chain and selective mass scaling
(src/solver.rs:213-231,288-300, src/kernel.rs:207-258).

- The raw-array models of test_oracle_dt.py on the GPU: the initial step of
  a uniform block, the contact limit over every contact node, and the
  crushed cube's step and nodal-mass chain read back from the GPU's own
  per-step taps, against the same closed forms.
- crash 70 70 (Silverado vs Neon presets at +-70 mph), where selective mass
  scaling holds the step at the floor from step 40 on: per step for the
  first 100 steps the parity bars of test_gpu_parity, plus the nodal masses
  the two solvers divide by, read back from each one's trajectory; then the
  full solve's delta-v, crush, steps and largest added mass.
"""
import ctypes as C

import numpy as np
import pytest

import crushrs_amd as ca
import oracle_ffi as of
from test_gpu_parity import end_parity, step_parity
from test_oracle_dt import (Raw, c_d, check_crushed_cube, crushed_cube, elastic, honeycomb)

pytestmark = pytest.mark.gpu


def gpu_taps(mv, n_nodes, n_steps, k=0):
    """GPU per-step dt, f_int, f_ext and displacement (crush_debug_run_steps)."""
    n3 = 3 * n_nodes
    fi, fx, u = np.zeros((n_steps, n3)), np.zeros((n_steps, n3)), np.zeros((n_steps, n3))
    dt = np.zeros(n_steps)
    done = C.c_int32()
    ca.check(ca.load().crush_debug_run_steps(C.byref(mv), n_steps, of.ptr(fi), of.ptr(fx), of.ptr(u), None, None,
                                             of.ptr(dt), C.byref(done)))
    s = done.value
    return dt[:s], fi[:s], fx[:s], u[:s]


def test_initial_and_contact_steps_on_the_gpu():
    print("case: uniform blocks and a distant contact, first step", flush=True)
    r = Raw()
    r.block((3, 2, 2), (0.3, 0.2, 0.25), (0.0, 0.0, 0.0), elastic(2e9, 0.3, 1000.0))
    r.block((2, 2, 2), (0.3, 0.3, 0.3), (5.0, 0.0, 0.0), honeycomb(20e6, 200.0, 0.2e6, 0.5e6))
    r.settings.end_time = 1e-3
    dt, _, _, _ = gpu_taps(r.view(), r.n_nodes, 1)
    want = 0.85 * min(0.2 * (1.0 - 0.36 * 0.3) / c_d(2e9, 0.3, 1000.0), 0.3 / c_d(20e6, 0.0, 200.0))
    assert abs(dt[0] - want) < 1e-12 * want
    r = Raw()
    hc = honeycomb(20e6, 200.0, 0.2e6, 0.5e6)
    a = r.block((2, 2, 2), (0.3, 0.3, 0.3), (0.0, 0.0, 0.0), hc)
    b = r.block((2, 2, 2), (0.3, 0.3, 0.3), (1.6, 0.0, 0.0), hc)
    faces = [r.face(b[0, j, k], b[0, j + 1, k], b[0, j + 1, k + 1], b[0, j, k + 1]) for j in range(2) for k in range(2)]
    r.contact(a[-1].ravel(), faces, 2e9, 0.3, scale=np.full(9, 5.0))
    r.settings.end_time = 1e-3
    dt, _, _, _ = gpu_taps(r.view(), r.n_nodes, 1)
    want = 0.85 * 2.0 * np.sqrt(200.0 * 0.3 ** 3 / 8.0 / 2e9)
    assert abs(dt[0] - want) < 1e-12 * want


def test_crushed_cube_step_and_mass_chain_on_the_gpu():
    print("case: crushed honeycomb cube through mass scaling to inversion", flush=True)
    r, top = crushed_cube()
    mv = r.view()
    dt, fi, _, u = gpu_taps(mv, r.n_nodes, 400)
    assert len(dt) > 50
    # the GPU's corner forces are f32: the top nodes part by a few nm, so
    # the element is a box to 1e-7 m and its limit h / c_d to 1e-5
    extras, scaled, capped = check_crushed_cube(dt, fi, u, top, r.settings.end_time, rtol_mass=1e-5, rtol_dt=1e-5,
                                                flat_tol=1e-7)
    assert scaled >= 3 and capped >= 1
    res, uu, vv, mm, gg = ca.alloc_results(r.n_nodes, 0)
    ca.check(ca.load().crush_run(C.byref(mv), C.byref(res)))
    assert abs(res.added_mass_max - max(extras)) < 1e-6 * max(extras)
    assert res.steps == len(dt)


def effective_masses(dt, fi, fx, u):
    """Nodal mass per DOF and step that the two half steps divided by, from
    a trajectory (steps whose next step is no adaptive check, so both half
    steps use one mass): the change of the half-step velocity equals
    dt_{s+1}/2 (f_ext_{s+1} - f_int_s)/m + dt_s/2 (f_ext_s - f_int_s)/m.
    Returns (steps, masses, weights) with weights |change| / max |change|."""
    out_s, out_m, out_w = [], [], []
    for s in range(1, len(dt) - 1):
        if (s + 1) % 5 == 0:
            continue
        dvh = (u[s + 1] - u[s]) / dt[s + 1] - (u[s] - u[s - 1]) / dt[s]
        den = 0.5 * dt[s + 1] * (fx[s + 1] - fi[s]) + 0.5 * dt[s] * (fx[s] - fi[s])
        with np.errstate(divide="ignore", invalid="ignore"):
            m = den / dvh
        out_s.append(s)
        out_m.append(m)
        out_w.append(np.abs(dvh) / max(np.abs(dvh).max(), 1e-300))
    return out_s, np.array(out_m), np.array(out_w)


def test_crash_70_70_mass_scaling_matches_oracle():
    print("case: crash 70 70 (mass scaling from step 40), 100 steps and full solve", flush=True)
    m = ca.Model.crash(70.0, 70.0)
    step_parity(m, 100)
    k = m.contact_node_count()
    gdt, gfi, gfx, gu = gpu_taps(m.view, m.n_nodes, 100, k)
    oi, ox, _, _, odt = of.run_steps(m, 100, k)
    ou = np.zeros((100, 3 * m.n_nodes))
    done = C.c_int32()
    assert of.load().oracle_run_displacements(m.ptr, 100, 0, of.ptr(ou), C.byref(done)) == 0
    gs, gm, gw = effective_masses(gdt, gfi, gfx, gu)
    os_, om, _ = effective_masses(odt, oi, ox, ou[:done.value])
    assert gs == os_
    lumped = np.repeat(of.run(ca.Model.crash(70.0, 70.0).set(end_time=1e-9)).masses, 3)
    sel = gw > 1e-2  # DOFs whose velocity changes enough to read the mass
    rel = np.abs(gm[sel] - om[sel]) / om[sel]
    assert np.all(rel < 1e-4), rel.max()
    # scaling is exercised: from step 40 on some nodes carry added mass
    late = np.array(gs) >= 45
    ratio = gm[late] / lumped
    assert np.nanmax(np.where(sel[late], ratio, np.nan)) > 1.05
    # before the first scaling every mass is the lumped one
    early = np.array(gs) < 35
    assert np.all(np.abs(np.where(sel[early], gm[early] / lumped, 1.0) - 1.0) < 1e-6)
    g, o = end_parity(m)
    assert o.added_mass_max > 1.0 and abs(g.added_mass_max - o.added_mass_max) < 1e-2 * o.added_mass_max
