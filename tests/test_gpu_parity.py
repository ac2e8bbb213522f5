"""GPU path vs the f64 oracle on the same models (parity tests proper)."""
import numpy as np
import pytest

import crushrs_amd as ca
import oracle_ffi as of

pytestmark = pytest.mark.gpu


def _step_parity(model, n_steps, force_tol=1e-4):
    k = model.contact_node_count()
    gi, gx, gh, ge, gdt = ca.debug_run_steps(model, n_steps)
    oi, ox, oh, oe, odt = of.run_steps(model, n_steps, k)
    assert len(gdt) == len(odt) == n_steps
    # erosion and contact-pair decisions identical step by step
    assert np.array_equal(ge, oe)
    mism = np.argwhere(gh != oh)
    assert mism.size == 0, f"contact decisions differ at (step, node) {mism[:5].tolist()}"
    # nodal internal-force L2 relative error per step
    for s in range(n_steps):
        on = np.linalg.norm(oi[s])
        if on == 0.0:
            assert np.linalg.norm(gi[s]) < 1e-6
            continue
        rel = np.linalg.norm(gi[s] - oi[s]) / on
        assert rel < force_tol, f"step {s}: f_int rel err {rel:.2e}"
    np.testing.assert_allclose(gdt, odt, rtol=1e-5)
    return gi, oi


def test_crash_30_30_first_100_steps():
    print("case: crash 30 30, 100 steps", flush=True)
    m = ca.Model.crash(30.0, 30.0)
    _step_parity(m, 100)


def test_crash_30_30_delta_v_and_crush():
    print("case: crash 30 30 full solve", flush=True)
    m = ca.Model.crash(30.0, 30.0)
    g = ca.run(m)
    o = of.run(m)
    assert g.eroded_elements == o.eroded_elements
    for a, b in zip(g.groups, o.groups):
        assert abs(a.delta_v[0] - b.delta_v[0]) < 0.05, (a.delta_v, b.delta_v)
        assert abs(a.crush_end - b.crush_end) < 2e-3
        assert abs(a.crush_max - b.crush_max) < 2e-3
