"""crush_run_noalloc (no device memory allocated) against crush_run.

Two calls in a row share the GPU's static scratch, the second with a
smaller model, and each result must equal crush_run's bitwise: steps,
displacements, velocities, group delta-v and crush, pulses.  A model too
large for the scratch and a request for frames are refused before any
kernel runs."""
import numpy as np
import pytest

import crushrs_amd as ca
from crushrs_amd import _abi

pytestmark = pytest.mark.gpu


def same(a, b):
    assert a.steps == b.steps and a.eroded_elements == b.eroded_elements
    assert np.array_equal(a.displacement, b.displacement)
    assert np.array_equal(a.velocity, b.velocity)
    assert np.array_equal(a.masses, b.masses)
    assert a.added_mass_max == b.added_mass_max
    for x, y in zip(a.groups, b.groups):
        assert np.array_equal(x.delta_v, y.delta_v)
        assert x.crush_max == y.crush_max and x.crush_end == y.crush_end
    if a.pulse is not None:
        assert np.array_equal(a.pulse, b.pulse)


def test_noalloc_equals_run_on_two_calls_the_second_smaller():
    big = ca.Model.headon(ca.renamed(ca.vehicle_file("expedition_1999"), "expedition_a"),
                          ca.renamed(ca.vehicle_file("silverado_2007"), "silverado_b"), 30 * ca.MPH, 25 * ca.MPH,
                          offset=0.2, element_size=0.2, end_time=0.05).set(pulse_dt=1e-4)
    small = ca.Model.crash(30.0, 30.0).set(end_time=0.03)
    assert small.n_hexes < big.n_hexes
    print(f"case: no-alloc head-on {big.n_hexes} hexes, then crash {small.n_hexes} hexes", flush=True)
    a1 = ca.run_noalloc(big)
    a2 = ca.run_noalloc(small)
    same(a1, ca.run(big))
    same(a2, ca.run(small))
    # and once more after crush_run, whose buffers came and went meanwhile
    same(ca.run_noalloc(small), a2)


def test_noalloc_refuses_what_does_not_fit():
    print("case: no-alloc refusals (frames, oversized model)", flush=True)
    m = ca.Model.crash(30.0, 30.0).set(end_time=0.002, frame_steps=5)
    r, u, v, ms, g = ca.alloc_model_results(m)
    rc = ca.load().crush_run_noalloc(m.ptr, ca.C.byref(r))
    assert rc == _abi.CRUSH_E_UNSUPPORTED, rc
    # about 290k hexes: the prepared model alone exceeds the scratch
    v = ca.vehicle("silverado")
    huge = ca.Model.barrier(v, element_size=0.045, speed=10 * ca.MPH).set(end_time=1e-5)
    assert huge.n_hexes > 150_000
    with pytest.raises(ca.CrushError) as e:
        ca.run_noalloc(huge)
    assert e.value.code == _abi.CRUSH_E_INPUT and "scratch" in str(e.value)
    # the scratch still works after a refusal
    small = ca.Model.crash(30.0, 30.0).set(end_time=0.005)
    same(ca.run_noalloc(small), ca.run(small))
