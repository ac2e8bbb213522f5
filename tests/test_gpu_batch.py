"""Batched solves: every instance of a batch must reproduce the one-model
GPU path (crush_run) on the same model, whatever the batch size, mix,
workgroup size or steps per launch, and the one-model path is held to the
oracle by test_gpu_parity.py.

Per-instance arithmetic is the same code in both paths, so displacements,
velocities, step counts and erosion must be bitwise equal; the reported
delta-v and crush come from block reductions whose summation order depends
on the workgroup size, so they are compared to 1e-9.
"""
import numpy as np
import pytest

import crushrs_amd as ca
import oracle_ffi as of

pytestmark = pytest.mark.gpu


def toml_pair(a, b, va, vb, offset=0.0, h=0.3, end_time=0.15):
    va_ = ca.renamed(ca.vehicle_file(a), f"{a}_a")
    vb_ = ca.renamed(ca.vehicle_file(b), f"{b}_b")
    return ca.Model.headon(va_, vb_, va * ca.MPH, vb * ca.MPH, offset=offset, element_size=h, end_time=end_time)


def mix():
    return [
        ca.Model.crash(30.0, 30.0),
        toml_pair("neon_1996", "expedition_1999", 25, 20, offset=0.3),
        ca.Model.crash(20.0, 35.0),
        toml_pair("silverado_2007", "navigator_1999", 30, 15, offset=0.1, end_time=0.08),
        ca.Model.barrier(ca.vehicle_file("silverado_2007"), end_time=0.05),
    ]


def same_as_plain(br, pr, tol=1e-9):
    assert br.steps == pr.steps
    assert br.eroded_elements == pr.eroded_elements
    assert br.time == pr.time
    assert np.array_equal(br.displacement, pr.displacement)
    assert np.array_equal(br.velocity, pr.velocity)
    for a, b in zip(br.groups, pr.groups):
        assert np.all(np.abs(a.delta_v - b.delta_v) <= tol), (a.delta_v, b.delta_v)
        assert abs(a.crush_end - b.crush_end) <= tol
        assert abs(a.crush_max - b.crush_max) <= tol


@pytest.mark.parametrize("tpi,tile,spg", [(256, 256, 32), (64, 64, 7), (512, 128, 1), (128, 256, 100)])
def test_batch_matches_plain(tpi, tile, spg):
    print(f"case: batch mix threads_per_instance={tpi} force_tile={tile} steps_per_graph={spg}", flush=True)
    models = mix()
    plain = [ca.run(m) for m in models]
    b = ca.Batch(models, threads_per_instance=tpi, force_tile=tile, steps_per_graph=spg)
    st = b.run()
    assert st.instances == len(models)
    for i, p in enumerate(plain):
        same_as_plain(b.result(i), p)


def test_batch_alone_vs_mix_and_reset():
    print("case: batch alone vs mix, reset", flush=True)
    models = mix()
    alone = ca.Batch([models[1]], steps_per_graph=16)
    alone.run()
    ra = alone.result(0)
    crowd = ca.Batch(models + [ca.Model.crash(25.0, 25.0) for _ in range(11)], steps_per_graph=16)
    crowd.run()
    rm = crowd.result(1)
    assert np.array_equal(ra.displacement, rm.displacement)
    assert np.array_equal(ra.velocity, rm.velocity)
    assert ra.steps == rm.steps
    for a, b in zip(ra.groups, rm.groups):
        assert np.array_equal(a.delta_v, b.delta_v)
        assert a.crush_max == b.crush_max
    # a reset batch runs again to the same results
    crowd.reset()
    crowd.run()
    rr = crowd.result(1)
    assert np.array_equal(rr.displacement, rm.displacement)


def test_batch_autotune_keeps_results():
    print("case: batch autotune", flush=True)
    models = mix()[:3]
    ref = ca.Batch(models)
    ref.run()
    tuned = ca.Batch(models, autotune=1)
    o = tuned.options
    assert o.autotune == 0 and o.threads_per_instance in (64, 128, 256, 512) and o.force_tile in (64, 128, 256)
    tuned.run()
    for i in range(len(models)):
        a, b = ref.result(i), tuned.result(i)
        assert a.steps == b.steps
        assert np.array_equal(a.displacement, b.displacement)


def test_batch_size_class_autotune_keeps_results():
    print("case: batch autotune with size classes (8 instances)", flush=True)
    models = mix()[:3] + [ca.Model.crash(20.0 + 2.0 * k, 25.0) for k in range(5)]
    ref = ca.Batch(models)
    ref.run()
    tuned = ca.Batch(models, autotune=1)
    classes = tuned.class_scales()
    assert len(classes) == 4 and all(cs in (0.5, 1.0, 2.0) for _, cs in classes)
    assert [e for e, _ in classes] == sorted(e for e, _ in classes)
    assert ref.class_scales() == []
    tuned.run()
    for i in range(len(models)):
        a, b = ref.result(i), tuned.result(i)
        assert a.steps == b.steps
        assert np.array_equal(a.displacement, b.displacement)
        assert np.array_equal(a.velocity, b.velocity)


def test_batch_vs_oracle_delta_v():
    print("case: batch crash 30 30 vs oracle", flush=True)
    m = ca.Model.crash(30.0, 30.0)
    b = ca.Batch([m, ca.Model.crash(30.0, 30.0)])
    b.run()
    o = of.run(m)
    for i in range(2):
        g = b.result(i)
        assert abs(g.steps - o.steps) <= max(1, 0.01 * o.steps)
        for a, r in zip(g.groups, o.groups):
            assert np.all(np.abs(a.delta_v - r.delta_v) < 0.05)
            assert abs(a.crush_max - r.crush_max) < 2e-3


def test_batch_separation_retires_early():
    print("case: batch separation retirement", flush=True)
    full = ca.Model.crash(30.0, 30.0)
    early = ca.Model.crash(30.0, 30.0).set(separation_steps=20)
    b = ca.Batch([full, early])
    b.run()
    rf, re = b.result(0), b.result(1)
    assert not rf.separated
    if re.separated:
        assert re.steps < rf.steps
    pe = ca.run(ca.Model.crash(30.0, 30.0).set(separation_steps=20))
    same_as_plain(re, pe)
