"""Serving (crush_server_*): continuous batching with admission by an HBM
budget.  Every result must be bitwise the one-model GPU solve of the same
model, whether the request ran alone, in a mix that changed while it was
resident, or waited for memory; retirement at separation must match the
one-model path with the same settings."""
import threading
import time

import numpy as np
import pytest

import crushrs_amd as ca

pytestmark = pytest.mark.gpu


def toml_pair(a, b, va, vb, offset=0.0, h=0.3, end_time=0.15):
    return ca.Model.headon(ca.renamed(ca.vehicle_file(a), f"{a}_a"), ca.renamed(ca.vehicle_file(b), f"{b}_b"),
                           va * ca.MPH, vb * ca.MPH, offset=offset, element_size=h, end_time=end_time)


def mix(n):
    out = []
    names = ["neon_1996", "silverado_2007", "expedition_1999", "navigator_1999"]
    for i in range(n):
        k = i % 4
        if k == 0:
            out.append(ca.Model.crash(20.0 + i, 25.0 + (i % 7)))
        elif k == 1:
            out.append(toml_pair(names[i % 4], names[(i + 1) % 4], 20 + i % 10, 15 + i % 9, offset=0.05 * (i % 5),
                                 end_time=0.06))
        elif k == 2:
            out.append(ca.Model.tbone(ca.vehicle("silverado"), ca.side_profile(ca.vehicle("neon")), 12.0 + i % 5, 0.0))
        else:
            out.append(ca.Model.barrier(ca.vehicle_file(names[i % 4]), end_time=0.04))
    return out


def same(a, b):
    assert a.steps == b.steps
    assert a.eroded_elements == b.eroded_elements
    assert np.array_equal(a.displacement, b.displacement)
    assert np.array_equal(a.velocity, b.velocity)
    for ga, gb in zip(a.groups, b.groups):
        assert np.all(np.abs(ga.delta_v - gb.delta_v) <= 1e-9)
        assert abs(ga.crush_max - gb.crush_max) <= 1e-9


def test_server_results_equal_plain():
    print("case: server mix of 12 vs plain", flush=True)
    models = mix(12)
    plain = [ca.run(m) for m in models]
    with ca.Server(steps_per_graph=8) as s:
        tickets = [s.submit(m, fields=True) for m in models]
        res = [s.wait(t, timeout_ms=120000) for t in tickets]
        st = s.stats()
    assert st.completed == 12 and st.failed == 0
    for r, p in zip(res, plain):
        same(r, p)


def test_server_alone_vs_changing_mix():
    print("case: server alone vs changing mix", flush=True)
    x = toml_pair("neon_1996", "expedition_1999", 25, 20, offset=0.3, end_time=0.08)
    with ca.Server() as s:
        alone = s.wait(s.submit(x, fields=True), timeout_ms=120000)
    others = mix(16)
    with ca.Server(steps_per_graph=4) as s:
        early = [s.submit(m) for m in others[:8]]
        t = s.submit(x, fields=True)
        late = []

        def feed():
            for m in others[8:]:
                time.sleep(0.01)
                late.append(s.submit(m))
        th = threading.Thread(target=feed)
        th.start()
        inmix = s.wait(t, timeout_ms=120000)
        th.join()
        for k in early + late:
            s.wait(k, timeout_ms=120000)
    same(inmix, alone)


def test_server_admission_by_hbm_budget():
    print("case: server with a budget for two requests", flush=True)
    models = mix(6)
    plain = [ca.run(m) for m in models]
    with ca.Server(hbm_budget_bytes=6e6, max_resident=2, steps_per_graph=8) as s:
        tickets = [s.submit(m, fields=True) for m in models]
        st = s.stats()
        assert st.queued + st.resident + st.completed == 6
        res = [s.wait(t, timeout_ms=120000) for t in tickets]
        st = s.stats()
    assert st.completed == 6 and st.resident == 0
    for r, p in zip(res, plain):
        same(r, p)


def test_server_retires_at_separation():
    print("case: server separation retirement", flush=True)
    m = ca.Model.crash(30.0, 30.0).set(separation_steps=20)
    p = ca.run(ca.Model.crash(30.0, 30.0).set(separation_steps=20))
    with ca.Server() as s:
        r = s.wait(s.submit(m, fields=True), timeout_ms=120000)
    same(r, p)
    assert r.separated == p.separated


def test_server_rejects_what_never_fits_and_unknown_tickets():
    print("case: server errors", flush=True)
    with ca.Server(hbm_budget_bytes=1e5) as s:
        with pytest.raises(ca.CrushError) as e:
            s.submit(ca.Model.crash(30.0, 30.0))
        assert e.value.code == 6  # CRUSH_E_BUSY
        with pytest.raises(ca.CrushError):
            s.wait(12345)
