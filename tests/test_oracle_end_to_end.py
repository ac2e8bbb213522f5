"""The f64 oracle against the end-to-end checks of the reference's own
tests (tests/impact.rs), which hold the only reference numbers for whole
solves of this path.

tests/impact.rs:284-309 reads delta-v from the integrated rear
accelerometer pulse; here the accelerometer node's own final velocity is
used, which is that integral before the output filter (the history and
filter layer is not part of this path).
"""
import numpy as np

import crushrs_amd as ca
import oracle_ffi as of


def test_head_on_conserves_momentum_with_plausible_delta_v():
    """tests/impact.rs:33-47 on `crash 30 30` (src/bin/crushrs.rs:403-432)."""
    m = ca.Model.crash(30.0, 30.0)
    r = of.run(m)
    (mt, dvt), (mc, dvc) = [(g.mass, g.delta_v[0]) for g in r.groups]
    v0 = 30.0 * ca.MPH
    assert abs(mt * dvt + mc * dvc) < 1e-6 * mt * v0
    plastic = (mt * v0 - mc * v0) / (mt + mc) + v0
    assert plastic < dvc < 1.3 * plastic
    e = ((dvc - v0) - (v0 + dvt)) / (2.0 * v0)
    assert 0.0 < e < 0.3
    assert r.eroded_elements == 0
    # the group delta-v is the mass-weighted mean of the node velocities
    nodes = m.node_set("silverado")
    _, vbar = r.mean_velocity(nodes)
    assert abs((vbar[0] - v0) - dvt) < 1e-9


def test_navigator_neon_head_on_matches_nhtsa_4429():
    """tests/impact.rs:284-309: Navigator vs pulse-tuned Neon (1378 kg) at
    30.14 mph each on the road; NHTSA test 4429 delta-v 9.89 +- 1.0 and
    19.57 +- 2.0 m/s, the Neon's crush zone not diving, and the vehicles
    separating by 150 ms."""
    a = ca.renamed(ca.vehicle("navigator"), "navigator_a")
    b = ca.vehicle("neon-pulse")
    b.mass = 1378.0
    b = ca.renamed(b, "neon_b")
    v0 = 30.14 * ca.MPH
    m = ca.Model.headon(a, b, v0, v0)
    r = of.run(m)
    v = m.view
    ga, gb = v.groups[0].accel_node, v.groups[1].accel_node
    dv_a = v0 - r.velocity[ga][0]
    dv_b = v0 + r.velocity[gb][0]
    assert abs(dv_a - 9.89) < 1.0, dv_a
    assert abs(dv_b - 19.57) < 2.0, dv_b
    names = m.part_names()
    crush = [i for i, p in enumerate(names) if p.startswith("neon_b") and "crush" in p]
    assert crush, names
    parts = np.ctypeslib.as_array(v.hex_part, shape=(v.n_hexes,))
    hexes = m.hexes()
    nodes = np.unique(hexes[np.isin(parts, crush)])
    _, vc = r.mean_velocity(nodes)
    assert -3.0 < vc[2] < 7.0, vc
    _, va = r.mean_velocity(m.node_set("navigator_a"))
    _, vb = r.mean_velocity(m.node_set("neon_b"))
    assert vb[0] > va[0]
