"""Scenario builders (host side) against the counts and parameters the
reference holds: the `crash` command's mesh (README.md:33-35), the material
values of examples/crash_impact.toml, and the per-vehicle element counts of
the tuned vehicle TOMLs."""
import numpy as np
import pytest

import crushrs_amd as ca


def part_counts(model, prefix):
    v = model.view
    hp = np.ctypeslib.as_array(v.hex_part, shape=(v.n_hexes,))
    hc = el = 0
    for p, nm in enumerate(model.part_names()):
        if nm != prefix and not nm.startswith(prefix + "_"):
            continue
        cnt = int((hp == p).sum())
        if v.materials[p].model == 1:
            hc += cnt
        else:
            el += cnt
    return hc, el


def test_crash_mesh_counts():
    m = ca.Model.crash(30.0, 30.0)
    assert m.n_hexes == 1290            # README.md:34
    assert m.n_nodes == 1848
    v = m.view
    assert v.n_contacts == 2
    # pair contact at half of min(E A / (n h)) per side (src/bin/crushrs.rs:420-421)
    k_car = 23.1e6 * 1.71 * 1.35 / (42 * 0.3)
    k_truck = 27.8e6 * 2.03 * 1.87 / (42 * 0.3)
    for c in range(2):
        assert abs(v.contacts[c].stiffness - 0.5 * min(k_car, k_truck)) < 1e-6 * k_car
        assert v.contacts[c].max_distance == 0.3
    assert (v.contacts[0].n_nodes, v.contacts[0].n_faces) == (56, 30)
    assert (v.contacts[1].n_nodes, v.contacts[1].n_faces) == (42, 42)
    assert len(m.node_set("silverado")) + len(m.node_set("neon")) == 1848


def test_crash_materials_match_example_toml():
    # examples/crash_impact.toml:24-43 lists the same presets as plain numbers
    m = ca.Model.crash(30.0, 30.0)
    v = m.view
    names = m.part_names()
    s = v.materials[names.index("silverado")]
    n = v.materials[names.index("neon")]
    assert abs(s.youngs_modulus - 27.8e6) < 1 and abs(n.youngs_modulus - 23.1e6) < 1
    # the example file rounds these to about 5 significant digits
    assert abs(s.density - 118.06) < 0.02 and abs(n.density - 134.53) < 0.02
    assert abs(s.yield_stress - 13490.0) < 1.0 and abs(n.yield_stress - 7248.0) < 1.0
    assert abs(s.hardening / 861711.0 - 1) < 1e-4 and abs(n.hardening / 553354.0 - 1) < 1e-4
    iv = m.initial_velocity()
    assert np.allclose(iv[m.node_set("silverado"), 0], 30 * ca.MPH)
    assert np.allclose(iv[m.node_set("neon"), 0], -30 * ca.MPH)


@pytest.mark.parametrize("h,name,want", [
    (0.3, "neon_1996", (420, 330)), (0.3, "silverado_2007", (840, 0)),
    (0.3, "expedition_1999", (672, 546)), (0.3, "navigator_1999", (672, 546)),
    (0.2, "neon_1996", (882, 945)), (0.2, "silverado_2007", (2610, 0)),
    (0.2, "expedition_1999", (1440, 1710)), (0.2, "navigator_1999", (1440, 1620)),
])
def test_tuned_toml_element_counts(h, name, want):
    a = ca.renamed(ca.vehicle_file(name), "a")
    b = ca.renamed(ca.vehicle("silverado"), "b")
    m = ca.Model.headon(a, b, 10.0, 10.0, element_size=h)
    assert part_counts(m, "a") == want
    # road: 32 x 8 x 1 hexes (src/vehicle.rs:829-833 with h = w / 8)
    assert "ground" in m.part_names()


def test_toml_matches_preset():
    # vehicles/silverado_2007.toml is the tuned preset written as TOML
    t = ca.vehicle_file("silverado_2007")
    p = ca.vehicle("silverado")
    for f in ("mass", "crush_zone_length", "modulus", "body_modulus", "transverse_factor"):
        assert getattr(t, f) == pytest.approx(getattr(p, f))
    assert list(t.curve) == pytest.approx(list(p.curve))


def test_side_and_rear_profiles():
    neon = ca.vehicle("neon")
    side = ca.side_profile(neon)
    assert side.name == b"neon_side"
    assert list(side.size) == pytest.approx([1.71, 4.36, 1.35])
    assert side.modulus == pytest.approx(125.8e6)
    with pytest.raises(ca.CrushError) as e:
        ca.rear_profile(neon)   # no rear curve: the reference refuses (src/vehicle.rs:297-298)
    assert e.value.code == 7
    neon.has_rear_curve = 1
    neon.rear_curve[0], neon.rear_curve[1] = 5e4, 1e6
    rear = ca.rear_profile(neon)
    assert rear.name == b"neon_rear" and rear.crush_zone_length == pytest.approx(4.36)


def test_unknown_vehicle_and_bad_toml(tmp_path):
    with pytest.raises(ca.CrushError) as e:
        ca.vehicle("trabant")
    assert e.value.code == 7
    p = tmp_path / "bad.toml"
    p.write_text('name = "x"\nmass = 1000.0\n')
    with pytest.raises(ca.CrushError) as e:
        ca.vehicle(str(p))
    assert e.value.code == 1


def test_tbone_and_barrier_build():
    m = ca.Model.tbone(ca.vehicle("silverado"), ca.vehicle("neon"), 30 * ca.MPH, 0.0)
    assert "neon_side" in m.part_names() and m.view.n_groups == 2
    b = ca.Model.barrier(ca.vehicle_file("neon_1996"))
    names = b.part_names()
    assert {"wall", "ground", "neon", "neon_crush", "neon_rails", "neon_bumper"} <= set(names)
